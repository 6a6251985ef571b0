"""CPU generation of the seeded star-schema tables (TEST INFRASTRUCTURE).

Drives the oracle's generator functions (oracle/sdb_oracle.c) with a column
specification passed in by the caller (the product's schema module), so the
oracle never imports the product.  Produces {table: {column: int32 array}}.
"""
import numpy as np

from . import oracle as O


def generate_table(table, nrows, seed, spec, order, stream_id, dates, first_row=0,
                   columns=None):
    """spec(col) -> generator tuple; order: column generation order;
    dates: (datekey, year, yearmonthnum) arrays of the date dimension."""
    need = None
    if columns is not None:
        # the requested columns and the columns they are derived from
        need = set(columns)
        for col in reversed(order):
            if col in need and spec(col)[0] == "child":
                need.add(spec(col)[1])
    out = {}
    for col in order:
        if need is not None and col not in need:
            continue
        kind = spec(col)
        s = stream_id(col)
        if kind[0] == "rowid":
            out[col] = np.arange(first_row + 1, first_row + nrows + 1, dtype=np.int32)
        elif kind[0] == "uniform":
            out[col] = O.gen_uniform(nrows, seed, s, kind[1], kind[2], first_row)
        elif kind[0] == "child":
            out[col] = O.gen_child(out[kind[1]], seed, s, kind[2], kind[3], kind[4], first_row)
        elif kind[0] == "lookup":
            assert kind[1] == "ddate"
            out[col] = O.gen_lookup(nrows, seed, s, dates[{0: 0, 4: 1, 5: 2}[kind[2]]], first_row)
        elif kind[0] == "date":
            arr = dates[{0: 0, 4: 1, 5: 2}[kind[1]]]
            out[col] = np.ascontiguousarray(arr[first_row:first_row + nrows])
        else:
            raise ValueError(kind)
    if columns is not None:
        out = {c: v for c, v in out.items() if c in columns}
    return out


def generate_tables(schema, fact_rows, seed, sizes=None, fact_columns=None, first_row=0):
    """All five tables.  ``schema`` is the product's schema module (passed in)."""
    sizes = dict(schema.DIM_SIZES, **(sizes or {}))
    dates = O.gen_dates(sizes["ddate"])
    tables = {}
    for t in ("ddate", "customer", "supplier", "part", "lineorder"):
        n = fact_rows if t == "lineorder" else sizes[t]
        tables[t] = generate_table(
            t, n, seed,
            spec=lambda c, t=t: schema.column_spec(t, c, sizes),
            order=schema.generation_order(t),
            stream_id=lambda c, t=t: schema.stream_id(t, c),
            dates=dates,
            first_row=first_row if t == "lineorder" else 0,
            columns=fact_columns if t == "lineorder" else None)
    return tables
