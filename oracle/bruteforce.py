"""Independent brute-force evaluator of the benchmark templates.

TEST INFRASTRUCTURE ONLY.  Written straight from the SQL of each template
(server/queries/original/q11.sql, queries/transformed/q21.sql, q31.sql,
q41.sql) in numpy, sharing no code with the operator restatement
(oracle/sdb_oracle.c) or the plan restatement (oracle/plan_oracle.py).  It pins
the oracle: both must give the same rows.

Result conventions of the reference that SQL does not fix, applied here as
the final step only:
  * SUM is reported as the uint64 holding the two's-complement int64 sum
    (kernels/aggregation.hpp:341, executor/result_writer.cpp:59-60);
  * group-by rows come in dense-index order with the first GROUP BY column
    varying fastest (kernels/aggregation.hpp:315-321), i.e. sorted by the
    reversed key tuple when every key lies in its dimension's range.
"""
import numpy as np

U64 = 2**64


def _u64(total):
    return int(total) % U64


def _date_mask(dd, p):
    year, ymn = dd[4], dd[5]
    if p[0] == "year":
        return year == p[1]
    if p[0] in ("year_range", "year_range_and"):
        return (year >= p[1]) & (year <= p[2])
    if p[0] == "yearmonthnum":
        return ymn == p[1]
    raise ValueError(p)


def _ent_mask(t, p):
    if p[0] == "region":
        return t[5] == p[1]
    if p[0] == "nation":
        return t[4] == p[1]
    if p[0] == "city":
        return t[3] == p[1]
    if p[0] == "city2":
        return (t[3] == p[1]) | (t[3] == p[2])
    raise ValueError(p)


def _part_mask(pt, p):
    if p[0] == "category":
        return pt[3] == p[1]
    if p[0] == "brand":
        return pt[4] == p[1]
    if p[0] == "brand_range":
        return (pt[4] >= p[1]) & (pt[4] <= p[2])
    if p[0] == "mfgr2":
        return (pt[2] == p[1]) | (pt[2] == p[2])
    raise ValueError(p)


def _lookup(keys, dim_keys, dim_mask, dim_vals=None):
    """Equi-join of fact keys with the selected dimension rows by sorting."""
    sel_keys = dim_keys[dim_mask]
    order = np.argsort(sel_keys, kind="stable")
    sk = sel_keys[order]
    pos = np.searchsorted(sk, keys)
    pos_c = np.minimum(pos, max(len(sk) - 1, 0))
    hit = (len(sk) > 0) & (sk[pos_c] == keys) if len(sk) else np.zeros(len(keys), bool)
    if dim_vals is None:
        return hit
    vals = dim_vals[dim_mask][order]
    return hit, vals[pos_c]


def _group_rows(keys, values):
    """SQL GROUP BY with SUM, keys: list of int arrays; rows ordered by the
    reversed key tuple (dense order of the reference)."""
    if len(values) == 0:
        return []
    kt = np.stack([k.astype(np.int64) for k in keys], axis=1)
    uniq, inv = np.unique(kt, axis=0, return_inverse=True)
    sums = np.zeros(len(uniq), dtype=object)
    vals = values.astype(np.int64)
    order = np.argsort(inv.ravel(), kind="stable")
    inv_s = inv.ravel()[order]
    v_s = vals[order]
    bounds = np.flatnonzero(np.diff(inv_s)) + 1
    starts = np.concatenate([[0], bounds])
    ends = np.concatenate([bounds, [len(inv_s)]])
    for g, (a, b) in enumerate(zip(starts, ends)):
        sums[inv_s[a]] = int(v_s[a:b].sum(dtype=np.int64))
    rows = [tuple(int(x) for x in uniq[i]) + (_u64(sums[i]),) for i in range(len(uniq))]
    rows.sort(key=lambda r: tuple(reversed(r[:-1])))
    return rows


def t1(tables, date=("year", 1993), disc=(1, 3), qty=("lt", 25)):
    lo, dd = tables["lineorder"], tables["ddate"]
    m = _lookup(lo[5], dd[0], _date_mask(dd, date))
    m &= (lo[11] >= disc[0]) & (lo[11] <= disc[1])
    if qty[0] == "lt":
        m &= lo[8] < qty[1]
    else:
        m &= (lo[8] >= qty[1]) & (lo[8] <= qty[2])
    total = (lo[9][m].astype(np.int64) * lo[11][m].astype(np.int64)).sum(dtype=np.int64)
    return [(_u64(total),)]


def t2(tables, part=("category", 1), s_region=1):
    lo, pt, su, dd = tables["lineorder"], tables["part"], tables["supplier"], tables["ddate"]
    hp, brand = _lookup(lo[3], pt[0], _part_mask(pt, part), pt[4])
    hs = _lookup(lo[4], su[0], su[5] == s_region)
    hd, year = _lookup(lo[5], dd[0], np.ones(len(dd[0]), bool), dd[4])
    m = hp & hs & hd
    return _group_rows([year[m], brand[m]], lo[12][m])


def t3(tables, cust=("region", 2), supp=("region", 2), date=("year_range", 1992, 1997),
       group="nation"):
    lo, cu, su, dd = tables["lineorder"], tables["customer"], tables["supplier"], tables["ddate"]
    g = 4 if group == "nation" else 3
    hc, cg = _lookup(lo[2], cu[0], _ent_mask(cu, cust), cu[g])
    hs, sg = _lookup(lo[4], su[0], _ent_mask(su, supp), su[g])
    hd, year = _lookup(lo[5], dd[0], _date_mask(dd, date), dd[4])
    m = hc & hs & hd
    return _group_rows([cg[m], sg[m], year[m]], lo[12][m])


def t4(tables, c_region=1, s_region=1, mfgr=(0, 1)):
    lo, cu, su, pt, dd = (tables["lineorder"], tables["customer"], tables["supplier"],
                          tables["part"], tables["ddate"])
    hc, nation = _lookup(lo[2], cu[0], cu[5] == c_region, cu[4])
    hs = _lookup(lo[4], su[0], su[5] == s_region)
    hp = _lookup(lo[3], pt[0], (pt[2] == mfgr[0]) | (pt[2] == mfgr[1]))
    hd, year = _lookup(lo[5], dd[0], np.ones(len(dd[0]), bool), dd[4])
    m = hc & hs & hp & hd
    profit = lo[12][m].astype(np.int64) - lo[13][m].astype(np.int64)
    return _group_rows([year[m], nation[m]], profit)


TEMPLATES = {"t1": t1, "t2": t2, "t3": t3, "t4": t4}


def evaluate(template, tables, **params):
    return TEMPLATES[template](tables, **params)
