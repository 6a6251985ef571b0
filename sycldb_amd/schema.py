"""Star-schema layout and the synthetic data specification.

Table and column numbering follow the reference's SSB loader:
``table_column_numbers`` / ``table_column_indices`` (kernels/types.cpp:3-21)
and the Calcite schema (server/src/main/java/com/eurecom/calcite/
SycldbSchema.java:138-262).  Column files are named ``<TABLE><index>`` in
upper case (models/models.hpp:1676-1679).

The generator specification below is the single definition of the seeded
synthetic data; the device generator (``Engine.generate``) and the CPU one used
by the tests consume the same entries.  Every value is a pure function of
(seed, stream, row), so any row range can be produced independently.
"""

TABLE_COLUMN_NUMBERS = {"lineorder": 17, "part": 9, "supplier": 7, "customer": 8, "ddate": 17}
TABLE_COLUMN_INDICES = {
    "lineorder": (2, 3, 4, 5, 8, 9, 11, 12, 13),
    "part": (0, 2, 3, 4),
    "supplier": (0, 3, 4, 5),
    "customer": (0, 3, 4, 5),
    "ddate": (0, 4, 5),
}
TABLE_IDS = {"lineorder": 1, "part": 2, "supplier": 3, "customer": 4, "ddate": 5}

# lineorder columns
LO_CUSTKEY, LO_PARTKEY, LO_SUPPKEY, LO_ORDERDATE = 2, 3, 4, 5
LO_QUANTITY, LO_EXTENDEDPRICE, LO_DISCOUNT, LO_REVENUE, LO_SUPPLYCOST = 8, 9, 11, 12, 13
# part columns
P_PARTKEY, P_MFGR, P_CATEGORY, P_BRAND1 = 0, 2, 3, 4
# supplier / customer columns
S_SUPPKEY, S_CITY, S_NATION, S_REGION = 0, 3, 4, 5
C_CUSTKEY, C_CITY, C_NATION, C_REGION = 0, 3, 4, 5
# date columns
D_DATEKEY, D_YEAR, D_YEARMONTHNUM = 0, 4, 5

# Dimension cardinalities of the benchmark configurations.  2,556 days from
# 1992-01-01 (to 1998-12-30), as the planner's ddate row count
# (SycldbSchema.java:259).
DIM_SIZES = {"customer": 300_000, "supplier": 20_000, "part": 1_400_000, "ddate": 2556}
DATE_FIRST_YEAR, DATE_LAST_YEAR = 1992, 1998


def stream_id(table, col):
    return TABLE_IDS[table] * 64 + col


def column_spec(table, col, sizes=None):
    """Generator of one column.

    Returns one of
      ("rowid",)                      value = row + 1 (dimension keys)
      ("uniform", lo, hi)             uniform integer in [lo, hi]
      ("child", base_col, mult, lo, hi)  base * mult + uniform[lo, hi]
      ("lookup", table, col)          a uniformly drawn row's value of that
                                      dimension column (lo_orderdate)
      ("date", col)                   the date dimension's own columns
    """
    sizes = dict(DIM_SIZES, **(sizes or {}))
    if table == "lineorder":
        return {
            LO_CUSTKEY: ("uniform", 1, sizes["customer"]),
            LO_PARTKEY: ("uniform", 1, sizes["part"]),
            LO_SUPPKEY: ("uniform", 1, sizes["supplier"]),
            LO_ORDERDATE: ("lookup", "ddate", D_DATEKEY),
            LO_QUANTITY: ("uniform", 1, 50),
            LO_EXTENDEDPRICE: ("uniform", 90_000, 10_500_000),
            LO_DISCOUNT: ("uniform", 0, 10),
            LO_REVENUE: ("uniform", 100, 1_000_000),
            LO_SUPPLYCOST: ("uniform", 100, 100_000),
        }[col]
    if table == "part":
        return {
            P_PARTKEY: ("rowid",),
            P_MFGR: ("uniform", 0, 4),
            P_CATEGORY: ("child", P_MFGR, 5, 0, 4),
            P_BRAND1: ("child", P_CATEGORY, 40, 0, 39),
        }[col]
    if table in ("supplier", "customer"):
        return {
            0: ("rowid",),
            5: ("uniform", 0, 4),
            4: ("child", 5, 5, 0, 4),
            3: ("child", 4, 10, 0, 9),
        }[col]
    if table == "ddate":
        return ("date", col)
    raise KeyError(table)


def generation_order(table):
    """Column order in which a table can be generated (parents first)."""
    return {
        "lineorder": TABLE_COLUMN_INDICES["lineorder"],
        "part": (P_PARTKEY, P_MFGR, P_CATEGORY, P_BRAND1),
        "supplier": (0, 5, 4, 3),
        "customer": (0, 5, 4, 3),
        "ddate": (D_DATEKEY, D_YEAR, D_YEARMONTHNUM),
    }[table]


def table_rows(table, fact_rows, sizes=None):
    sizes = dict(DIM_SIZES, **(sizes or {}))
    return fact_rows if table == "lineorder" else sizes[table]
