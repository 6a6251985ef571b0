"""Hand-built physical plans for the benchmark templates.

The reference receives its plans from a Calcite planner over Thrift
(server/thrift/calciteserver.thrift:1-80; JSON -> Thrift conversion in
server/src/main/java/com/eurecom/calcite/SycldbJsonConverter.java:10-163).
That planner cannot run offline, so plans are built here as JSON documents
with the same field names as the Thrift structs (``RelNode``, ``ExprType``,
``LiteralType``, ``AggType``, ``CollationType``).  Column indices follow the
reference executor's row model: a scan exposes every column of the table
(``table_column_numbers``), a join exposes left ++ right columns (also for
semi joins, transient_table.hpp:1766-1767), a project exposes its exprs.

Templates (the SSB query flights the configurations are drawn from):
  * T1 -- q1.x (server/queries/original/q11.sql): semi-join date, filter
    discount/quantity, SUM(extendedprice * discount)
  * T2 -- q2.x (queries/transformed/q21.sql): part, supplier, date;
    GROUP BY year, brand
  * T3 -- q3.x (queries/transformed/q31.sql): customer, supplier, date;
    GROUP BY c_nation|c_city, s_nation|s_city, year
  * T4 -- q4.x (queries/transformed/q41.sql): customer, supplier, part,
    date; GROUP BY year, c_nation; SUM(revenue - supplycost)
"""
import json

from . import schema as S


def col(i, name=None):
    return {"exprType": "COLUMN", "input": int(i), "name": name or "$%d" % i}


def lit(v, typ="INTEGER"):
    return {"exprType": "LITERAL", "literal": {"literalOption": "LITERAL", "value": int(v)},
            "type": typ}


def closed_range(lo, hi):
    """A Sarg literal with one closed range (SEARCH reads rangeSet[0][1..2],
    models/models.hpp:633-636)."""
    return {"exprType": "LITERAL", "type": "INTEGER",
            "literal": {"literalOption": "RANGE", "rangeSet": [["closed", str(lo), str(hi)]]}}


def two_points(a, b):
    """A Sarg literal with two points (models/models.hpp:669-672)."""
    return {"exprType": "LITERAL", "type": "INTEGER",
            "literal": {"literalOption": "RANGE",
                        "rangeSet": [["singleton", str(a)], ["singleton", str(b)]]}}


def expr(op, *operands, typ="BOOLEAN"):
    return {"exprType": "EXPR", "op": op, "operands": list(operands), "type": typ}


def search(column, sarg):
    return expr("SEARCH", column, sarg)


class PlanBuilder:
    """Appends RelNodes with consecutive ids and tracks each node's width."""

    def __init__(self):
        self.rels = []
        self.width = []

    def _add(self, rel, width):
        rel["id"] = len(self.rels)
        self.rels.append(rel)
        self.width.append(width)
        return rel["id"]

    def scan(self, table):
        return self._add({"relOp": "TABLE_SCAN", "tables": ["ssb", table], "inputs": []},
                         S.TABLE_COLUMN_NUMBERS[table])

    def filter(self, condition):
        return self._add({"relOp": "FILTER", "condition": condition}, self.width[-1])

    def join(self, left, right, left_col, right_col, join_type="inner"):
        cond = expr("=", col(left_col), col(self.width[left] + right_col))
        return self._add({"relOp": "JOIN", "inputs": [left, right], "condition": cond,
                          "joinType": join_type}, self.width[left] + self.width[right])

    def offset(self, node):
        """Column offset of a join's right side (the width of its left)."""
        return self.width[node]

    def project(self, exprs, fields=None):
        fields = fields or ["f%d" % i for i in range(len(exprs))]
        return self._add({"relOp": "PROJECT", "exprs": exprs, "fields": fields}, len(exprs))

    def aggregate(self, group, operand, name="agg"):
        aggs = [{"agg": "SUM", "operands": [operand], "name": name, "type": "BIGINT",
                 "distinct": False}]
        return self._add({"relOp": "AGGREGATE", "group": list(group), "aggs": aggs},
                         len(group) + 1)

    def sort(self, fields):
        coll = [{"field": f, "direction": "ASCENDING", "nulls": "LAST"} for f in fields]
        return self._add({"relOp": "SORT", "collation": coll}, self.width[-1])

    def plan(self):
        return {"rels": self.rels}


# ----------------------------------------------------------------- predicates

def _date_pred(p):
    kind = p[0]
    if kind == "year":
        return expr("=", col(S.D_YEAR), lit(p[1]))
    if kind == "year_range":
        return search(col(S.D_YEAR), closed_range(p[1], p[2]))
    if kind == "year_range_and":
        return expr("AND", expr(">=", col(S.D_YEAR), lit(p[1])), expr("<=", col(S.D_YEAR), lit(p[2])))
    if kind == "yearmonthnum":
        return expr("=", col(S.D_YEARMONTHNUM), lit(p[1]))
    raise ValueError(p)


def _ent_pred(p, region_col, nation_col, city_col):
    kind = p[0]
    if kind == "region":
        return expr("=", col(region_col), lit(p[1]))
    if kind == "nation":
        return expr("=", col(nation_col), lit(p[1]))
    if kind == "city":
        return expr("=", col(city_col), lit(p[1]))
    if kind == "city2":
        return search(col(city_col), two_points(p[1], p[2]))
    raise ValueError(p)


def _part_pred(p):
    kind = p[0]
    if kind == "category":
        return expr("=", col(S.P_CATEGORY), lit(p[1]))
    if kind == "brand":
        return expr("=", col(S.P_BRAND1), lit(p[1]))
    if kind == "brand_range":
        return search(col(S.P_BRAND1), closed_range(p[1], p[2]))
    if kind == "mfgr2":
        return search(col(S.P_MFGR), two_points(p[1], p[2]))
    raise ValueError(p)


# ------------------------------------------------------------------ templates

def t1_plan(date=("year", 1993), disc=(1, 3), qty=("lt", 25)):
    """q1.x: SUM(lo_extendedprice * lo_discount) over a semi-join with date."""
    b = PlanBuilder()
    lo = b.scan("lineorder")
    d_pred = search(col(S.LO_DISCOUNT), closed_range(disc[0], disc[1]))
    if qty[0] == "lt":
        q_pred = expr("<", col(S.LO_QUANTITY), lit(qty[1]))
    else:
        q_pred = search(col(S.LO_QUANTITY), closed_range(qty[1], qty[2]))
    lof = b.filter(expr("AND", d_pred, q_pred))
    b.scan("ddate")
    ddf = b.filter(_date_pred(date))
    j = b.join(lof, ddf, S.LO_ORDERDATE, S.D_DATEKEY, "semi")
    assert j == ddf + 1
    b.project([expr("*", col(S.LO_EXTENDEDPRICE), col(S.LO_DISCOUNT), typ="INTEGER")], ["revenue"])
    b.aggregate([], 0, "revenue")
    return b.plan()


def t2_plan(part=("category", 1), s_region=1):
    """q2.x: GROUP BY d_year, p_brand1; SUM(lo_revenue)."""
    b = PlanBuilder()
    lo = b.scan("lineorder")
    b.scan("part")
    pf = b.filter(_part_pred(part))
    j1 = b.join(lo, pf, S.LO_PARTKEY, S.P_PARTKEY)
    brand_at = b.offset(lo) + S.P_BRAND1
    b.scan("supplier")
    sf = b.filter(_ent_pred(("region", s_region), S.S_REGION, S.S_NATION, S.S_CITY))
    j2 = b.join(j1, sf, S.LO_SUPPKEY, S.S_SUPPKEY, "semi")
    dd = b.scan("ddate")
    year_at = b.offset(j2) + S.D_YEAR
    j3 = b.join(j2, dd, S.LO_ORDERDATE, S.D_DATEKEY)
    b.project([col(year_at), col(brand_at), col(S.LO_REVENUE)], ["d_year", "p_brand1", "lo_revenue"])
    b.aggregate([0, 1], 2, "revenue")
    b.sort([0, 1])
    return b.plan()


def t3_plan(cust=("region", 2), supp=("region", 2), date=("year_range", 1992, 1997),
            group="nation"):
    """q3.x: GROUP BY c_<group>, s_<group>, d_year; SUM(lo_revenue)."""
    gcol = S.C_NATION if group == "nation" else S.C_CITY
    b = PlanBuilder()
    lo = b.scan("lineorder")
    b.scan("customer")
    cf = b.filter(_ent_pred(cust, S.C_REGION, S.C_NATION, S.C_CITY))
    j1 = b.join(lo, cf, S.LO_CUSTKEY, S.C_CUSTKEY)
    c_at = b.offset(lo) + gcol
    b.scan("supplier")
    sf = b.filter(_ent_pred(supp, S.S_REGION, S.S_NATION, S.S_CITY))
    s_at = b.offset(j1) + gcol
    j2 = b.join(j1, sf, S.LO_SUPPKEY, S.S_SUPPKEY)
    b.scan("ddate")
    df = b.filter(_date_pred(date))
    y_at = b.offset(j2) + S.D_YEAR
    b.join(j2, df, S.LO_ORDERDATE, S.D_DATEKEY)
    b.project([col(c_at), col(s_at), col(y_at), col(S.LO_REVENUE)],
              ["c_" + group, "s_" + group, "d_year", "lo_revenue"])
    b.aggregate([0, 1, 2], 3, "revenue")
    b.sort([2, 3])
    return b.plan()


def t4_plan(c_region=1, s_region=1, mfgr=(0, 1)):
    """q4.1: GROUP BY d_year, c_nation; SUM(lo_revenue - lo_supplycost)."""
    b = PlanBuilder()
    lo = b.scan("lineorder")
    b.scan("customer")
    cf = b.filter(expr("=", col(S.C_REGION), lit(c_region)))
    j1 = b.join(lo, cf, S.LO_CUSTKEY, S.C_CUSTKEY)
    n_at = b.offset(lo) + S.C_NATION
    b.scan("supplier")
    sf = b.filter(expr("=", col(S.S_REGION), lit(s_region)))
    j2 = b.join(j1, sf, S.LO_SUPPKEY, S.S_SUPPKEY, "semi")
    b.scan("part")
    pf = b.filter(_part_pred(("mfgr2", mfgr[0], mfgr[1])))
    j3 = b.join(j2, pf, S.LO_PARTKEY, S.P_PARTKEY, "semi")
    dd = b.scan("ddate")
    y_at = b.offset(j3) + S.D_YEAR
    b.join(j3, dd, S.LO_ORDERDATE, S.D_DATEKEY)
    b.project([col(y_at), col(n_at),
               expr("-", col(S.LO_REVENUE), col(S.LO_SUPPLYCOST), typ="INTEGER")],
              ["d_year", "c_nation", "profit"])
    b.aggregate([0, 1], 2, "profit")
    b.sort([0, 1])
    return b.plan()


TEMPLATES = {"t1": t1_plan, "t2": t2_plan, "t3": t3_plan, "t4": t4_plan}


def build(template, **params):
    return TEMPLATES[template](**params)


def to_json(plan):
    return json.dumps(plan, separators=(",", ":"))


# ------------------------------------------------------ benchmark configurations

def config1_plan():
    """6M-row CPU configuration: server/queries/original/q11.sql."""
    return t1_plan(("year", 1993), (1, 3), ("lt", 25))


def config2_plan(region=2):
    """60M-row configuration: both dimensions to one region, 6-year range."""
    return t3_plan(("region", region), ("region", region), ("year_range", 1992, 1997), "nation")


def config3_plan(category=1, region=1):
    """600M-row configuration: part category (4 %) and supplier region (20 %)."""
    return t2_plan(("category", category), region)


# ----------------------------------------------------------- random pipelines

def random_pipelines(n, seed):
    """n (template, params) pairs drawn from the four templates with seeded
    random predicate constants, spanning fact selectivities from about 1e-6
    (a city pair x a city x one month) to about 0.2 (one region)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        t = ("t1", "t2", "t3", "t4")[i % 4]
        if t == "t1":
            kind = rng.integers(0, 3)
            if kind == 0:
                date = ("year", int(rng.integers(1992, 1999)))
            elif kind == 1:
                y = int(rng.integers(1992, 1999))
                date = ("yearmonthnum", y * 100 + int(rng.integers(1, 13)))
            else:
                a = int(rng.integers(1992, 1998))
                date = ("year_range_and", a, a + int(rng.integers(0, 3)))
            lo = int(rng.integers(0, 9))
            disc = (lo, lo + int(rng.integers(0, 3)))
            if rng.integers(0, 2):
                qty = ("lt", int(rng.integers(2, 51)))
            else:
                q = int(rng.integers(1, 45))
                qty = ("range", q, q + int(rng.integers(0, 6)))
            out.append(("t1", {"date": date, "disc": disc, "qty": qty}))
        elif t == "t2":
            kind = rng.integers(0, 3)
            if kind == 0:
                part = ("category", int(rng.integers(0, 25)))
            elif kind == 1:
                part = ("brand", int(rng.integers(0, 1000)))
            else:
                b0 = int(rng.integers(0, 990))
                part = ("brand_range", b0, b0 + int(rng.integers(1, 9)))
            out.append(("t2", {"part": part, "s_region": int(rng.integers(0, 5))}))
        elif t == "t3":
            def ent():
                k = rng.integers(0, 4)
                if k == 0:
                    return ("region", int(rng.integers(0, 5)))
                if k == 1:
                    return ("nation", int(rng.integers(0, 25)))
                if k == 2:
                    return ("city", int(rng.integers(0, 250)))
                c = int(rng.integers(0, 248))
                return ("city2", c, c + int(rng.integers(1, 3)))
            if rng.integers(0, 3):
                a = int(rng.integers(1992, 1998))
                date = ("year_range", a, a + int(rng.integers(0, 6 - (a - 1992) + 1)))
            else:
                date = ("yearmonthnum", int(rng.integers(1992, 1999)) * 100 + int(rng.integers(1, 13)))
            out.append(("t3", {"cust": ent(), "supp": ent(), "date": date,
                               "group": "city" if rng.integers(0, 2) else "nation"}))
        else:
            m = int(rng.integers(0, 4))
            out.append(("t4", {"c_region": int(rng.integers(0, 5)), "s_region": int(rng.integers(0, 5)),
                               "mfgr": (m, m + 1)}))
    return out
