"""Pins the plan-level oracle (oracle/plan_oracle.py): against the independent
brute-force SQL evaluator on every template, under segment splits and row
permutations, and against hand-computed results for the reference's filter
quirks (left-fold of logical operators, NONE overwriting the flags).  CPU."""
import numpy as np
import pytest

from oracle import bruteforce, datagen, plan_oracle as PO
from sycldb_amd import plans as PL
from sycldb_amd import schema as S

CASES = [
    ("t1", {}),
    ("t1", {"date": ("yearmonthnum", 199401), "disc": (4, 6), "qty": ("range", 26, 35)}),
    ("t1", {"date": ("year_range_and", 1992, 1993), "disc": (0, 10), "qty": ("lt", 51)}),
    ("t2", {}),
    ("t2", {"part": ("brand", 260), "s_region": 3}),
    ("t2", {"part": ("brand_range", 221, 228), "s_region": 2}),
    ("t3", {}),
    ("t3", {"cust": ("nation", 14), "supp": ("nation", 14), "date": ("year_range_and", 1992, 1997), "group": "city"}),
    ("t3", {"cust": ("city2", 141, 145), "supp": ("region", 2), "date": ("yearmonthnum", 199712),
            "group": "city"}),
    ("t4", {}),
    ("t4", {"c_region": 3, "s_region": 0, "mfgr": (2, 4)}),
    # ~15,000 live groups: overflows the fused kernel's LDS hash into the global path
    ("t3", {"cust": ("region", 1), "supp": ("region", 1), "date": ("year_range", 1992, 1997), "group": "city"}),
]


@pytest.fixture(scope="module")
def tables():
    return datagen.generate_tables(S, 300_000, seed=44)


@pytest.mark.parametrize("tmpl,params", CASES)
def test_oracle_matches_bruteforce(tables, tmpl, params):
    plan = PL.build(tmpl, **params)
    got = PO.execute_plan(plan, tables).rows()
    want = bruteforce.evaluate(tmpl, tables, **params)
    assert got == want
    assert len(got) > 0


@pytest.mark.parametrize("tmpl,params", [CASES[0], CASES[3], CASES[6], CASES[9]])
@pytest.mark.parametrize("seg", [1000, 4097, 65536])
def test_segment_split_invariance(tables, tmpl, params, seg):
    plan = PL.build(tmpl, **params)
    assert PO.execute_plan(plan, tables, segment_size=seg).rows() == PO.execute_plan(plan, tables).rows()


def test_fact_permutation_invariance(tables):
    perm = np.random.default_rng(9).permutation(len(tables["lineorder"][5]))
    shuffled = dict(tables)
    shuffled["lineorder"] = {c: np.ascontiguousarray(v[perm]) for c, v in tables["lineorder"].items()}
    for tmpl, params in CASES[:: 3]:
        plan = PL.build(tmpl, **params)
        assert PO.execute_plan(plan, shuffled).rows() == PO.execute_plan(plan, tables).rows()


def test_group_rows_in_dense_order(tables):
    rows = PO.execute_plan(PL.build("t3"), tables).rows()
    # first group column varies fastest: sorted by (year, s_nation, c_nation)
    assert rows == sorted(rows, key=lambda r: (r[2], r[1], r[0]))


def _tiny_lineorder(q, d, p):
    n = len(q)
    return {8: np.array(q, np.int32), 11: np.array(d, np.int32), 9: np.array(p, np.int32),
            5: np.array([19930101] * n, np.int32), 12: np.arange(n, dtype=np.int32)}


def test_left_fold_of_logical_filter():
    # OR(a, AND(b, c)) is applied as ((a OR b) AND c): the parent op goes to
    # the first operand, the node's own op to the rest
    # (models/transient_table.hpp:845-860)
    lo = _tiny_lineorder([1, 0, 1, 0], [0, 2, 9, 0], [0, 3, 3, 3])
    b = PL.PlanBuilder()
    b.scan("lineorder")
    b.filter(PL.expr("OR", PL.expr("=", PL.col(8), PL.lit(1)),
                     PL.expr("AND", PL.expr("=", PL.col(11), PL.lit(2)), PL.expr("=", PL.col(9), PL.lit(3)))))
    b.project([PL.col(12)])
    res = PO.execute_plan(b.plan(), {"lineorder": lo})
    # SQL would select rows 0, 1, 2; the reference's fold selects 1 and 2
    assert res.rows() == [(1,), (2,)]


def test_none_parent_overwrites_join_result():
    # a FILTER after a join starts with parent op "" -> NONE: its first
    # comparison replaces the flags, dropping the semi-join's result
    lo = {5: np.array([19930101, 19940101, 19930102, 19950101], np.int32),
          11: np.array([2, 2, 1, 1], np.int32), 9: np.array([10, 20, 30, 40], np.int32)}
    dd = {0: np.array([19930101, 19930102, 19940101, 19950101], np.int32),
          4: np.array([1993, 1993, 1994, 1995], np.int32), 5: np.array([199301, 199301, 199401, 199501], np.int32)}
    b = PL.PlanBuilder()
    lo_id = b.scan("lineorder")
    b.scan("ddate")
    ddf = b.filter(PL.expr("=", PL.col(4), PL.lit(1993)))
    b.join(lo_id, ddf, 5, 0, "semi")
    b.filter(PL.expr("=", PL.col(11), PL.lit(2)))
    b.project([PL.col(9)])
    b.aggregate([], 0)
    res = PO.execute_plan(b.plan(), {"lineorder": lo, "ddate": dd})
    assert res.rows() == [(30,)]          # rows with discount 2: 10 + 20


def test_dag_order_defers_lineorder():
    plan = PL.build("t2")
    order = PO.dag_topological_sort(plan["rels"])
    assert order == [1, 4, 7, 2, 5, 0, 3, 6, 8, 9, 10, 11]
    assert sorted(order) == list(range(len(plan["rels"])))


def test_parse_execution_info_group_columns():
    gb, _ = PO.parse_execution_info(PL.build("t3", group="city")["rels"])
    assert gb == {"customer": 3, "supplier": 3, "ddate": 4}
    gb, _ = PO.parse_execution_info(PL.build("t2")["rels"])
    assert gb == {"ddate": 4, "part": 4}


def test_full_join_without_group_column_fails():
    # the reference dereferences a null group_by_column there
    b = PL.PlanBuilder()
    lo = b.scan("lineorder")
    b.scan("supplier")
    sf = b.filter(PL.expr("=", PL.col(5), PL.lit(1)))
    b.join(lo, sf, 4, 0, "inner")
    tables = {"lineorder": {4: np.array([1, 2], np.int32)},
              "supplier": {0: np.array([1, 2], np.int32), 5: np.array([1, 0], np.int32)}}
    with pytest.raises(PO.OracleError):
        PO.execute_plan(b.plan(), tables)


# ---------------------------------------------------------------- quirk plans
# The plan oracle on the quirk plans the GPU engines are also checked with
# (tests/quirk_plans.py), against the rule each quirk fixes, written here as
# plain numpy over the same seeded tables: the left fold, the NONE overwrite
# after a join, literal fills and literal-operand arithmetic, and group slots
# of keys outside their recorded range.

def _lo(tables):
    lo = tables["lineorder"]
    return (lo[S.LO_QUANTITY].astype(np.int64), lo[S.LO_DISCOUNT].astype(np.int64),
            lo[S.LO_EXTENDEDPRICE].astype(np.int64), lo[S.LO_REVENUE].astype(np.int64))


def _scalar(v):
    return [(int(v) % 2**64,)]


def test_quirk_left_fold_or_and(tables):
    import quirk_plans as Q
    q, d, p, r = _lo(tables)
    mask = ((d == 2) | (q < 10)) & (p > 5_000_000)          # not (d == 2) | ((q < 10) & (p > 5e6))
    sql = (d == 2) | ((q < 10) & (p > 5_000_000))
    assert mask.sum() != sql.sum()
    assert PO.execute_plan(Q.PLANS["left_fold_or_and"], tables).rows() == _scalar(r[mask].sum())


def test_quirk_left_fold_and_or(tables):
    import quirk_plans as Q
    q, d, p, r = _lo(tables)
    # a gets NONE, the OR node gets AND: b gets AND, c and e get OR
    mask = (((q < 30) & (d == 1)) | (d == 9)) | (p > 9_000_000)
    want = [(k, int(q[mask & (d == k)].sum())) for k in range(11) if (mask & (d == k)).any()]
    assert PO.execute_plan(Q.PLANS["left_fold_and_or"], tables).rows() == want


def test_quirk_filter_after_join_overwrites(tables):
    import quirk_plans as Q
    q, d, p, r = _lo(tables)
    assert PO.execute_plan(Q.PLANS["filter_after_join"], tables).rows() == _scalar(p[d == 2].sum())


def test_quirk_literal_projections(tables):
    import quirk_plans as Q
    q, d, p, r = _lo(tables)
    sel = q < 25
    assert PO.execute_plan(Q.PLANS["literal_first"], tables).rows() == _scalar(7 * sel.sum())
    assert PO.execute_plan(Q.PLANS["col_times_lit"], tables).rows() == _scalar((3 * q[sel]).sum())
    assert PO.execute_plan(Q.PLANS["lit_minus_col"], tables).rows() == _scalar((100 - d[sel]).sum())
    assert PO.execute_plan(Q.PLANS["col_div_lit"], tables).rows() == _scalar((r[sel] // 1000).sum())


def test_quirk_group_slots_outside_recorded_range(tables):
    # 100 - discount has the recorded range of discount (0..10): the slot of
    # key k is k % 11, and groups come out in slot order
    import quirk_plans as Q
    q, d, p, r = _lo(tables)
    sel = q < 25
    keys = sorted({int(k) for k in (100 - d[sel])}, key=lambda k: k % 11)
    want = [(k, int((3 * q[sel & (100 - d == k)]).sum())) for k in keys]
    assert [k for k, _ in want][:3] == [99, 100, 90]
    assert PO.execute_plan(Q.PLANS["group_by_lit_minus_col"], tables).rows() == want
