"""Oracle parity at the benchmark configurations' full sizes: q1.1 on the
6,000,000-row table, the q3.1 join on 60,000,000 rows and the q2.x join on
the 600,000,000-row table, each through the compiled fused path (tuned in
its first execution), the unfused per-operator path and, at 600M rows, a
batch of one pipeline per template -- against the plan oracle on the same
seeded tables, generated on the host with only the columns the plans read."""
import pytest

pytestmark = pytest.mark.gpu

from oracle import datagen, plan_oracle as PO  # noqa: E402
from sycldb_amd import plans as PL  # noqa: E402
from sycldb_amd import schema as S  # noqa: E402

SEED = 45

CONFIGS = [
    ("c1", 6_000_000, PL.config1_plan, (S.LO_ORDERDATE, S.LO_QUANTITY, S.LO_EXTENDEDPRICE, S.LO_DISCOUNT)),
    ("c2", 60_000_000, PL.config2_plan, (S.LO_CUSTKEY, S.LO_SUPPKEY, S.LO_ORDERDATE, S.LO_REVENUE)),
]


@pytest.mark.parametrize("name,rows,builder,cols", CONFIGS)
def test_config_matches_oracle(name, rows, builder, cols):
    from sycldb_amd.engine import Engine
    plan = builder()
    want = PO.execute_plan(plan, datagen.generate_tables(S, rows, SEED, fact_columns=cols)).rows()
    assert want
    print(name, rows, "rows:", len(want), "result rows", flush=True)
    with Engine(codegen="compile") as eng:
        eng.generate(rows, SEED, fact_columns=cols)
        assert eng.execute(plan).rows() == want
        assert eng.execute(plan).rows() == want
        eng.set_fusion(False)
        assert eng.execute(plan).rows() == want


def test_c3_600m_matches_oracle():
    from sycldb_amd.engine import Engine
    rows = 600_000_000
    plan = PL.config3_plan()
    batch = [PL.build("t1", date=("year", 1994), disc=(4, 6), qty=("range", 26, 35)),
             PL.build("t2", part=("brand_range", 221, 228), s_region=2),
             PL.build("t3", cust=("nation", 14), supp=("nation", 14), date=("year_range_and", 1992, 1997),
                      group="city"),
             PL.build("t4", c_region=3, s_region=0, mfgr=(2, 4))]
    cols = (S.LO_CUSTKEY, S.LO_PARTKEY, S.LO_SUPPKEY, S.LO_ORDERDATE, S.LO_QUANTITY, S.LO_EXTENDEDPRICE,
            S.LO_DISCOUNT, S.LO_REVENUE, S.LO_SUPPLYCOST)
    tables = datagen.generate_tables(S, rows, SEED, fact_columns=cols)
    want = PO.execute_plan(plan, tables).rows()
    want_batch = [PO.execute_plan(p, tables).rows() for p in batch]
    del tables
    print("c3 600M: oracle done", len(want), [len(w) for w in want_batch], flush=True)
    with Engine(codegen="compile") as eng:
        eng.generate(rows, SEED, fact_columns=cols)
        assert eng.execute(plan).rows() == want
        assert eng.execute(plan).rows() == want
        print("c3 600M: fused ok", flush=True)
        results, nb = eng.execute_batch(batch)
        assert nb == len(batch)
        assert [r.rows() for r in results] == want_batch
        print("c3 600M: batch ok", flush=True)
        eng.set_fusion(False)
        assert eng.execute(plan).rows() == want
