"""The reference executor's quirks on the GPU engines: the left fold of
nested logical filters and the NONE overwrite of a filter after a join
(models/transient_table.hpp:845-860), LITERAL and literal-operand
projections (1003-1024, 1109-1203), and group slots of keys outside their
recorded range -- through the compiled fused, interpreted, unfused,
segmented and batched engines, against the plan oracle on the same seeded
tables (the oracle is pinned on these plans by tests/test_oracle_plan.py)."""
import pytest

pytestmark = pytest.mark.gpu

from oracle import datagen, plan_oracle as PO  # noqa: E402
from sycldb_amd import plans as PL  # noqa: E402
from sycldb_amd import schema as S  # noqa: E402

import quirk_plans as Q  # noqa: E402

FACT_ROWS = 400_003
SEED = 46
QUEUED = 16


@pytest.fixture(scope="module")
def cpu_tables():
    return datagen.generate_tables(S, FACT_ROWS, seed=SEED)


@pytest.fixture(scope="module")
def expected(cpu_tables):
    return {k: PO.execute_plan(p, cpu_tables).rows() for k, p in Q.PLANS.items()}


@pytest.mark.parametrize("kw", [dict(codegen="compile"), dict(), dict(fusion=False), dict(codegen="interpret"),
                                dict(segment_size=77_777, codegen="compile"), dict(unroll=4 | QUEUED),
                                dict(segment_size=50_000, fusion=False)])
def test_quirk_plans_match_oracle(expected, kw):
    from sycldb_amd.engine import Engine
    print("engine", kw, flush=True)
    with Engine(**kw) as eng:
        eng.generate(FACT_ROWS, SEED)
        for name, plan in Q.PLANS.items():
            print("plan", name, flush=True)
            assert eng.execute(plan).rows() == expected[name], (kw, name)


def test_quirk_plans_batched(expected, cpu_tables):
    """All quirk plans in one batch, mixed with two template plans."""
    from sycldb_amd.engine import Engine
    names = list(Q.PLANS)
    plans = [Q.PLANS[k] for k in names] + [PL.build("t1"), PL.build("t2")]
    want = [expected[k] for k in names] + [PO.execute_plan(p, cpu_tables).rows() for p in plans[len(names):]]
    with Engine() as eng:
        eng.generate(FACT_ROWS, SEED)
        for rep in range(2):
            print("batch", rep, flush=True)
            results, nb = eng.execute_batch(plans)
            assert nb >= 2
            for name, r, w in zip(names + ["t1", "t2"], results, want):
                assert r.rows() == w, name
