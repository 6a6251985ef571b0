"""Batched execution (sdb_engine_execute_batch): many pipelines whose
fact-table passes share one scan.  Every pipeline's result must equal the
oracle's and the plain engine's result for that pipeline alone."""
import pytest

pytestmark = pytest.mark.gpu

from oracle import datagen, plan_oracle as PO  # noqa: E402
from sycldb_amd import plans as PL  # noqa: E402
from sycldb_amd import schema as S  # noqa: E402

FACT_ROWS = 300_007
SEED = 46


@pytest.fixture(scope="module")
def cpu_tables():
    return datagen.generate_tables(S, FACT_ROWS, seed=SEED)


@pytest.fixture(scope="module")
def pipelines():
    return PL.random_pipelines(32, 11)


@pytest.fixture(scope="module")
def expected(cpu_tables, pipelines):
    return [PO.execute_plan(PL.build(t, **p), cpu_tables).rows() for t, p in pipelines]


def _engine(**kw):
    from sycldb_amd.engine import Engine
    e = Engine(**kw)
    e.generate(FACT_ROWS, SEED)
    return e


def test_batch_of_32_matches_oracle(pipelines, expected):
    eng = _engine()
    try:
        plans = [PL.build(t, **p) for t, p in pipelines]
        print("batch of", len(plans), flush=True)
        results, batched = eng.execute_batch(plans)
        assert batched >= 24, batched
        for i, r in enumerate(results):
            assert r.rows() == expected[i], (i, pipelines[i])
        # the same batch again (allocator reset between batches)
        results2, _ = eng.execute_batch(plans)
        assert [r.rows() for r in results2] == expected
        # each plan alone through the plain path agrees
        for i in range(0, len(plans), 5):
            assert eng.execute(plans[i]).rows() == expected[i]
    finally:
        eng.close()


@pytest.mark.parametrize("kw", [dict(segment_size=77_777), dict(unroll=1), dict(codegen="interpret")])
def test_batch_variants_match_oracle(pipelines, expected, kw):
    """Several segments per pipeline; another chunking; and an engine whose
    fused bundles are interpreted (the batch then runs each plan plainly)."""
    eng = _engine(**kw)
    try:
        plans = [PL.build(t, **p) for t, p in pipelines[:12]]
        print("batch", kw, flush=True)
        results, batched = eng.execute_batch(plans)
        for i, r in enumerate(results):
            assert r.rows() == expected[i], (kw, i, pipelines[i])
    finally:
        eng.close()


def test_batch_with_unfused_engine_runs_plainly(pipelines, expected):
    eng = _engine(fusion=False)
    try:
        plans = [PL.build(t, **p) for t, p in pipelines[:4]]
        results, batched = eng.execute_batch(plans)
        assert batched == 0
        assert [r.rows() for r in results] == expected[:4]
    finally:
        eng.close()


def test_batch_of_one_and_empty():
    eng = _engine()
    try:
        results, batched = eng.execute_batch([])
        assert results == [] and batched == 0
        plan = PL.build("t2")
        results, _ = eng.execute_batch([plan])
        assert results[0].rows() == eng.execute(plan).rows()
    finally:
        eng.close()


def test_batch_without_shared_masks(pipelines, expected, monkeypatch):
    """SDB_BATCH_MASKS=0: every probe stays inside its program."""
    monkeypatch.setenv("SDB_BATCH_MASKS", "0")
    eng = _engine()
    try:
        plans = [PL.build(t, **p) for t, p in pipelines[:12]]
        print("batch without masks", flush=True)
        results, batched = eng.execute_batch(plans)
        assert batched >= 8
        for i, r in enumerate(results):
            assert r.rows() == expected[i], (i, pipelines[i])
    finally:
        eng.close()


MASK_READS = [{}, {"SDB_BATCH_LAZY_MASKS": "1"}, {"SDB_BATCH_LAZY_MASKS": "0"},
              {"SDB_BATCH_DICT_MASKS": "1"}, {"SDB_BATCH_DICT_MASKS": "1", "SDB_BATCH_LAZY_MASKS": "1"},
              {"SDB_BATCH_DICT_MASKS": "1", "SDB_BATCH_DICT_MIN_LEN": "1024"}]


@pytest.mark.parametrize("env", MASK_READS)
def test_batch_mask_edge_cases(cpu_tables, monkeypatch, env):
    """Shared probe masks at the extremes: a pipeline no row passes (a year
    outside the date table), near-empty ones, identical pipelines side by
    side, and every template's widest predicates; with the mask tables read
    as the estimate decides, always lazily, always eagerly, and through rank
    dictionaries (the date table's; with a lower length bound also the
    supplier's, and the customer's, whose set keys overflow the dictionary so
    that the kernel reads the table instead)."""
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    lazy = env
    pipes = [
        ("t1", {"date": ("year", 2050), "disc": (1, 3), "qty": ("lt", 25)}),
        ("t2", {"part": ("brand", 999), "s_region": 4}),
        ("t2", {"part": ("brand", 999), "s_region": 4}),
        ("t3", {"cust": ("city", 249), "supp": ("city", 0), "date": ("yearmonthnum", 199812),
                "group": "city"}),
        ("t3", {"cust": ("region", 0), "supp": ("region", 0), "date": ("year_range", 1992, 1998),
                "group": "nation"}),
        ("t4", {"c_region": 0, "s_region": 4, "mfgr": (0, 4)}),
        ("t1", {"date": ("year_range_and", 1992, 1998), "disc": (0, 10), "qty": ("lt", 51)}),
        ("t2", {"part": ("brand_range", 0, 999), "s_region": 0}),
    ]
    expected = [PO.execute_plan(PL.build(t, **p), cpu_tables).rows() for t, p in pipes]
    eng = _engine()
    try:
        plans = [PL.build(t, **p) for t, p in pipes]
        print("mask edge batch", lazy, flush=True)
        results, batched = eng.execute_batch(plans)
        assert batched == len(plans)
        for i, r in enumerate(results):
            assert r.rows() == expected[i], (i, pipes[i])
    finally:
        eng.close()


@pytest.mark.parametrize("env", [{}, {"SDB_BATCH_LAZY_MASKS": "1"}, {"SDB_BATCH_DICT_MASKS": "1"},
                                 {"SDB_BATCH_DICT_MASKS": "1", "SDB_BATCH_LAZY_MASKS": "1"}])
def test_batch_irregular_date_keys(cpu_tables, pipelines, monkeypatch, env):
    """Fact date keys that are no dates (day 32, month 13, zero, negative,
    far years), inside and outside the probed key range, and a date table
    holding such a key: the shared masks' range tests and default words must
    still match the oracle (mask tables also read lazily, and the date table
    as a rank dictionary)."""
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    lazy = env
    import numpy as np
    from sycldb_amd.engine import Engine
    tables = {t: {c: np.array(a, dtype=np.int32, copy=True) for c, a in cols.items()} for t, cols in cpu_tables.items()}
    od = tables["lineorder"][S.LO_ORDERDATE]
    odd = [19920132, 19921301, 19920230, 19930000, 19991301, 0, -7, 2147483647, 19920101, 19981231]
    rng = np.random.default_rng(7)
    for k, v in enumerate(odd):
        od[rng.choice(len(od), 500, replace=False)] = v
    dk = tables["ddate"][S.D_DATEKEY]
    dk[-1] = 19921301            # a date-table key that no calendar slot names
    dk[-2] = 19920230            # one that a slot names but no generated row holds
    pipes = pipelines[:16] + [("t1", {"date": ("year", 1992), "disc": (0, 10), "qty": ("lt", 51)})]
    expected = [PO.execute_plan(PL.build(t, **p), tables).rows() for t, p in pipes]
    eng = Engine()
    try:
        eng.load_tables(tables)
        plans = [PL.build(t, **p) for t, p in pipes]
        print("calendar batch", lazy, flush=True)
        results, batched = eng.execute_batch(plans)
        assert batched == len(plans)
        for i, r in enumerate(results):
            assert r.rows() == expected[i], (i, pipes[i])
    finally:
        eng.close()


@pytest.mark.parametrize("kw", [dict(), dict(segment_size=77_777)])
def test_batch_graph_replays_match_oracle(pipelines, expected, cpu_tables, kw):
    """Batched launches captured once into a hipGraph and replayed: three
    batches in a row, with plain executions and synchronisations between
    them, each against the oracle; then the same batch as plain stream
    operations.  Several segments replay one graph per segment."""
    import torch
    from sycldb_amd.engine import batch_graph_stats, set_batch_graphs
    eng = _engine(**kw)
    try:
        plans = [PL.build(t, **p) for t, p in pipelines]
        set_batch_graphs(True)
        c0, r0 = batch_graph_stats()
        for rep in range(3):
            print("graph batch", kw, rep, flush=True)
            results, batched = eng.execute_batch(plans)
            assert batched >= 24
            assert [r.rows() for r in results] == expected, rep
            # eager work of the program between the replays
            assert eng.execute(plans[rep]).rows() == expected[rep]
            torch.cuda.synchronize()
        c1, r1 = batch_graph_stats()
        # the first batch of a geometry runs as stream operations, the later
        # ones replay its graph
        assert r1 - r0 >= 2
        assert 1 <= c1 - c0 <= 4
        # the same shapes with other constants: the replayed graph copies the
        # new programs and probe descriptors
        pipes2 = []
        for t, p in pipelines:
            p = dict(p)
            if t == "t1" and "qty" in p:
                q = p["qty"]
                p["qty"] = ("lt", q[1] + 3) if q[0] == "lt" else ("range", q[1], q[2] + 1)
            if t == "t2" and "s_region" in p:
                p["s_region"] = (p["s_region"] + 1) % 5
            pipes2.append((t, p))
        want2 = [PO.execute_plan(PL.build(t, **p), cpu_tables).rows() for t, p in pipes2]
        plans2 = [PL.build(t, **p) for t, p in pipes2]
        for rep in range(2):
            print("graph batch, other constants", kw, rep, flush=True)
            results, _ = eng.execute_batch(plans2)
            assert [r.rows() for r in results] == want2, rep
            results, _ = eng.execute_batch(plans)
            assert [r.rows() for r in results] == expected, rep
        c1, r1 = batch_graph_stats()
        set_batch_graphs(False)
        results, _ = eng.execute_batch(plans)
        assert [r.rows() for r in results] == expected
        c2, r2 = batch_graph_stats()
        assert r2 == r1
    finally:
        set_batch_graphs(None)
        eng.close()


@pytest.mark.parametrize("env", [{"SDB_BATCH_WIDE_MASKS": "1"}, {"SDB_BATCH_WIDE_MASKS": "1", "SDB_BATCH_BLOCK": "1024"},
                                 {"SDB_BATCH_BLOCK": "1024", "SDB_BATCH_DOUBLE_BUFFER": "1"},
                                 {"SDB_BATCH_STAGE_MASKS": "1", "SDB_BATCH_BLOCK": "256"},
                                 {"SDB_BATCH_STAGE_MASKS": "1", "SDB_BATCH_BLOCK": "1024"},
                                 {"SDB_BATCH_STAGE_MASKS": "1", "SDB_BATCH_WIDE_MASKS": "1"},
                                 {"SDB_BATCH_LAZY_MASKS": "1"},
                                 {"SDB_BATCH_LAZY_MASKS": "1", "SDB_BATCH_STAGE_MASKS": "1", "SDB_BATCH_BLOCK": "1024"},
                                 {"SDB_BATCH_LAZY_MASKS": "1", "SDB_BATCH_WIDE_MASKS": "1"},
                                 {"SDB_BATCH_DICT_MASKS": "1"}, {"SDB_BATCH_DICT_MASKS": "1", "SDB_BATCH_BLOCK": "1024"},
                                 {"SDB_BATCH_DICT_MASKS": "1", "SDB_BATCH_WIDE_MASKS": "1"},
                                 {"SDB_BATCH_DICT_MASKS": "1", "SDB_BATCH_DICT_MIN_LEN": "1024"}])
def test_batch_layouts_match_oracle(pipelines, expected, monkeypatch, env):
    """The autotuner's other candidates, forced: whole 32-bit mask words
    instead of narrow windows, mask tables staged in LDS, other launch
    shapes, and mask tables read lazily in the estimated order."""
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    eng = _engine()
    try:
        plans = [PL.build(t, **p) for t, p in pipelines]
        print("batch layout", env, flush=True)
        results, batched = eng.execute_batch(plans)
        assert batched >= 24
        for i, r in enumerate(results):
            assert r.rows() == expected[i], (env, i, pipelines[i])
    finally:
        eng.close()


@pytest.mark.parametrize("env", MASK_READS)
def test_batch_per_template_matches_oracle(pipelines, expected, monkeypatch, env):
    """Batches of one template each (the lazy form's case: every program
    probes every masked column), in one engine whose shape library already
    holds all four templates, so masked slots the batch does not use are
    skipped."""
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    lazy = env
    eng = _engine()
    try:
        plans = [PL.build(t, **p) for t, p in pipelines]
        results, _ = eng.execute_batch(plans)
        assert [r.rows() for r in results] == expected
        for tmpl in ("t1", "t2", "t3", "t4"):
            idx = [i for i, (t, _) in enumerate(pipelines) if t == tmpl]
            if len(idx) < 2:
                continue
            print("template batch", tmpl, lazy, flush=True)
            results, batched = eng.execute_batch([plans[i] for i in idx])
            assert batched == len(idx)
            for j, r in zip(idx, results):
                assert r.rows() == expected[j], (tmpl, j, pipelines[j])
    finally:
        eng.close()
