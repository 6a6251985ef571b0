"""Plans end to end on the GPU engine (TransientTable model, fused and
unfused bundles, several segment sizes) against the plan oracle on the same
seeded tables; the device generator against the CPU one; reference error
behaviour."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import datagen, plan_oracle as PO  # noqa: E402
from sycldb_amd import plans as PL  # noqa: E402
from sycldb_amd import schema as S  # noqa: E402

FACT_ROWS = 400_003
SEED = 45
CASES = [
    ("t1", {}),
    ("t1", {"date": ("yearmonthnum", 199401), "disc": (4, 6), "qty": ("range", 26, 35)}),
    ("t1", {"date": ("year_range_and", 1992, 1993), "disc": (0, 10), "qty": ("lt", 51)}),
    ("t2", {}),
    ("t2", {"part": ("brand", 260), "s_region": 3}),
    ("t2", {"part": ("brand_range", 221, 228), "s_region": 2}),
    ("t3", {}),
    ("t3", {"cust": ("nation", 14), "supp": ("nation", 14), "date": ("year_range_and", 1992, 1997), "group": "city"}),
    ("t3", {"cust": ("city2", 141, 145), "supp": ("region", 2), "date": ("yearmonthnum", 199712), "group": "city"}),
    ("t4", {}),
    ("t4", {"c_region": 3, "s_region": 0, "mfgr": (2, 4)}),
    # ~15,000 live groups: overflows the fused kernel's LDS hash into the global path
    ("t3", {"cust": ("region", 1), "supp": ("region", 1), "date": ("year_range", 1992, 1997), "group": "city"}),
]


@pytest.fixture(scope="module")
def cpu_tables():
    return datagen.generate_tables(S, FACT_ROWS, seed=SEED)


@pytest.fixture(scope="module")
def expected(cpu_tables):
    return {i: PO.execute_plan(PL.build(t, **p), cpu_tables).rows() for i, (t, p) in enumerate(CASES)}


QUEUED = 16  # sycldb_amd.engine.UNROLL_QUEUED (SDB_UNROLL_QUEUED; test_capi checks the value)


def make_engine(**kw):
    from sycldb_amd.engine import Engine
    eng = Engine(**kw)
    eng.generate(FACT_ROWS, SEED)
    return eng


@pytest.fixture(scope="module")
def fused():
    # compile mode: every execution runs the compiled kernels (the default
    # mode runs a new shape on the interpreter until they are ready)
    e = make_engine(codegen="compile")
    yield e
    e.close()


def test_device_generator_matches_cpu(fused, cpu_tables):
    for t, cols in cpu_tables.items():
        for c, want in cols.items():
            got = fused.download(t, c, len(want))
            assert (got == want).all(), (t, c)


@pytest.mark.parametrize("idx", range(len(CASES)))
def test_fused_matches_oracle(fused, expected, idx):
    tmpl, params = CASES[idx]
    print("case", idx, tmpl, params, flush=True)
    res = fused.execute(PL.build(tmpl, **params))
    assert res.rows() == expected[idx]


@pytest.mark.parametrize("kw", [dict(), dict(fusion=False), dict(segment_size=50_000), dict(segment_size=77_777),
                                dict(segment_size=77_777, fusion=False), dict(codegen="interpret"),
                                dict(codegen="interpret", segment_size=77_777), dict(unroll=1), dict(unroll=4),
                                dict(unroll=4, segment_size=77_777), dict(unroll=1 | QUEUED),
                                dict(unroll=2 | QUEUED), dict(unroll=4 | QUEUED),
                                dict(unroll=4 | QUEUED, segment_size=77_777)])
def test_unfused_and_segmented_match_oracle(expected, kw):
    print("engine", kw, flush=True)
    eng = make_engine(**kw)
    try:
        for idx, (tmpl, params) in enumerate(CASES):
            res = eng.execute(PL.build(tmpl, **params))
            assert res.rows() == expected[idx], (kw, idx)
    finally:
        eng.close()


def test_default_engine_compiles_fused_kernels(fused):
    """With hipRTC present the default engine runs runtime-compiled kernels
    (built on the background compiler thread; the interpreter runs a shape
    meanwhile, with the same results)."""
    from sycldb_amd import engine as E
    ok, why = E.codegen_available()
    assert ok, why
    plan = PL.build("t4", mfgr=(1, 5))
    first = fused.execute(plan).rows()
    assert E.codegen_wait(300), "background compilation did not finish"
    assert E.codegen_kernels_compiled() > 0
    assert fused.execute(plan).rows() == first


def test_fusion_launches_fewer_kernels(fused):
    from sycldb_amd.engine import Engine
    plan = PL.build("t2")
    fused.execute(plan)
    n_fused = fused.last_launches()
    with Engine(fusion=False) as u:
        u.generate(FACT_ROWS, SEED)
        u.execute(plan)
        n_unfused = u.last_launches()
    assert n_fused < n_unfused


BLOCK512, BLOCK1024 = 256, 512  # SDB_UNROLL_BLOCK512 / SDB_UNROLL_BLOCK1024
TUNE_ROWS = 4_500_011  # over the autotuner's 4M-row threshold


@pytest.fixture(scope="module")
def tune_tables():
    return datagen.generate_tables(S, TUNE_ROWS, seed=SEED + 1)


@pytest.fixture(scope="module")
def tune_case(tune_tables):
    plans = [PL.build("t2"), PL.build("t3")]
    return plans, [PO.execute_plan(p, tune_tables).rows() for p in plans]


def test_every_tuner_candidate_matches_oracle(tune_case):
    """Every launch configuration the autotuner chooses among (chunks per
    thread x threads per block, with and without the queued row tail),
    forced explicitly, returns the oracle's result."""
    from sycldb_amd.engine import Engine
    plans, want = tune_case
    eng = Engine()
    try:
        eng.generate(TUNE_ROWS, SEED + 1)
        for unroll in (1, 2, 4):
            for blk in (0, BLOCK512, BLOCK1024):
                for q in (0, QUEUED):
                    eng.set_codegen("auto", unroll | blk | q)
                    for i, p in enumerate(plans):
                        print("candidate", unroll, blk, q, i, flush=True)
                        assert eng.execute(p).rows() == want[i], (unroll, blk, q, i)
    finally:
        eng.close()


def _tuning_launches(report):
    """Launches that ran autotuner trials, over every program shape."""
    import re
    return sum(int(m) for m in re.findall(r"(\d+) tuning launches", report))


def test_autotuner_settles_in_one_execution(tune_case):
    """In compile mode (waiting for the background compiler) the first
    execution over 4M rows runs every trial on a scratch copy of the
    accumulators, then the program itself once: its result is the oracle's,
    and the shape is settled afterwards."""
    from sycldb_amd.engine import Engine, codegen_tuning_report
    plans, want = tune_case
    eng = Engine(codegen="compile")
    try:
        eng.generate(TUNE_ROWS, SEED + 1)
        for i, p in enumerate(plans):
            print("tuning execution", i, flush=True)
            assert eng.execute(p).rows() == want[i]
        report = codegen_tuning_report()
        print(report, flush=True)
        # settled: executing the plans again runs no further trial
        trials = _tuning_launches(report)
        for i, p in enumerate(plans):
            assert eng.execute(p).rows() == want[i]
        assert _tuning_launches(codegen_tuning_report()) == trials
    finally:
        eng.close()


def test_auto_mode_results_while_compiling(tune_tables):
    """The default mode runs a new shape on the interpreter while its kernels
    compile, then the tuned kernels: every execution returns the oracle's
    result (the q4.x template has not run over 4M rows before here)."""
    from sycldb_amd.engine import Engine, codegen_wait
    eng = Engine()
    try:
        eng.generate(TUNE_ROWS, SEED + 1)
        p = PL.build("t4", c_region=2, s_region=1, mfgr=(1, 3))
        w = PO.execute_plan(p, tune_tables).rows()
        for rep in range(3):
            assert eng.execute(p).rows() == w, rep
        assert codegen_wait(300)
        for rep in range(3):
            assert eng.execute(p).rows() == w, rep
    finally:
        eng.close()


def test_repeated_execution_is_stable(fused, expected):
    plan = PL.build("t3")
    for _ in range(3):
        assert fused.execute(plan).rows() == expected[6]


def test_result_text_matches_oracle(fused, cpu_tables, tmp_path):
    import ctypes
    from sycldb_amd import _lib as L
    plan = PL.build("t2")
    want = PO.execute_plan(plan, cpu_tables).text()
    res = fused.execute(plan)
    assert res.text() == want


@pytest.mark.parametrize("idx", [3, 7])
def test_result_arrays_match_oracle(fused, cpu_tables, tmp_path, idx):
    """The full host view (flags, sums, keys at selected rows) and the .res
    text, for a small group-by table (copied whole) and a 437,500-slot one
    (compacted on the device before the copy)."""
    from sycldb_amd import _lib as L
    tmpl, params = CASES[idx]
    plan = PL.build(tmpl, **params)
    want = PO.execute_plan(plan, cpu_tables)
    res = fused.execute(plan)
    assert res.nrows == want.nrows
    assert (res.flags == want.flags).all()
    sel = np.nonzero(want.flags)[0]
    assert (res.selected_rows == sel).all()
    assert len(res.columns) == len(want.columns)
    for (u64, got), (agg, ref) in zip(res.columns, want.columns):
        assert u64 == agg
        if agg:
            assert (got == ref.astype(np.uint64)).all()      # unselected slots sum to 0
        else:
            assert (got[sel] == ref[sel]).all()
    path = tmp_path / "out.res"
    L.check(L.lib.sdb_result_write(res._h, str(path).encode()))
    assert path.read_text() == want.text()


def test_errors_follow_reference(fused):
    bad_op = PL.build("t1")
    bad_op["rels"][5]["exprs"][0]["op"] = "%"     # unknown arithmetic op -> invalid_argument
    with pytest.raises(ValueError):
        fused.execute(bad_op)
    bad_join = PL.build("t1")
    bad_join["rels"][4]["condition"]["operands"][1]["input"] = 999
    with pytest.raises(ValueError):
        fused.execute(bad_join)
    # the engine keeps working after errors
    assert len(fused.execute(PL.build("t1")).rows()) == 1


def test_small_budget_raises_memory_error():
    from sycldb_amd.engine import Engine
    with Engine(device_budget_bytes=1 << 20) as e:
        e.generate(FACT_ROWS, SEED)
        with pytest.raises(MemoryError):
            e.execute(PL.build("t3"))


def test_autotuned_launches_match_oracle():
    """Launches over >= 4M rows go through the autotuner, which tries each
    candidate configuration on successive executions: every one of them must
    return the oracle's result, and a configuration is chosen at the end."""
    from sycldb_amd import engine as E
    rows = 4_500_000
    tables = datagen.generate_tables(S, rows, seed=SEED + 1)
    plan = PL.build("t3", cust=("region", 1), supp=("region", 1), date=("year_range", 1992, 1997), group="city")
    want = PO.execute_plan(plan, tables).rows()
    with E.Engine(unroll=0) as eng:
        eng.generate(rows, SEED + 1)
        for rep in range(12):
            assert eng.execute(plan).rows() == want, rep
    report = E.codegen_tuning_report()
    print(report)
    assert "unroll" in report


@pytest.mark.parametrize("span", [0xFFFE, 0xFFFF])
def test_value_table_width_boundary(cpu_tables, span):
    """Part brands spanning exactly 0xFFFE above their (negative) minimum:
    the widest 16-bit value table, whose largest offset sits next to the
    marker of unset slots; and 0xFFFF, one more, where the probe falls back
    to the reference pairs.  The selected parts hold both ends of the range."""
    from sycldb_amd.engine import Engine
    tables = {t: {c: np.array(a, dtype=np.int32, copy=True) for c, a in cols.items()} for t, cols in cpu_tables.items()}
    part = tables["part"]
    n = len(part[S.P_BRAND1])
    base = -5
    brand = (base + (np.arange(n, dtype=np.int64) * 7919) % (span + 1)).astype(np.int32)
    sel = np.flatnonzero(part[S.P_CATEGORY] == 1)
    brand[sel[0]] = base + span
    brand[sel[1]] = base
    part[S.P_BRAND1] = brand
    # fact rows that reach both ends: their part, and a supplier of region 1
    lo = tables["lineorder"]
    supp = tables["supplier"]
    skey = supp[S.S_SUPPKEY][np.flatnonzero(supp[S.S_REGION] == 1)[0]]
    rows = np.random.default_rng(11).choice(len(lo[S.LO_PARTKEY]), 100, replace=False)
    lo[S.LO_PARTKEY][rows[:50]] = part[S.P_PARTKEY][sel[0]]
    lo[S.LO_PARTKEY][rows[50:]] = part[S.P_PARTKEY][sel[1]]
    lo[S.LO_SUPPKEY][rows] = skey
    plan = PL.build("t2")
    want = PO.execute_plan(plan, tables).rows()
    assert any(r[1] == base + span for r in want) and any(r[1] == base for r in want)
    eng = Engine()
    try:
        eng.load_tables(tables)
        print("value table span", span, flush=True)
        assert eng.execute(plan).rows() == want
    finally:
        eng.close()


@pytest.mark.gpu
def test_exit_while_kernels_compile():
    """A process that exits while its plans' kernels still compile in the
    background exits cleanly (the compiler is stopped when the interpreter
    exits, before the compiler libraries are torn down), and a plan shape
    first met after the compiler stopped runs on the interpreter kernel with
    the interpreter's result."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "import torch\n"
        "from sycldb_amd import plans, _lib as L\n"
        "from sycldb_amd.engine import Engine, codegen_pending\n"
        "with Engine() as eng:\n"
        "    eng.generate(200_000, 7)\n"
        "    a = eng.execute(plans.t2_plan(('category', 3), 2)).rows()\n"
        "    busy = codegen_pending()\n"
        "    L.check(L.lib.sdb_codegen_shutdown())\n"
        "    b = eng.execute(plans.t4_plan(2, 3, (0, 1))).rows()\n"
        "with Engine(codegen='interpret') as eng:\n"
        "    eng.generate(200_000, 7)\n"
        "    want_a = eng.execute(plans.t2_plan(('category', 3), 2)).rows()\n"
        "    want_b = eng.execute(plans.t4_plan(2, 3, (0, 1))).rows()\n"
        "assert a == want_a and b == want_b and len(a) > 0 and len(b) > 0\n"
        "print('ok', busy, flush=True)\n"
        "with Engine() as eng:\n"
        "    eng.generate(200_000, 7)\n"
        "    eng.execute(plans.t3_plan()).rows()\n"
    ) % root
    print("exit while compiling: child process", flush=True)
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=240)
    assert out.returncode == 0, (out.returncode, out.stderr[-3000:])
    assert out.stdout.split()[0] == "ok", out.stdout


@pytest.mark.gpu
def test_joins_with_exact_statistics():
    """SDB_JOIN_STATS=1 (read once per process, hence a child process): joins
    read their compact tables' set-slot counts (one stream sync) instead of
    the build side's estimate and bounds; semi-join (q1.x) and full-join
    (q2.x, q3.x) plans keep the plan oracle's results."""
    import json
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    variants = [("t1", dict(date=("year", 1993))),
                ("t1", dict(date=("yearmonthnum", 199401), disc=(4, 6), qty=("between", 26, 35))),
                ("t1", dict(date=("year_range", 1994, 1995), disc=(5, 7), qty=("between", 26, 35))),
                ("t2", dict(part=("category", 3), s_region=2)),
                ("t3", dict(cust=("nation", 7), supp=("nation", 7), date=("year_range", 1992, 1997)))]
    code = (
        "import sys, json; sys.path.insert(0, %r)\n"
        "import torch\n"
        "from sycldb_amd import plans\n"
        "from sycldb_amd.engine import Engine\n"
        "out = []\n"
        "with Engine() as eng:\n"
        "    eng.generate(300_001, 11)\n"
        "    for t, kw in json.loads(sys.argv[1]):\n"
        "        kw = {k: tuple(v) if isinstance(v, list) else v for k, v in kw.items()}\n"
        "        out.append([list(map(int, r)) for r in eng.execute(plans.build(t, **kw)).rows()])\n"
        "print(json.dumps(out), flush=True)\n"
    ) % root
    env = dict(os.environ, SDB_JOIN_STATS="1")
    print("joins with exact statistics: child process", flush=True)
    out = subprocess.run([sys.executable, "-c", code, json.dumps(variants)], capture_output=True, text=True,
                         timeout=240, env=env)
    assert out.returncode == 0, out.stderr[-3000:]
    got = json.loads(out.stdout.strip().splitlines()[-1])
    tables = datagen.generate_tables(S, 300_001, 11)
    for (t, kw), rows in zip(variants, got):
        want = [list(map(int, r)) for r in PO.execute_plan(PL.build(t, **kw), tables).rows()]
        assert rows == want, (t, kw)
