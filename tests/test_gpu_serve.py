"""Serving loop (sdb_server_*): requests over three engines with resident
tables of different sizes.  Every served request's result must equal the
plan oracle's result for that plan on the same seeded tables, and the same
plan run alone on its engine: with continuous batching (Poisson arrivals), a
burst of requests that queue up behind the first ones (which must have been
batched together), and the serial baseline one request at a time."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import datagen, plan_oracle as PO  # noqa: E402
from sycldb_amd import plans as PL  # noqa: E402
from sycldb_amd import schema as S  # noqa: E402

SIZES = (100_003, 200_003, 300_007)


@pytest.fixture(scope="module")
def engines():
    from sycldb_amd.engine import Engine
    es = []
    for i, n in enumerate(SIZES):
        e = Engine()
        e.generate(n, 50 + i)
        es.append(e)
    yield es
    for e in es:
        e.close()


@pytest.fixture(scope="module")
def trace():
    rng = np.random.default_rng(5)
    pipes = PL.random_pipelines(48, 21)
    reqs = [(int(rng.integers(0, len(SIZES))), PL.build(t, **p)) for t, p in pipes]
    arrivals = np.cumsum(rng.exponential(0.0002, len(reqs)))  # ~5000 requests/s: arrivals overlap running batches
    return reqs, arrivals


@pytest.fixture(scope="module")
def oracle_rows(trace):
    """Each request's rows from the plan oracle on its engine's tables."""
    reqs, _ = trace
    tables = [datagen.generate_tables(S, n, seed=50 + i) for i, n in enumerate(SIZES)]
    return [PO.execute_plan(plan, tables[e]).rows() for e, plan in reqs]


@pytest.fixture(scope="module")
def alone(engines, trace, oracle_rows):
    """Each request run alone on its engine; equal to the oracle's rows."""
    reqs, _ = trace
    rows = [engines[e].execute(plan).rows() for e, plan in reqs]
    for i, (r, w) in enumerate(zip(rows, oracle_rows)):
        assert r == w, i
    return rows


def test_batched_serving_matches_oracle(engines, trace, alone, oracle_rows):
    from sycldb_amd.engine import Server, replay
    reqs, arrivals = trace
    print("serving", len(reqs), "requests", flush=True)
    with Server(engines, max_batch=16) as srv:
        out = replay(srv, reqs, arrivals)
        stats = srv.stats()
    for i, (res, info) in enumerate(out):
        assert res.rows() == oracle_rows[i] == alone[i], i
        assert info["latency_ms"] >= info["run_ms"] >= 0
    assert stats["requests"] == len(reqs)


def test_burst_is_batched(engines, trace, alone, oracle_rows):
    """48 requests at once to one worker for all engines: ~16 per engine
    queue behind the first ones, more than the initial batching threshold, so
    they share scans.  (With a worker per engine these small tables drain
    as fast as the requests are submitted, and nothing needs batching.)"""
    from sycldb_amd.engine import Server, replay
    reqs, arrivals = trace
    with Server(engines, max_batch=16, single_worker=True) as srv:
        out = replay(srv, reqs, arrivals * 0)
        stats = srv.stats()
    for i, (res, info) in enumerate(out):
        assert res.rows() == oracle_rows[i], i
    assert stats["requests"] == len(reqs)
    assert max(info["batch_size"] for _, info in out) > 1
    assert stats["batches"] < len(reqs)


def test_serial_serving_matches_oracle(engines, trace, alone, oracle_rows):
    from sycldb_amd.engine import Server, replay
    reqs, arrivals = trace
    with Server(engines, serial=True) as srv:
        out = replay(srv, reqs[:16], arrivals[:16])
    for i, (res, info) in enumerate(out):
        assert res.rows() == oracle_rows[i]
        assert info["batch_size"] == 1


def test_tight_memory_budget_still_serves(engines, trace, alone, oracle_rows):
    """A budget below one request's footprint admits one request per batch."""
    from sycldb_amd.engine import Server, replay
    reqs, arrivals = trace
    with Server(engines, max_batch=16, memory_budget_bytes=1) as srv:
        out = replay(srv, reqs[:12], arrivals[:12] * 0)
    for i, (res, info) in enumerate(out):
        assert res.rows() == oracle_rows[i]
        assert info["batch_size"] == 1


def test_bad_request_fails_alone(engines):
    """An invalid plan is rejected at submission; the server keeps serving."""
    from sycldb_amd.engine import Server
    with Server(engines) as srv:
        bad = PL.build("t1")
        bad["rels"][5]["exprs"][0]["op"] = "%"
        t_bad = srv.submit(0, bad)
        t_ok = srv.submit(0, PL.build("t1"))
        with pytest.raises(ValueError):
            srv.wait(t_bad)
        res, _ = srv.wait(t_ok)
        assert res.rows() == engines[0].execute(PL.build("t1")).rows()


def test_single_worker_serving_matches_oracle(engines, trace, alone, oracle_rows):
    """The one-worker form (no overlap between engines) gives the same
    results."""
    from sycldb_amd.engine import Server, replay
    reqs, arrivals = trace
    with Server(engines, max_batch=16, single_worker=True) as srv:
        out = replay(srv, reqs, arrivals)
    for i, (res, info) in enumerate(out):
        assert res.rows() == oracle_rows[i], i


def test_burst_with_a_worker_per_engine(engines, trace, alone, oracle_rows):
    """The same burst with a worker per engine: every result the oracle's."""
    from sycldb_amd.engine import Server, replay
    reqs, arrivals = trace
    with Server(engines, max_batch=16) as srv:
        out = replay(srv, reqs, arrivals * 0)
        stats = srv.stats()
    for i, (res, info) in enumerate(out):
        assert res.rows() == oracle_rows[i], i
    assert stats["requests"] == len(reqs)
