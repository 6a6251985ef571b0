"""Committed expected results (tests/golden/plans_seed7_20k.json, written by
tests/golden/make_golden.py from the plan oracle and the brute-force SQL loop
in agreement).  The CPU tests keep the oracle and the brute-force loop pinned
to the committed vectors; the GPU test runs the same plans on the engine, on
tables from the device generator with the same seed."""
import json
import os

import pytest

from sycldb_amd import plans as PL
from sycldb_amd import schema as S

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "plans_seed7_20k.json")


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def tables(golden):
    from oracle import datagen
    return datagen.generate_tables(S, golden["fact_rows"], seed=golden["seed"])


def _rows(case):
    return [tuple(r) for r in case["rows"]]


def test_fixture_is_nontrivial(golden):
    assert len(golden["cases"]) >= 8
    assert all(len(c["rows"]) > 0 for c in golden["cases"])


def test_plan_oracle_matches_golden(golden, tables):
    from oracle import plan_oracle as PO
    for case in golden["cases"]:
        res = PO.execute_plan(PL.build(case["template"], **case["params"]), tables)
        assert res.rows() == _rows(case), case["template"]
        assert res.text() == case["text"]


def test_bruteforce_matches_golden(golden, tables):
    from oracle import bruteforce
    for case in golden["cases"]:
        assert bruteforce.evaluate(case["template"], tables, **case["params"]) == _rows(case)


@pytest.mark.gpu
def test_engine_matches_golden(golden):
    from sycldb_amd.engine import Engine
    with Engine() as eng:
        eng.generate(golden["fact_rows"], golden["seed"])
        for case in golden["cases"]:
            print("golden", case["template"], case["params"], flush=True)
            res = eng.execute(PL.build(case["template"], **case["params"]))
            assert res.rows() == _rows(case)
            assert res.text() == case["text"]
