"""The columnar binary format (models/models.hpp:1659-1708): the committed
fixtures under tests/golden/columnar/ (written by make_columnar.py) hold
the seeded generator's tables, and sdb_engine_table_load_dir loads them --
including the reference's handling of an unreadable file (0 entries) and of
a column whose length differs from the first column's (left unloaded) --
with plan results equal to the oracle's on the same files."""
import json
import os
import shutil

import numpy as np
import pytest

from oracle import bruteforce, datagen, plan_oracle as PO
from sycldb_amd import plans as PL
from sycldb_amd import schema as S

DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "columnar")
TABLES = ("ddate", "customer", "supplier", "part", "lineorder")
PLANS = [("t1", {}), ("t2", {"part": ("category", 3), "s_region": 2}), ("t3", {"group": "city"}),
         ("t4", {"c_region": 1, "s_region": 3, "mfgr": (0, 4)})]


def read_dir(d):
    out = {}
    for t in TABLES:
        out[t] = {}
        for c in sorted(S.TABLE_COLUMN_INDICES[t]):
            path = os.path.join(d, "%s%d" % (t.upper(), c))
            out[t][c] = np.fromfile(path, dtype="<i4").astype(np.int32)
    return out


def test_fixtures_are_the_seeded_tables():
    m = json.load(open(os.path.join(DIR, "manifest.json")))
    want = datagen.generate_tables(S, m["fact_rows"], m["seed"], sizes=m["sizes"])
    got = read_dir(DIR)
    for t in TABLES:
        assert sorted(got[t]) == sorted(S.TABLE_COLUMN_INDICES[t])
        for c, v in got[t].items():
            assert (v == want[t][c]).all(), (t, c)


@pytest.mark.parametrize("tmpl,params", PLANS)
def test_oracle_on_fixtures_matches_bruteforce(tmpl, params):
    tables = read_dir(DIR)
    plan = PL.build(tmpl, **params)
    rows = PO.execute_plan(plan, tables).rows()
    assert rows == bruteforce.evaluate(tmpl, tables, **params)


@pytest.mark.gpu
def test_load_dir_runs_plans_like_the_oracle(dev):
    from sycldb_amd.engine import Engine
    tables = read_dir(DIR)
    with Engine() as eng:
        for t in TABLES:
            eng.load_dir(t, DIR)
        for t in TABLES:
            for c, v in tables[t].items():
                assert (eng.download(t, c, len(v)) == v).all(), (t, c)
        for tmpl, params in PLANS:
            plan = PL.build(tmpl, **params)
            assert eng.execute(plan).rows() == PO.execute_plan(plan, tables).rows(), tmpl


@pytest.mark.gpu
def test_load_dir_unreadable_and_short_files(dev, tmp_path, capfd):
    """A missing file is 0 entries: a missing later column stays unloaded
    (warning); a column of another length stays unloaded (warning); a
    missing first column gives the table 0 rows."""
    from sycldb_amd.engine import Engine
    d = tmp_path / "cols"
    shutil.copytree(DIR, d)
    os.remove(d / "LINEORDER13")                       # supplycost: missing
    short = np.fromfile(d / "LINEORDER9", dtype="<i4")[:-7]
    short.tofile(d / "LINEORDER9")                      # extendedprice: 7 rows short
    os.remove(d / "SUPPLIER0")                          # the supplier table's first column
    tables = read_dir(DIR)
    with Engine() as eng:
        for t in TABLES:
            eng.load_dir(t, str(d))
        err = capfd.readouterr().err
        assert "LINEORDER13" in err and "LINEORDER9" in err and "SUPPLIER3" in err
        # the loaded lineorder columns are intact and serve a plan that
        # reads only them (every template reads a column or table left
        # unloaded here, which the reference would read uninitialised)
        lo = tables["lineorder"]
        for c in (S.LO_CUSTKEY, S.LO_QUANTITY, S.LO_REVENUE):
            assert (eng.download("lineorder", c, len(lo[c])) == lo[c]).all()
        b = PL.PlanBuilder()
        b.scan("lineorder")
        b.filter(PL.expr("<", PL.col(S.LO_QUANTITY), PL.lit(25)))
        b.project([PL.col(S.LO_REVENUE)])
        b.aggregate([], 0)
        want = int(lo[S.LO_REVENUE][lo[S.LO_QUANTITY] < 25].astype(np.int64).sum()) % 2**64
        assert eng.execute(b.plan()).rows() == [(want,)]
