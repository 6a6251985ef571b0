"""Writes tests/golden/plans_seed7_20k.json: expected results of the plan
template cases on a small seeded star schema.

Inputs are fixed by (seed, fact_rows): the tables come from the seeded
generator (oracle/datagen.py, the CPU twin of the device generator).  Each
expected result is computed twice -- by the operator-sequence restatement of
the reference executor (oracle/plan_oracle.py) and by the independent
brute-force SQL loop (oracle/bruteforce.py) -- and written only when both
agree.  Run from the repository root:  python tests/golden/make_golden.py
"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import bruteforce, datagen, plan_oracle as PO  # noqa: E402
from sycldb_amd import plans as PL  # noqa: E402
from sycldb_amd import schema as S  # noqa: E402

SEED = 7
FACT_ROWS = 20_000
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "plans_seed7_20k.json")

CASES = [
    ("t1", {}),
    ("t1", {"date": ("yearmonthnum", 199401), "disc": (4, 6), "qty": ("range", 26, 35)}),
    ("t2", {}),
    ("t2", {"part": ("brand_range", 221, 228), "s_region": 2}),
    ("t3", {}),
    ("t3", {"cust": ("city2", 141, 145), "supp": ("region", 2), "date": ("year_range", 1992, 1997),
            "group": "city"}),
    ("t4", {}),
    ("t4", {"c_region": 3, "s_region": 0, "mfgr": (2, 4)}),
]


def main():
    tables = datagen.generate_tables(S, FACT_ROWS, seed=SEED)
    cases = []
    for tmpl, params in CASES:
        plan = PL.build(tmpl, **params)
        res = PO.execute_plan(plan, tables)
        rows = res.rows()
        brute = bruteforce.evaluate(tmpl, tables, **params)
        if rows != brute:
            raise SystemExit("oracle and brute force disagree on %s %s" % (tmpl, params))
        cases.append({"template": tmpl, "params": params, "rows": [list(r) for r in rows], "text": res.text()})
    doc = {"seed": SEED, "fact_rows": FACT_ROWS,
           "note": "expected results; inputs are the seeded tables of oracle/datagen.py",
           "cases": cases}
    with open(OUT, "w") as f:
        json.dump(doc, f, indent=1)
    print("wrote %s (%d cases, %d rows)" % (OUT, len(cases), sum(len(c["rows"]) for c in cases)))


if __name__ == "__main__":
    main()
