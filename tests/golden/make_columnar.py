"""Writes tests/golden/columnar/: a tiny seeded star schema in the
reference's columnar binary format -- one raw little-endian int32 file per
loaded column, named <TABLE><column number> in upper case
(models/models.hpp:1676-1700), for the columns table_column_indices lists
(kernels/types.cpp:3-11) -- and manifest.json with the seed and sizes.

The tables come from the seeded generator (oracle/datagen.py) with small
dimensions, so the files total a few tens of KB.  Run from the repository
root:  python tests/golden/make_columnar.py
"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from oracle import datagen  # noqa: E402
from sycldb_amd import schema as S  # noqa: E402

SEED = 13
FACT_ROWS = 1500
SIZES = {"customer": 300, "supplier": 40, "part": 200, "ddate": 2556}
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "columnar")


def main():
    tables = datagen.generate_tables(S, FACT_ROWS, SEED, sizes=SIZES)
    os.makedirs(OUT, exist_ok=True)
    for t, cols in tables.items():
        for c in sorted(S.TABLE_COLUMN_INDICES[t]):
            np.ascontiguousarray(cols[c], dtype="<i4").tofile(os.path.join(OUT, "%s%d" % (t.upper(), c)))
    with open(os.path.join(OUT, "manifest.json"), "w") as f:
        json.dump({"seed": SEED, "fact_rows": FACT_ROWS, "sizes": SIZES}, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main()
