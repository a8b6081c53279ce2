"""Oracle results on the full 4D interlaced bench stage, for the GPU test at
that size: MAP cost after ICD iterations 1..3 and norms of the object after
iteration 3, from the CPU oracle (float arrays, colour schedule) on the same
seeded scan.  The oracle is our own restatement; this file only caches its
output so that the GPU test does not spend minutes recomputing it.

    python tests/golden/make_fullsize.py   # writes tests/golden/fullsize_interlaced4d.json
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pyorc  # noqa: E402
from timbir_amd import datagen  # noqa: E402

STAGE = dict(mult_xy=1, mult_z=1, stage_idx=2, prev_mult_z=2)


def main():
    scan = datagen.interlaced4d()
    o = pyorc.Problem(scan, opt=pyorc.options(schedule=pyorc.SCHED_COLOUR), **STAGE)
    o.init_error_sino()
    costs = []
    for it in range(1, 4):
        o.icd_iteration(it)
        costs.append(float(o.cost()[0]))
        print(it, costs[-1], flush=True)
    x = o.obj().astype(np.float64)
    out = dict(stage=STAGE, iterations=3, costs=costs, obj_sum=float(x.sum()),
               obj_l2=float(np.sqrt((x * x).sum())), obj_max=float(x.max()),
               scan_sha256=datagen.sha256(scan["projections"]))
    with open(os.path.join(HERE, "fullsize_interlaced4d.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
