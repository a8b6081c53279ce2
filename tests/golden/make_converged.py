"""The converged MAP point of the plain 3D workload (128 x 128 x 4, 256
views, quadratic data term, qGGMRF c = 1e-2) from the FP64 CPU oracle, for
the GPU test that runs reconstruct()'s stages to it.

The problem is convex (quad_convex = 1, no rings, no variance update), so
its MAP point does not depend on the visiting order.  The oracle runs the
reference's multi-resolution schedule from zero: the (2,2) stage to the
reference's stop rule, then the finest stage until an ICD iteration changes
the MAP cost by less than 1e-12 of its value (computeCost,
XT_ICD_update.c:186-319).  This file caches C* and x* (float32) with the
scan's SHA-256 so the GPU test does not spend minutes recomputing them.

    python tests/golden/make_converged.py   # writes tests/golden/converged_plain3d.npz
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pyorc  # noqa: E402
from timbir_amd import datagen  # noqa: E402

REL_STOP = 1e-12
CAP = 20000


def main():
    scan = datagen.plain3d()
    opt = pyorc.options(schedule=pyorc.SCHED_COLOUR, compute_cost=1)
    t0 = time.time()
    c0 = pyorc.Problem(scan, mult_xy=2, mult_z=2, stage_idx=0, opt=opt, precision="f64")
    _, costs0, _ = c0.stage_run(0)
    print(f"stage (2,2): {len(costs0) - 1} iterations, cost {costs0[-1]:.12e}", flush=True)
    p = pyorc.Problem(scan, mult_xy=1, mult_z=1, stage_idx=1, prev_mult_z=2,
                      var_est=c0.var_est, opt=opt, precision="f64")
    p.init_from_coarse(c0)
    p.init_error_sino()
    last = p.cost()[0]
    trace = [last]
    it = 0
    for it in range(1, CAP + 1):
        p.icd_iteration(it)
        c = p.cost()[0]
        trace.append(c)
        rel = abs(last - c) / abs(c)
        if it % 50 == 0:
            print(f"  it {it}: cost {c:.15e} rel change {rel:.3e} ({time.time() - t0:.0f} s)",
                  flush=True)
        last = c
        if rel < REL_STOP and it > 2:
            break
    x = p.obj().astype(np.float32).copy()
    print(f"finest: {it} iterations, C* = {last:.15e}", flush=True)
    np.savez_compressed(os.path.join(HERE, "converged_plain3d.npz"), cost=np.float64(last),
                        x=x, iterations=np.int32(it), rel_stop=np.float64(REL_STOP),
                        trace=np.asarray(trace, np.float64),
                        scan_sha256=np.array(datagen.sha256(scan["projections"])))


if __name__ == "__main__":
    main()
