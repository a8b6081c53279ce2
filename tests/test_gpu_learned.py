"""What the trained learned initialiser gains (timbir_amd.learned, weights in
timbir_amd/weights/unet_init.npz, trained by timbir_amd.train_init on the
pipeline's own reconstructions of seeded phantoms).

On an interlaced 4D scan of the deployment's model and acquisition (256 x
256 x 8 slices x 16 frames, 2048 views in 8 sub-frames, zingers and rings;
a seed the training never saw), three starts of the same ICD run to the
workload's "1e-4 relative cost change" (timbir_amd.converge): zero and the
filtered back-projection under reconstruct()'s multi-resolution schedule,
and the finest stage started from the network's output.  The learned start
must get there in fewer finest-stage iterations and less wall time than
either plain start, and end no further from the MAP point (a lower cost:
the problem is the same, so the cost orders the iterates)."""
import os
import sys

import numpy as np
import pytest

import timbir_amd as tb
from timbir_amd import converge, datagen

pytestmark = pytest.mark.gpu

WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "timbir_amd",
                       "weights", "unet_init.npz")


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.timeout(900)
def test_learned_start_converges_sooner():
    if not os.path.exists(WEIGHTS):
        pytest.skip("no trained weights")
    from timbir_amd.learned import UNet
    o = tb.get_options()
    saved = {k: getattr(o, k) for k, _ in o._fields_}
    tb.set_options(verbose=0, write_files=0)
    try:
        scan = datagen.interlaced4d(N=256, nz=8, n_frames=16, n_views=2048, seed=20269999)
        net = UNet(params=UNet.load_params(WEIGHTS))
        res = {}
        for start in ("zero", "fbp", "learned"):
            r = converge.run(scan, start, tol=1e-4, cap=150, net=net)
            rc = r["rel_change"]
            log(f"  {start:7s}: {rc and rc['iterations']} finest iterations, "
                f"{rc and rc['seconds']:.2f} s, cost {rc and rc['cost']:.8e}; "
                f"cost after finest iteration 1: {r['finest_cost_at'].get('1'):.8e}")
            assert rc is not None, start
            res[start] = rc
        for plain in ("zero", "fbp"):
            assert res["learned"]["iterations"] < res[plain]["iterations"], plain
            assert res["learned"]["seconds"] < res[plain]["seconds"], plain
            assert res["learned"]["cost"] <= res[plain]["cost"] * (1 + 1e-5), plain
    finally:
        tb.set_options(**saved)
