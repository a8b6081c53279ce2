"""Precision study on the precision-study workload's model (microstructure
plus ellipsoid sections, ring offsets and ring correction) at the largest
size the FP64 CPU oracle runs in a test: 128 x 128 x 8 slices, 360 views
over a half turn (the full workload's views per detector width), 8 ICD
iterations of the finest stage from zero, colour schedule, NHICD off.

The variants, each against the FP64 ICD oracle (the reference point):
  fp32c  GPU, theta1 and theta2 in FP32 with compensated (TwoSum)
         accumulation (tb_options.theta_fp = 1)
  fp32   GPU, theta sums in plain FP32 (theta_fp = 2)
  fp64c  GPU default: theta1 in FP64, theta2 FP32 (theta_fp = 0); in all GPU
         variants the arrays, A values, voxel update and error sinogram are
         FP32 and the cost, convergence check, offset and variance solves
         FP64
  f32    the CPU oracle with FP32 arrays and FP64 A values and sums (the
         reference's own types, Real_arr_t = float, Real_t = double)
  f64    the CPU oracle with FP64 arrays: the reference point
Each final (object, offsets) is costed by the FP64 oracle (computeCost,
src/MBIR_4D/XT_ICD_update.c:186-319), so every variant is measured with one
FP64 cost function; the object error is the RMSE to the FP64 run over its
range."""
import sys

import numpy as np
import pytest

import timbir_amd as tb
from oracle import pyorc
from timbir_amd import datagen

pytestmark = pytest.mark.gpu

ITERS = 8
N, NZ, VIEWS = 128, 8, 360


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture
def opts():
    o = tb.get_options()
    saved = {k: getattr(o, k) for k, _ in o._fields_}
    tb.set_options(no_nhicd=1, kernel=4)
    yield
    tb.set_options(**saved)


def run_oracle(scan, precision):
    o = pyorc.Problem(scan, opt=pyorc.options(schedule=pyorc.SCHED_COLOUR, no_nhicd=1),
                      precision=precision)
    o.init_error_sino()
    for it in range(1, ITERS + 1):
        o.icd_iteration(it)
    return np.array(o.obj(), np.float64), np.array(o.off(), np.float64)


def run_gpu(scan, theta_fp):
    tb.set_options(theta_fp=theta_fp)
    g = tb.Stage(scan)
    g.init_error_sino()
    for it in range(1, ITERS + 1):
        g.icd_iteration(it)
    st = g.stats()
    x, d = g.object().astype(np.float64), g.offset().astype(np.float64)
    g.close()
    assert st["family_launches"][tb.K_MWG] == st["voxel_launches"] > 0, st
    return x, d


def fp64_cost(ref, x, d):
    ref.obj()[...] = x
    ref.off()[...] = d
    ref.init_error_sino()
    return ref.cost()[0]


def test_precision_variants(opts):
    log(f"precision study: {N} x {N} x {NZ} slices, {VIEWS} views, {ITERS} iterations")
    scan = datagen.precision3d(N=N, nz=NZ, n_views=VIEWS, convg_thresh=0.0)
    runs = {"fp32c": run_gpu(scan, 1), "fp32": run_gpu(scan, 2), "fp64c": run_gpu(scan, 0)}
    runs["f32"] = run_oracle(scan, "f32")
    x64, d64 = run_oracle(scan, "f64")
    ref = pyorc.Problem(scan, opt=pyorc.options(schedule=pyorc.SCHED_COLOUR), precision="f64")
    c64 = fp64_cost(ref, x64, d64)
    span = x64.max() - x64.min()
    out = {}
    for name, (x, d) in runs.items():
        c = fp64_cost(ref, x, d)
        out[name] = (abs(c - c64) / abs(c64), float(np.sqrt(np.mean((x - x64) ** 2)) / span))
        log(f"  {name:6s}: MAP cost vs f64 {out[name][0]:.2e} relative, rmse vs f64 "
            f"{out[name][1]:.2e} of range")
    log(f"  f64 MAP cost {c64:.10e}")
    # the MAP cost bar of the path (1e-5 relative to the FP64 run), and the
    # object well inside 0.5% of the range
    for name in ("fp32c", "fp32", "fp64c"):
        assert out[name][0] < 1e-5 and out[name][1] < 2e-4, name
    # the compensated FP32 sums are no worse than the FP64-theta1 default
    # beyond FP32 rounding of the A values, which all GPU variants share
    assert out["fp32c"][1] <= 1.5 * out["fp64c"][1] + 1e-6
