"""Filtered back-projection, the optional initial object (tb_init_fbp).

The reference starts every reconstruction from zero (OBJECT_INIT_VAL,
src/MBIR_4D/XT_Constants.h:61), so FBP has no reference counterpart: the numpy
statement in tests/mathref.py is pinned by mathematics (a disk and an
off-centre ellipse from their closed-form projections come back at their
value, in their place), and the GPU kernels are held to it on stage data."""
import os
import sys

import numpy as np
import pytest

from timbir_amd import datagen

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import mathref  # noqa: E402


# --------------------------------------------------------------------------
# the numpy statement (CPU)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("ell", [[0.0, 0.0, 20.0, 20.0, 0.0, 1.0],
                                 [6.0, -9.0, 14.0, 7.0, 0.5, 0.8]],
                         ids=["centred_disk", "offset_ellipse"])
def test_fbp_recovers_an_ellipse(ell):
    # 360 exact views of one ellipse, box-averaged bins: inside the ellipse
    # (shrunk by 20 %) FBP returns its value within 0.1 %, and well outside
    # it (grown by 30 %) stays below 10 % of it.  A wrong filter scale, a
    # mirrored axis or an angle sign error fails one of these.
    N = 64
    ell = np.array([ell])
    ang = np.arange(360) * np.pi / 360
    p = datagen.project_ellipses(ell, ang, N, N / 2, profile="box")[:, None, :]
    x = mathref.fbp(p, ang, [np.arange(360)], N, N, -N / 2, -N / 2, 1.0, -N / 2, 1.0)[0, 0]
    grow = np.array([1, 1, 0.8, 0.8, 1, 1])
    inside = datagen.rasterize_ellipses(ell * grow, N, N / 2, sub=1) > 0
    outside = datagen.rasterize_ellipses(ell * np.array([1, 1, 1.3, 1.3, 1, 1]), N, N / 2,
                                         sub=1) == 0
    v = ell[0, 5]
    assert abs(x[inside].mean() - v) < 1e-3 * v
    assert x[inside].std() < 5e-3 * v
    assert np.abs(x[outside]).max() < 0.1 * v


def test_ramp_filter_removes_the_mean_of_a_long_row():
    # the Ram-Lak kernel sums to zero over all offsets: far from the ends of
    # a constant row its output is ~0 (the tail beyond the row decays as
    # 1/n), and it is linear
    R = mathref.ramp_matrix(512, 1.0)
    q = R @ np.ones(512)
    assert np.abs(q[200:312]).max() < 2e-3
    a, b = np.random.default_rng(0).normal(size=(2, 512))
    np.testing.assert_allclose(R @ (2 * a + b), 2 * (R @ a) + R @ b, atol=1e-12)


# --------------------------------------------------------------------------
# the GPU kernels against it
# --------------------------------------------------------------------------

def stage_fbp_reference(scan, st, off):
    """mathref.fbp on what the stage holds: raw projections minus the offset,
    averaged over each slice's mult_z rows, each frame from its own views."""
    from oracle import pyorc
    i = st.info
    mz = i["N_t"] // i["N_z"]
    y = np.asarray(scan["projections"], np.float64) - off.T[None].astype(np.float64)
    y = y.reshape(i["N_p"], i["N_z"], mz, i["N_r"]).mean(axis=2)
    o = pyorc.Problem(scan, mult_xy=i["mult_xy"], mult_z=i["mult_z"], stage_idx=0)
    voff, vidx = o.views()
    views = [vidx[voff[f]:voff[f + 1]] for f in range(i["N_time"])]
    return mathref.fbp(y, np.asarray(scan["proj_angles"], np.float64), views, i["N_x"], i["N_y"],
                       i["x0"], i["y0"], i["delta_xy"], i["R0"], i["delta_r"], o.mask())


@pytest.mark.gpu
@pytest.mark.parametrize("mxy,mz,interlace", [(1, 1, 0), (2, 2, 0), (1, 1, 8)])
def test_fbp_matches_numpy(mxy, mz, interlace):
    """tb_init_fbp on a 2-frame scan with ring offsets set: the object equals
    the FP64 numpy FBP within 1e-5 of its largest value."""
    import timbir_amd as tb
    print(f"fbp mxy={mxy} mz={mz} interlace={interlace}", flush=True)
    scan = datagen.tiny(N=32, nz=4, n_frames=2, n_views=48, rings=0.01, interlace_K=interlace)
    st = tb.Stage(scan, mult_xy=mxy, mult_z=mz, stage_idx=0)
    off = np.random.default_rng(3).normal(0, 0.01, (st.info["N_r"], st.info["N_t"]))
    st.set_offset(off.astype(np.float32))
    st.init_fbp()
    got = st.object()
    ref = stage_fbp_reference(scan, st, off.astype(np.float32))
    st.close()
    scale = np.abs(ref).max()
    assert scale > 0
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-5 * scale)
