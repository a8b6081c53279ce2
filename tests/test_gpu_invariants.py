"""Size-independent checks of the voxel kernels at sizes the CPU oracle cannot
run in a test.

ICD keeps the error sinogram e = y - A x - d up to date voxel by voxel
(XT_VoxUpdate.c:94-114): after any number of updates, the e the kernels
maintain must equal y - A x - d recomputed from the current object with the
forward projector (XT_ForwardProject.c:8-25) and the current offsets, up to
FP32 rounding.  A lost update, a race between cells or a wrong entry shows
as an O(1) error somewhere; the test reports the worst entry."""
import sys

import numpy as np
import pytest

import timbir_amd as tb
from timbir_amd import datagen

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def check_invariant(scan, mult_xy, mult_z, iters=2):
    st = tb.Stage(scan, mult_xy=mult_xy, mult_z=mult_z, stage_idx=0)
    i = st.info
    st.init_error_sino()
    st.reset_stats()
    for it in range(1, iters + 1):
        st.icd_iteration(it)
    s = st.stats()
    e = st.error_sino().astype(np.float64)            # [N_p][N_r][N_t]
    ax = st.forward_project().astype(np.float64)
    d = st.offset().astype(np.float64)                # [N_r][N_t]
    y = np.asarray(scan["projections"], np.float64)   # [N_p][rows][cols]
    y = np.transpose(y, (0, 2, 1))                    # [N_p][N_r][N_t]
    ref = y - ax - d[None]
    st.close()
    err = np.abs(e - ref)
    rel = float(np.sqrt(np.mean((e - ref) ** 2)) / np.sqrt(np.mean(ref ** 2)))
    log(f"  stage ({mult_xy},{mult_z}) N_z={i['N_z']} N_time={i['N_time']} views/frame="
        f"{i['N_p'] // i['N_time']}: {s['updates']} updates, {s['lds_launches']} LDS / "
        f"{s['global_launches']} global launches; e vs y - Ax - d: rms {rel:.2e} of rms, "
        f"max {err.max():.2e} (|y| max {np.abs(y).max():.2f})")
    return s, rel, float(err.max()), float(np.abs(y).max())


def test_error_sinogram_stays_consistent_global_e():
    """Finest stage whose cells (128 views x 256 bins) exceed LDS: the
    global-e kernel, 256 x 256 x 16 slices x 8 frames, 1024 views."""
    log("invariant: global-e kernel")
    scan = datagen.interlaced4d(N=256, nz=16, n_frames=8, n_views=1024, ch_steps=16,
                                frame_steps=4, seed=11)
    s, rel, emax, ymax = check_invariant(scan, 1, 1)
    assert s["global_launches"] == s["voxel_launches"] > 0
    assert s["updates"] > 0
    assert rel < 1e-4 and emax < 1e-3 * ymax


def test_error_sinogram_stays_consistent_wide_columns():
    """Coarsest stage of a 512-wide scan: 17-bin columns over 8 detector rows
    (the wide-column kernel), 512 x 512 x 16 slices x 2 frames, 128 views."""
    log("invariant: wide-column kernel")
    scan = datagen.interlaced4d(N=512, nz=16, n_frames=2, n_views=128, ch_steps=8,
                                frame_steps=2, seed=12)
    s, rel, emax, ymax = check_invariant(scan, 8, 8)
    assert s["global_launches"] == s["voxel_launches"] > 0
    assert s["updates"] > 0
    assert rel < 1e-4 and emax < 1e-3 * ymax


def test_error_sinogram_stays_consistent_many_views():
    """Single-frame scan with 720 views at its (2,2) stage: more views than
    any lane layout holds, so the pass-based kernel runs, 256 x 256 x 4."""
    log("invariant: pass-based kernel")
    scan = datagen.batched(n_problems=1, sizes=(256,), views=(720,), seed=13)[0]
    s, rel, emax, ymax = check_invariant(scan, 2, 2)
    assert s["global_launches"] == s["voxel_launches"] > 0
    assert s["updates"] > 0
    assert rel < 1e-4 and emax < 1e-3 * ymax


def test_error_sinogram_stays_consistent_single_row_passes():
    """Finest stage of the precision-study model with 1440 views in one frame
    (more than any lane layout holds): single-row passes, 128 x 128 x 8."""
    log("invariant: single-row pass-based kernel")
    scan = datagen.precision3d(N=128, nz=8, n_views=1440, seed=14)
    s, rel, emax, ymax = check_invariant(scan, 1, 1)
    assert s["global_launches"] == s["voxel_launches"] > 0
    assert s["updates"] > 0
    assert rel < 1e-4 and emax < 1e-3 * ymax


@pytest.mark.parametrize("mxy,mz", [(4, 4), (8, 8)])
def test_error_sinogram_stays_consistent_grouped_passes(mxy, mz):
    """Coarse stages of the precision-study model with 1440 views in one
    frame: more passes than registers hold, so the grouped-pass kernel runs,
    256 x 256 x 16."""
    log(f"invariant: grouped-pass kernel ({mxy},{mz})")
    scan = datagen.precision3d(N=256, nz=16, n_views=1440, seed=15)
    s, rel, emax, ymax = check_invariant(scan, mxy, mz)
    assert s["global_launches"] == s["voxel_launches"] > 0
    assert s["updates"] > 0
    assert rel < 1e-4 and emax < 1e-3 * ymax


@pytest.mark.parametrize("mxy,mz", [(1, 1), (4, 4)])
def test_forward_projection_conserves_mass(mxy, mz):
    """Sum over the detector bins of every view of A x equals the object's
    mass in that slice times delta_xy^2 / delta_r, whatever the view angle
    (each A column, XT_AMatrix.c:50-110, spreads one voxel's area over the
    bins; the 3-bin cap of the finest columns and the tabulated detector
    response keep it within 1%).  Checks the device forward projector
    (XT_ForwardProject.c:8-25) at a size and view count the oracle does not
    run: 256 x 256 x 16 slices, 1440 views, object inside a disc so that no
    footprint leaves the detector."""
    log(f"forward projector mass ({mxy},{mz})")
    scan = datagen.precision3d(N=256, nz=16, n_views=1440, seed=16)
    st = tb.Stage(scan, mult_xy=mxy, mult_z=mz, stage_idx=0)
    i = st.info
    n, nz = i["N_x"], i["N_z"]
    c = (np.arange(n) + 0.5) - n / 2
    disc = (c[:, None] ** 2 + c[None, :] ** 2) < (0.8 * n / 2) ** 2
    x = (np.random.default_rng(17).random((i["N_time"], nz, n, n)) * disc).astype(np.float32)
    st.set_object(x)
    ax = st.forward_project().astype(np.float64)      # [N_p][N_r][N_t]
    st.close()
    per_row = ax.sum(axis=1)                          # [N_p][N_t]
    rows_per_slice = i["N_t"] // nz
    mass = x[0].astype(np.float64).sum(axis=(1, 2)) * i["delta_xy"] ** 2 / i["delta_r"]
    # every detector row of slice z sees the slice's mass
    want = np.repeat(mass, rows_per_slice)[None, :]
    ratio = per_row / want
    log(f"  {i['N_p']} views x {i['N_t']} rows: sum over bins / mass in "
        f"[{ratio.min():.5f}, {ratio.max():.5f}]")
    assert np.all(np.abs(ratio - 1.0) < 1e-2)
    # the spread over views is the A columns' angular dependence only
    assert ratio.max() - ratio.min() < 1.5e-2
