"""The deployment workload at its full size on the GPU: 512 x 512 x 32
slices x 16 frames, 2048 interlaced views in 8 sub-frames, 512 bins,
Poisson I0 = 1e4, zingers and ring offsets, generalized-Huber data term and
ring correction (datagen.deploy4d).  The CPU oracle cannot run it in a
test, so each of reconstruct()'s four stages is held to what does not
depend on size: after two ICD iterations (with the offset and variance
updates of that stage) the error sinogram the voxel kernels maintain equals
y - A x - d recomputed with the forward projector (XT_VoxUpdate.c:94-114,
XT_ForwardProject.c:8-25), and the forward projector conserves mass."""
import sys

import numpy as np
import pytest

import timbir_amd as tb
from timbir_amd import datagen

pytestmark = pytest.mark.gpu

# reconstruct()'s schedule for 512 columns x 32 rows (XT_Engine.c:111-141)
# and the kernel family that serves each stage
STAGES = [((8, 8), tb.K_SPLIT), ((4, 4), tb.K_MWG), ((2, 2), tb.K_MWG), ((1, 1), tb.K_MWG)]


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def scan():
    log("generating the deployment scan (512 x 512 x 32 x 16 frames, 2048 views)")
    return datagen.deploy4d()


@pytest.mark.timeout(900)
@pytest.mark.parametrize("k", range(4))
def test_deployment_stage_keeps_the_error_sinogram(scan, k):
    (mxy, mz), family = STAGES[k]
    prev_mz = STAGES[k - 1][0][1] if k else 0
    log(f"deployment stage {k} ({mxy},{mz})")
    st = tb.Stage(scan, mult_xy=mxy, mult_z=mz, stage_idx=k, prev_mult_z=prev_mz)
    i = st.info
    if k:
        # a start with structure: the generator's truth, block-averaged to
        # the stage's grid (the object a coarser stage would hand over)
        t = scan["truth"].astype(np.float64)
        t = t.reshape(i["N_time"], i["N_z"], mz, i["N_y"], mxy, i["N_x"], mxy).mean(axis=(2, 4, 6))
        st.set_object(t.astype(np.float32))
    st.init_error_sino()
    st.reset_stats()
    c0 = st.cost()[0]
    for it in (1, 2):
        st.icd_iteration(it)
    c2 = st.cost()[0]
    s = st.stats()
    fam = s["family_launches"]
    e = st.error_sino().astype(np.float64)
    ax = st.forward_project().astype(np.float64)
    d = st.offset().astype(np.float64)
    st.close()
    y = np.transpose(np.asarray(scan["projections"], np.float64), (0, 2, 1))
    ref = y - ax - d[None]
    del ax
    rel = float(np.sqrt(np.mean((e - ref) ** 2)) / np.sqrt(np.mean(ref ** 2)))
    emax = float(np.abs(e - ref).max())
    ymax = float(np.abs(y).max())
    log(f"  N_z={i['N_z']} updates {s['updates']} in {s['voxel_ms']:.0f} ms "
        f"({s['updates'] / max(s['voxel_ms'], 1e-9) / 1e3:.1f} M/s), families {fam}, "
        f"cost {c0:.6e} -> {c2:.6e}, var_est {i['var_est']:.3f}; e vs y - Ax - d: rms "
        f"{rel:.2e} of rms, max {emax:.2e} (|y| max {ymax:.2f})")
    assert fam[family] == s["voxel_launches"] > 0, fam
    assert s["updates"] > 0
    assert c2 < c0
    assert rel < 1e-4 and emax < 1e-3 * ymax


@pytest.mark.timeout(600)
def test_deployment_forward_projection_conserves_mass(scan):
    """Every view's sum over bins of A x equals the slice's mass times
    delta_xy^2 / delta_r within 1%, at the finest stage of the full
    deployment geometry (2048 views, 512 bins, 32 rows)."""
    log("deployment forward projector mass")
    st = tb.Stage(scan, mult_xy=1, mult_z=1, stage_idx=0)
    i = st.info
    n, nz = i["N_x"], i["N_z"]
    c = (np.arange(n) + 0.5) - n / 2
    disc = (c[:, None] ** 2 + c[None, :] ** 2) < (0.8 * n / 2) ** 2
    x = (np.random.default_rng(18).random((i["N_time"], nz, n, n)) * disc).astype(np.float32)
    st.set_object(x)
    ax = st.forward_project().astype(np.float64)      # [N_p][N_r][N_t]
    off, idx = st_views(scan, i)
    st.close()
    per_row = ax.sum(axis=1)                          # [N_p][N_t]
    ratio = np.empty_like(per_row)
    for f in range(i["N_time"]):
        mass = x[f].astype(np.float64).sum(axis=(1, 2)) * i["delta_xy"] ** 2 / i["delta_r"]
        ratio[idx[off[f]:off[f + 1]]] = per_row[idx[off[f]:off[f + 1]]] / mass[None, :]
    log(f"  sum over bins / mass in [{ratio.min():.5f}, {ratio.max():.5f}]")
    assert np.all(np.abs(ratio - 1.0) < 1e-2)


def st_views(scan, info):
    """Views of each frame: recon_times[f] <= t < recon_times[f + 1]
    (XT_Init.c:69-96)."""
    t = np.asarray(scan["proj_times"], np.float64)
    rt = np.asarray(scan["recon_times"], np.float64)
    idx, off = [], [0]
    for f in range(info["N_time"]):
        sel = np.nonzero((t >= rt[f]) & (t < rt[f + 1]))[0]
        idx.extend(sel.tolist())
        off.append(len(idx))
    return np.asarray(off), np.asarray(idx)
