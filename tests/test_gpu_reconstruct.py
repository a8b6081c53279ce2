"""The drop-in entry point reconstruct() (src/include/mbir4d.h:34) and the
bin_data command-line driver on the GPU: the whole multi-resolution run
against the oracle's on each workload's data model, the files it writes,
restart from status.bin, and the driver against the library it wraps."""
import functools
import os
import subprocess
import sys

import numpy as np
import pytest

import timbir_amd as tb
from oracle import pyorc
from timbir_amd import bindata, datagen

pytestmark = pytest.mark.gpu

# the library's defaults (tb_get_options before any change)
DEFAULTS = dict(compute_cost=0, abort_on_cost_increase=0, max_iterations=1000, no_nhicd=0,
                write_files=1, order_seed=0, verbose=1, kernel=0, views_per_lane=0,
                init_fbp=0)


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def scan4d(**kw):
    # c = 1 keeps the prior smooth, so the GPU and oracle trajectories agree
    # to FP32 over whole stages (see test_gpu_parity.test_phases_track_oracle)
    base = dict(N=32, nz=4, n_frames=2, n_views=40, sig_s=0.005, sig_t=0.01, c_s=1.0, c_t=1.0,
                convg_thresh=0.0, rings=0.01, remove_rings=1, quad_convex=0)
    base.update(kw)
    return datagen.tiny(**base)


@pytest.fixture
def options():
    o = tb.get_options()
    saved = {k: getattr(o, k) for k, _ in o._fields_}
    tb.set_options(**DEFAULTS)
    yield
    tb.set_options(**saved)


def n_stages(scan):
    # XT_Engine.c:111-141 for these sizes: 32 columns give the (2,2), (1,1) schedule
    return 2


def test_reconstruct_follows_oracle(tmp_path, monkeypatch, options):
    """Two-stage reconstruct() with ring correction, the zinger model and the
    variance update: the final object against the oracle's whole
    reconstruct() on the same schedule, 4 iterations per stage, NHICD off.
    (NHICD picks lines by an order statistic of the update magnitudes; FP32
    rounding of those magnitudes can move a near-tie across the threshold,
    which changes the visiting order and the iterates by ~1e-3 while both
    runs head for the same MAP point: test_reconstruct_reaches_oracle_map.)
    The generalized-Huber weight uses delta = 1 here: with the workloads'
    delta = 0.5 the weight drops from w to delta w as |e sqrt(w)| crosses T,
    so an entry within ~1e-6 of T (this scan has one in stage 0's second
    iteration) flips on FP32-rounding differences in e and moves the
    iterates by ~1e-4; at delta = 1 the weight is continuous at T.  The
    files reconstruct() writes hold the object it returns."""
    monkeypatch.chdir(tmp_path)
    log("reconstruct vs oracle")
    scan = scan4d(huber_delta=1.0)
    tb.set_options(max_iterations=4, verbose=0, no_nhicd=1)
    x = tb.reconstruct(scan)
    xo, _, _ = pyorc.reconstruct(scan, pyorc.options(schedule=pyorc.SCHED_COLOUR,
                                                     max_iterations=4, no_nhicd=1))
    rmse = float(np.sqrt(np.mean((x - xo) ** 2)) / (xo.max() - xo.min()))
    log(f"  rmse/range {rmse:.2e}")
    assert x.shape == (scan["recon_num"], scan["proj_rows"], scan["proj_cols"],
                       scan["proj_cols"])
    assert rmse < 1e-5
    for f in range(scan["recon_num"]):
        got = np.fromfile(f"object_time_{f}.bin", np.float32).reshape(x.shape[1:])
        np.testing.assert_array_equal(got, x[f])
    assert np.fromfile("status.bin", np.int32)[0] == n_stages(scan) - 1
    assert os.path.getsize("variance_estimate.bin") == 8  # a double (Appendix quirk 13)
    assert os.path.getsize("proj_offset.bin") == 4 * scan["proj_rows"] * scan["proj_cols"]
    assert os.path.getsize("mag_update.bin") == 4 * x.size


def test_reconstruct_equals_stage_pipeline(tmp_path, monkeypatch, options):
    """With NHICD on, reconstruct() equals, bit for bit, the stage API driven
    through the same schedule: stage runs, the multires object / map handoff,
    offsets and variance carried over (XT_Engine.c:156-182).  Each stage's
    NHICD path is held to the oracle by the stage-level tests."""
    monkeypatch.chdir(tmp_path)
    log("reconstruct vs the stage pipeline")
    scan = scan4d()
    tb.set_options(max_iterations=5, verbose=0)
    x = tb.reconstruct(scan)
    g0 = tb.Stage(scan, mult_xy=2, mult_z=2, stage_idx=0)
    g0.stage_run(5, with_cost=False)
    g1 = tb.Stage(scan, mult_xy=1, mult_z=1, stage_idx=1, prev_mult_z=2, var_est=g0.var_est)
    g1.init_from_coarse(g0.object(), g0.info["N_z"], g0.magmap())
    g1.set_offset(g0.offset())
    g0.close()
    g1.stage_run(5, with_cost=False)
    np.testing.assert_array_equal(x, g1.object())
    assert float(np.fromfile("variance_estimate.bin", np.float64)[0]) == g1.var_est
    np.testing.assert_array_equal(np.fromfile("proj_offset.bin", np.float32),
                                  g1.offset().reshape(-1))
    g1.close()


def test_restart_reruns_the_recorded_stage_from_its_files(tmp_path, monkeypatch, options):
    """restart = 1 starts the stage loop again at the stage status.bin
    records (XT_Engine.c:143-157), from that stage's own files as they are
    (initObject's read without interpolation, XT_ICD_update.c:558-571), then
    runs the later stages as usual: bit for bit the stage pipeline doing the
    same.  A second restart (status.bin now 1) refines the last stage from
    its files."""
    log("restart")
    monkeypatch.chdir(tmp_path)
    scan = scan4d()
    tb.set_options(max_iterations=3, verbose=0)
    # the files an interrupted run leaves after stage 0, from the stage API
    st = tb.Stage(scan, mult_xy=2, mult_z=2, stage_idx=0)
    st.stage_run(3)
    obj0, off0, var0 = st.object(), st.offset(), st.var_est
    for f in range(scan["recon_num"]):
        obj0[f].tofile(f"object_time_{f}.bin")
    st.magmap().tofile("mag_update.bin")
    off0.tofile("proj_offset.bin")
    np.array([var0], np.float64).tofile("variance_estimate.bin")
    np.array([0], np.int32).tofile("status.bin")
    st.close()
    x_restart = tb.reconstruct(scan, restart=1)
    assert np.fromfile("status.bin", np.int32)[0] == 1
    # the same through the stage API
    g0 = tb.Stage(scan, mult_xy=2, mult_z=2, stage_idx=0, var_est=var0)
    g0.set_object(obj0)
    g0.set_offset(off0)
    g0.stage_run(3, with_cost=False)
    g1 = tb.Stage(scan, mult_xy=1, mult_z=1, stage_idx=1, prev_mult_z=2, var_est=g0.var_est)
    g1.init_from_stage(g0)
    g1.set_offset(g0.offset())
    g0.close()
    g1.stage_run(3, with_cost=False)
    np.testing.assert_array_equal(x_restart, g1.object())
    obj1, mag1, off1, var1 = g1.object(), g1.magmap(), g1.offset(), g1.var_est
    g1.close()
    # every stage done: the last one again, from its files
    x_again = tb.reconstruct(scan, restart=1)
    g2 = tb.Stage(scan, mult_xy=1, mult_z=1, stage_idx=1, prev_mult_z=2, var_est=var1)
    g2.set_object(obj1)
    g2.set_magmap(mag1)
    g2.set_offset(off1)
    g2.stage_run(3, with_cost=False)
    np.testing.assert_array_equal(x_again, g2.object())
    g2.close()


def test_bin_data_driver_matches_library(tmp_path, monkeypatch, options):
    """The driver reads the bin_data files and calls reconstruct() with the
    library's default options: its object_time_*.bin equal those of a
    reconstruct() call from Python on the same arrays."""
    log("bin_data driver")
    # float32-exact parameters, so the command line carries them unchanged
    scan = scan4d(convg_thresh=2.0, sig_s=0.005, sig_t=0.01)
    for k in ("vox_wid", "rot_center", "sig_s", "sig_t", "c_s", "c_t", "convg_thresh",
              "huber_delta", "huber_T"):
        scan[k] = float(np.float32(scan[k]))
    d1 = tmp_path / "driver"
    d2 = tmp_path / "library"
    d1.mkdir()
    d2.mkdir()
    bindata.write(scan, d1)
    r = subprocess.run(bindata.command_line(scan), cwd=d1, capture_output=True, text=True,
                       timeout=240)
    log(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout + r.stderr
    monkeypatch.chdir(d2)
    tb.set_options(verbose=0)
    x = tb.reconstruct(scan)
    for f in range(scan["recon_num"]):
        a = np.fromfile(d1 / f"object_time_{f}.bin", np.float32)
        b = np.fromfile(d2 / f"object_time_{f}.bin", np.float32)
        np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(b.reshape(x.shape[1:]), x[f])
    for name in ("proj_offset.bin", "variance_estimate.bin", "status.bin"):
        assert (d1 / name).read_bytes() == (d2 / name).read_bytes(), name


def follows_oracle(scan, iters=4):
    """reconstruct() against the oracle's reconstruct() on the same schedule,
    iters iterations per stage, NHICD off.  Returns rmse / range of the final
    object against the FP32 oracle, against the FP64 oracle, and between the
    two oracles (how far FP32 rounding alone moves this trajectory)."""
    tb.set_options(max_iterations=iters, verbose=0, no_nhicd=1)
    x = tb.reconstruct(scan)
    opt = dict(schedule=pyorc.SCHED_COLOUR, max_iterations=iters, no_nhicd=1)
    x32, _, _ = pyorc.reconstruct(scan, pyorc.options(**opt))
    x64, _, _ = pyorc.reconstruct(scan, pyorc.options(**opt), precision="f64")
    assert x.shape == x32.shape

    def rel(a, b):
        return float(np.sqrt(np.mean((a - b) ** 2)) / (b.max() - b.min()))
    out = rel(x, x32), rel(x, x64), rel(x32, x64)
    log("  rmse/range vs FP32 oracle %.2e, vs FP64 oracle %.2e, FP32 vs FP64 oracle %.2e" % out)
    return out


def check_workload(scan, iters=4):
    """Two runs of one workload's data.  With c_s = c_t = 1 the prior is
    smooth and FP32 rounding moves the trajectory by ~1e-7: the object must
    equal the FP32 oracle's to 1e-5 of its range.  With the workload's own
    prior (c = 1e-2 makes the qGGMRF curvature near zero differences large,
    so rounding differences grow over the sweeps; the FP32 and FP64 oracles
    part by ~1e-3 on these sizes) the object must be as close to the FP64
    oracle as the FP32 oracle is, within a factor 3 (plus 1e-5 of the
    range, the bar of the smooth run)."""
    smooth = dict(scan, c_s=1.0, c_t=1.0)
    r32, _, _ = follows_oracle(smooth, iters)
    assert r32 < 1e-5
    _, r64, ref = follows_oracle(scan, iters)
    assert r64 <= 3 * ref + 1e-5


def test_plain3d_follows_oracle(tmp_path, monkeypatch, options):
    """The plain 3D workload at its full size (128 x 128 x 4, 256 views,
    quadratic data term): two stages of 3 iterations."""
    monkeypatch.chdir(tmp_path)
    log("plain3d 128x128x4, 256 views")
    check_workload(datagen.plain3d(convg_thresh=0.0), iters=3)


@functools.lru_cache(maxsize=1)
def batch():
    # the batched workload's model at reduced sizes: uneven widths and views
    return datagen.batched(n_problems=3, sizes=(32, 48, 64), views=(24, 36, 48),
                           convg_thresh=0.0)


@pytest.mark.parametrize("which", range(3))
def test_batched_problem_follows_oracle(tmp_path, monkeypatch, options, which):
    """One problem of the batched workload (one frame, 4 rows, its own width,
    view count and I0) through reconstruct()."""
    monkeypatch.chdir(tmp_path)
    scan = batch()[which]
    log(f"batched problem {which}: N={scan['proj_cols']} views={scan['proj_num']}")
    check_workload(scan)


def test_precision_scan_follows_oracle(tmp_path, monkeypatch, options):
    """The precision-study model (microstructure plus ellipsoid sections,
    ring offsets and ring correction) at 32 x 32 x 8, 48 views."""
    monkeypatch.chdir(tmp_path)
    log("precision3d 32x32x8, 48 views")
    check_workload(datagen.precision3d(N=32, nz=8, n_views=48, convg_thresh=0.0))


def test_bin_data_driver_on_plain3d(tmp_path, monkeypatch, options):
    """The plain 3D workload at its full size (128 x 128 x 4, 256 views)
    through the bin_data files and the driver, against reconstruct() on the
    same arrays: identical objects and offsets."""
    log("bin_data driver, plain3d 128x128x4")
    scan = datagen.plain3d()
    for k in ("vox_wid", "rot_center", "sig_s", "sig_t", "c_s", "c_t", "convg_thresh",
              "huber_delta", "huber_T"):
        scan[k] = float(np.float32(scan[k]))
    d1 = tmp_path / "driver"
    d2 = tmp_path / "library"
    d1.mkdir()
    d2.mkdir()
    bindata.write(scan, d1)
    r = subprocess.run(bindata.command_line(scan), cwd=d1, capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    monkeypatch.chdir(d2)
    tb.set_options(verbose=0)
    x = tb.reconstruct(scan)
    out = bindata.read_outputs(d1, scan["proj_rows"], scan["proj_cols"], scan["recon_num"])
    np.testing.assert_array_equal(out["object"], x)
    np.testing.assert_array_equal(
        out["proj_offset"],
        bindata.read_outputs(d2, scan["proj_rows"], scan["proj_cols"], 1)["proj_offset"])
    assert out["status"] == 1


def test_cost_file_holds_the_last_stage(tmp_path, monkeypatch, options):
    """With the cost enabled (the reference built without -DNO_COST_CALCULATE)
    cost.bin holds the last stage's costs only: the initial cost, then one per
    iteration (XT_ICD_update.c:1168, :1205), equal to the stage pipeline's."""
    monkeypatch.chdir(tmp_path)
    log("cost.bin")
    scan = scan4d()
    tb.set_options(max_iterations=3, verbose=0, compute_cost=1)
    tb.reconstruct(scan)
    costs = np.fromfile("cost.bin", np.float64)
    assert len(costs) == 4
    tb.set_options(compute_cost=0)
    g0 = tb.Stage(scan, mult_xy=2, mult_z=2, stage_idx=0)
    g0.stage_run(3, with_cost=False)
    g1 = tb.Stage(scan, mult_xy=1, mult_z=1, stage_idx=1, prev_mult_z=2, var_est=g0.var_est)
    g1.init_from_stage(g0)
    g1.set_offset(g0.offset())
    g0.close()
    c1 = g1.stage_run(3, with_cost=True)
    g1.close()
    np.testing.assert_array_equal(costs, c1)
    assert costs[-1] < costs[0]


def oracle_pipeline(scan, sched, iters):
    """The oracle's stages in reconstruct()'s order (XT_Engine.c:156-182),
    colour schedule, NHICD off: each stage from the previous one's object,
    map, offsets and variance, run to the reference's stop rule or iters.
    Returns the per-stage ratio trajectories and the finest stage."""
    ratios, prev, var = [], None, 1.0
    for s, (mxy, mz) in enumerate(sched):
        p = pyorc.Problem(scan, mult_xy=mxy, mult_z=mz, stage_idx=s,
                          prev_mult_z=sched[s - 1][1] if s else 0, var_est=var,
                          opt=pyorc.options(schedule=pyorc.SCHED_COLOUR, no_nhicd=1))
        if prev is not None:
            p.init_from_coarse(prev)
            p.off()[...] = prev.off()
        _, _, r = p.stage_run(iters)
        ratios.append(r)
        var, prev = p.var_est, p
    return ratios, prev


def test_stop_rule_fires_at_the_oracle_iteration(tmp_path, monkeypatch, options):
    """reconstruct() ends every stage where the oracle does: the reference's
    rule stops a stage after the first iteration past the first whose ratio
    100 sum|dx| / sum|x| (last chunk) falls below convg_thresh
    (XT_ICD_update.c:1096-1104, ICD_BackProject :1188-1195).  The threshold
    is picked from the oracle's own trajectories so that no ratio of any
    stage lies within 2% of it; the GPU's ratios agree with the oracle's to
    about 1e-6, so the stop iterations must be equal.  Smooth prior, NHICD
    off, ring correction and the zinger model on (delta = 1, whose weight is
    continuous at the threshold: see test_reconstruct_follows_oracle)."""
    import re
    monkeypatch.chdir(tmp_path)
    sched = [(2, 2), (1, 1)]
    max_it = 40
    base = scan4d(huber_delta=1.0)
    chosen = None
    for thresh in np.geomspace(40.0, 0.5, 25):
        ratios, _ = oracle_pipeline(dict(base, convg_thresh=float(thresh)), sched, max_it)
        stops = [len(r) for r in ratios]
        near = min(np.min(np.abs(np.asarray(r) / thresh - 1)) for r in ratios)
        if all(2 < n < max_it for n in stops) and near > 0.02:
            chosen = (float(thresh), stops, ratios)
            break
    assert chosen, "no threshold with a clear margin"
    thresh, stops, ratios = chosen
    log(f"convg_thresh {thresh:.4f}: oracle stops after {stops} iterations, ratios "
        f"{[np.round(r, 4).tolist() for r in ratios]}")
    scan = dict(base, convg_thresh=thresh)
    tb.set_options(max_iterations=max_it, verbose=1, no_nhicd=1)
    dbg = tmp_path / "debug.log"
    with open(dbg, "w") as fh:
        x = tb.reconstruct(scan, debug=fh)
    text = dbg.read_text()
    got = [int(m.group(2)) for m in
           re.finditer(r"Completed multi-resolution stage (\d+) after (\d+) iterations", text)]
    log(f"  GPU stops after {got} iterations")
    assert got == stops
    _, fine = oracle_pipeline(scan, sched, max_it)
    xo = fine.obj()
    rmse = float(np.sqrt(np.mean((x - xo) ** 2)) / (xo.max() - xo.min()))
    log(f"  rmse/range vs the oracle {rmse:.2e}")
    assert rmse < 1e-5


def test_converges_to_the_fp64_map_point_at_full_size(options):
    """The plain 3D workload at its full size (128 x 128 x 4, 256 views,
    c = 1e-2): reconstruct()'s stages on the GPU from zero, the finest stage
    run on until its MAP cost is within 1e-5 of C*, the FP64 oracle's
    converged cost (relative change below 1e-12, cached by
    tests/golden/make_converged.py); the object must then be within 0.5% of
    the range of x*, the oracle's converged image (RMSE)."""
    import os
    gold = np.load(os.path.join(os.path.dirname(__file__), "golden", "converged_plain3d.npz"))
    scan = datagen.plain3d()
    assert datagen.sha256(scan["projections"]) == str(gold["scan_sha256"]), "generator changed"
    cstar, xstar = float(gold["cost"]), gold["x"]
    tb.set_options(verbose=0)
    g0 = tb.Stage(scan, mult_xy=2, mult_z=2, stage_idx=0)
    n0 = len(g0.stage_run(0, with_cost=False)) - 1
    g1 = tb.Stage(scan, mult_xy=1, mult_z=1, stage_idx=1, prev_mult_z=2, var_est=g0.var_est)
    g1.init_from_stage(g0)
    g1.set_offset(g0.offset())
    g0.close()
    g1.init_error_sino()
    gap = None
    for it in range(1, 1001):
        g1.icd_iteration(it)
        c = g1.cost()[0]
        gap = (c - cstar) / abs(cstar)
        if abs(gap) <= 1e-5:
            break
    x = g1.object()[0]
    rmse = float(np.sqrt(np.mean((x - xstar[0]) ** 2)) / (xstar.max() - xstar.min()))
    log(f"  (2,2): {n0} iterations; finest: {it} iterations to |C - C*| = {abs(gap):.2e} |C*|, "
        f"rmse/range {rmse:.2e} (oracle: {int(gold['iterations'])} iterations to 1e-12)")
    assert abs(gap) <= 1e-5
    assert rmse <= 5e-3
    g1.close()
