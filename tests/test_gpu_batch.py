"""Many reconstructions at once (timbir_amd.batch) on the GPU: concurrent
runs on separate streams, and scans advanced in lockstep through one
tb_batch (one voxel-update launch per kernel configuration over every
scan's cells, HIP-graph replay), give the same objects, bit for bit, as
one at a time and as reconstruct() (src/include/mbir4d.h:34) itself."""
import sys

import numpy as np
import pytest

import timbir_amd as tb
from timbir_amd import batch, datagen

pytestmark = pytest.mark.gpu


@pytest.fixture
def options():
    o = tb.get_options()
    saved = {k: getattr(o, k) for k, _ in o._fields_}
    tb.set_options(max_iterations=4, verbose=0)
    yield
    tb.set_options(**saved)


def test_concurrent_runs_equal_single_runs(tmp_path, monkeypatch, options):
    print("batch of 4 uneven problems, 4 workers", flush=True)
    sys.stdout.flush()
    scans = datagen.batched(n_problems=4, sizes=(32, 48, 64), views=(24, 36, 48),
                            convg_thresh=2.0)
    one = batch.reconstruct_many(scans, workers=1)
    many = batch.reconstruct_many(scans, workers=4)
    for s, a, b in zip(scans, one, many):
        assert a.shape == (1, s["proj_rows"], s["proj_cols"], s["proj_cols"])
        np.testing.assert_array_equal(a, b)
    monkeypatch.chdir(tmp_path)
    np.testing.assert_array_equal(tb.reconstruct(scans[1]), one[1])


@pytest.mark.parametrize("graph", [True, False], ids=["graph", "eager"])
def test_lockstep_batch_equals_single_runs(options, graph):
    """Uneven scans (three sizes, three view counts, so several kernel
    configurations and schedules of 1-2 stages) in one tb_batch: every
    object equals that of the scan run alone, bit for bit, after four ICD
    iterations per stage with NHICD; the batch made fewer launches than the
    members would alone, and with graph on replayed iterations from a
    captured graph."""
    print(f"lockstep batch of 6 uneven problems, graph={graph}", flush=True)
    sys.stdout.flush()
    scans = datagen.batched(n_problems=6, sizes=(32, 48, 64), views=(24, 36, 48),
                            convg_thresh=2.0, seed=20260011)
    alone = [batch.reconstruct_stages(s) for s in scans]
    st = {}
    together = batch.reconstruct_batched(scans, graph=graph, stats=st)
    print(f"  stats {st}", flush=True)
    for a, b in zip(alone, together):
        np.testing.assert_array_equal(a, b)
    assert st["groups"] >= 1 and st["iterations"] > 0
    if graph:
        assert st["captures"] >= 1 and st["graph_replays"] > 0
    else:
        assert st["captures"] == 0 and st["graph_replays"] == 0


def test_lockstep_batch_one_member_equals_icd_iteration(options):
    """A batch of one stage runs tb_icd_iteration's sequence: same object,
    error sinogram and stop flags, iteration by iteration."""
    scan = datagen.batched(n_problems=1, sizes=(48,), views=(36,), convg_thresh=2.0,
                           seed=20260012)[0]
    a = tb.Stage(scan)
    c = tb.Stage(scan)
    a.init_error_sino()
    c.init_error_sino()
    b = tb.Batch(graph=True)
    try:
        for it in range(1, 5):
            fa, ra = a.icd_iteration(it)
            fc, rc = b.iteration([c], it)
            assert (fa, ra) == (int(fc[0]), float(rc[0])), it
            np.testing.assert_array_equal(a.object(), c.object())
            np.testing.assert_array_equal(a.error_sino(), c.error_sino())
        sa, sc = a.stats(), c.stats()
        assert sa["voxel_launches"] == sc["voxel_launches"]
        assert sa["family_launches"] == sc["family_launches"]
    finally:
        b.close()
        a.close()
        c.close()


@pytest.mark.timeout(1200)
def test_full_batched_workload(tmp_path, monkeypatch):
    """The batched workload at its own sizes: 64 single-slice problems of
    uneven size (128-512 px, 180-720 views, 4 detector rows each, I0 in
    [5e3, 5e4]), reconstructed together, 3 ICD iterations per stage, NHICD
    off, first by host threads with a stream each, then in lockstep through
    one tb_batch, which must be faster and give the same bits (both with the
    kernel families the lockstep batch launches: the split-cell kernel,
    which a problem run alone takes for its few cells, adds the theta
    partials in another order and is off here).  Problems up
    to 192 px: the object against the CPU oracle's reconstruct() on the
    same schedule: as close to the FP64 oracle as the FP32 oracle is,
    within a factor 3 plus 1e-4 of the range (the FP32 bar of the
    trajectory tests; the workload's stiff prior amplifies rounding, and
    the GPU sums theta in another order than the oracle's FP32 run).
    Larger problems: the batched object equals a run of that problem
    alone, bit for bit."""
    from oracle import pyorc
    o = tb.get_options()
    saved = {k: getattr(o, k) for k, _ in o._fields_}
    tb.set_options(max_iterations=3, verbose=0, no_nhicd=1, split_segments=1)
    try:
        scans = datagen.batched()
        print(f"batched workload: {len(scans)} problems", flush=True)
        import time
        t0 = time.time()
        def done(k):
            s = scans[k]
            print(f"  problem {k} (N={s['proj_cols']}, {s['proj_num']} views) done at "
                  f"{time.time() - t0:.1f} s", flush=True)
        out = batch.reconstruct_many(scans, workers=8, progress=done)
        t_pool = time.time() - t0
        print(f"  64 problems in {t_pool:.1f} s (host threads, a stream each)", flush=True)
        t0 = time.time()
        st = {}
        lock = batch.reconstruct_batched(scans, stats=st)
        t_lock = time.time() - t0
        print(f"  64 problems in {t_lock:.1f} s in lockstep ({st['launches']} voxel launches, "
              f"{st['graph_replays']} graph replays, stages {st['stages']})", flush=True)
        for a, b in zip(out, lock):
            np.testing.assert_array_equal(a, b)
        assert t_lock < t_pool
        opt = dict(schedule=pyorc.SCHED_COLOUR, max_iterations=3, no_nhicd=1)
        checked = 0
        for k, (s, x) in enumerate(zip(scans, out)):
            assert x.shape == (1, s["proj_rows"], s["proj_cols"], s["proj_cols"])
            assert np.all(np.isfinite(x))
            if s["proj_cols"] <= 192:
                x32, _, _ = pyorc.reconstruct(s, pyorc.options(**opt))
                x64, _, _ = pyorc.reconstruct(s, pyorc.options(**opt), precision="f64")
                rel = lambda a, b: float(np.sqrt(np.mean((a - b) ** 2)) / (b.max() - b.min()))
                r64, ref = rel(x, x64), rel(x32, x64)
                print(f"  problem {k}: N={s['proj_cols']} views={s['proj_num']}: rmse/range "
                      f"vs FP64 oracle {r64:.2e} (FP32 oracle {ref:.2e})", flush=True)
                assert r64 <= 3 * ref + 1e-4
                checked += 1
        assert checked >= 20
        for k in sorted(range(len(scans)), key=lambda q: -scans[q]["proj_cols"] * scans[q]["proj_num"])[:2]:
            alone = batch.reconstruct_stages(scans[k])
            print(f"  problem {k}: N={scans[k]['proj_cols']} views={scans[k]['proj_num']} "
                  f"alone == batched", flush=True)
            np.testing.assert_array_equal(alone, out[k])
    finally:
        tb.set_options(**saved)
