"""The autotuner (tb_autotune*, tb_options.autotune) on the GPU: tuning a
stage leaves it exactly as it was, a tuned run equals a run with the chosen
configuration asked for explicitly, and the cache round-trips through a
file."""
import sys

import numpy as np
import pytest

import timbir_amd as tb
from timbir_amd import datagen

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture
def opts():
    o = tb.get_options()
    saved = {k: getattr(o, k) for k, _ in o._fields_}
    tb.autotune_clear()
    yield
    tb.set_options(**saved)
    tb.autotune_clear()


def scan_small():
    return datagen.interlaced4d(N=48, nz=4, n_frames=2, n_views=96, K=4, seed=20261111)


def test_tuning_leaves_the_stage_unchanged(opts):
    st = tb.Stage(scan_small())
    try:
        st.init_error_sino()
        st.icd_iteration(1)
        x, e, m = st.object(), st.error_sino(), st.magmap()
        s0 = st.stats()
        t = st.autotune()
        log(f"tuned {t}")
        assert t["kernel"] in (0, 2, 3, 4, 5, 6, 7, 8, 9) and t["ms"] >= 0
        np.testing.assert_array_equal(st.object(), x)
        np.testing.assert_array_equal(st.error_sino(), e)
        np.testing.assert_array_equal(st.magmap(), m)
        s1 = st.stats()
        for k in ("voxel_launches", "lds_launches", "global_launches", "family_launches",
                  "voxel_ms", "updates"):
            assert s1[k] == s0[k], k
        assert tb.autotune_entries() == [t]
        assert st.autotune() == t  # cached: no second timing
    finally:
        st.close()


def test_tuned_run_equals_the_chosen_configuration(opts):
    scan = scan_small()
    tb.set_options(autotune=1, kernel=0)
    a = tb.Stage(scan)
    a.init_error_sino()
    for it in range(1, 4):
        a.icd_iteration(it)
    (t,) = tb.autotune_entries()
    log(f"tuned {t}")
    tb.set_options(autotune=0, kernel=t["kernel"], views_per_lane=t["views_per_lane"])
    b = tb.Stage(scan)
    b.init_error_sino()
    for it in range(1, 4):
        b.icd_iteration(it)
    try:
        np.testing.assert_array_equal(a.object(), b.object())
        np.testing.assert_array_equal(a.error_sino(), b.error_sino())
        assert a.stats()["family_launches"] == b.stats()["family_launches"]
    finally:
        a.close()
        b.close()


def test_cache_round_trips_through_a_file(opts, tmp_path):
    st = tb.Stage(scan_small())
    try:
        st.init_error_sino()
        t = st.autotune()
    finally:
        st.close()
    path = str(tmp_path / "tune.json")
    tb.autotune_save(path)
    tb.autotune_clear()
    assert tb.autotune_entries() == []
    tb.autotune_load(path)
    got = tb.autotune_entries()
    assert len(got) == 1
    g = got[0]
    assert (g["shape"], g["kernel"], g["views_per_lane"]) == (t["shape"], t["kernel"],
                                                              t["views_per_lane"])
    with pytest.raises(tb.TimbirError):
        tb.autotune_put([{"shape": [1] * 8, "kernel": 11, "views_per_lane": 0}])
