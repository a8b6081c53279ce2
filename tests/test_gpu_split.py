"""The split-cell kernel's bounded waits on the GPU (k_icd_split, kernel
option 10): a wait that runs out of polls (tb_options.split_spin_limit, set
from the host so that it runs out at once) ends its launch normally and
fails the call; the stage that failed is then discarded (its object and
error sinogram are invalid: workgroups that gave up stored other values
than those that did not) and a fresh stage runs as before, bit for bit; an
abort request (tb_stage_abort) while no split-cell launch runs changes
nothing.  The scan is a single-frame one with 4 cells per colour, the
shape the kernel serves."""
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
    tb.set_options(kernel=10, verbose=0, write_files=0)
    yield
    tb.set_options(**saved)


@pytest.fixture(scope="module")
def scan():
    return datagen.precision3d(N=128, nz=8, n_views=360)


def run(scan, iters=2, abort_between=False):
    st = tb.Stage(scan)
    try:
        if abort_between:
            st.abort()  # before any split-cell launch: nothing to end
        st.init_error_sino()
        for it in range(1, iters + 1):
            st.icd_iteration(it)
            if abort_between:
                st.synchronize()
                st.abort()  # between launches: the next launch clears it
        fam = st.stats()["family_launches"]
        return st.object(), st.error_sino(), fam
    finally:
        st.close()


def test_wait_out_of_polls_fails_the_call_and_later_runs_are_unchanged(scan, opts):
    x0, e0, fam = run(scan)
    log(f"split launches {fam[tb.K_SPLIT]}")
    assert fam[tb.K_SPLIT] > 0 and sum(fam) == fam[tb.K_SPLIT], fam
    tb.set_options(split_spin_limit=1)  # one poll per wait
    st = tb.Stage(scan)
    try:
        st.init_error_sino()
        with pytest.raises(tb.TimbirError, match="timed out"):
            st.icd_iteration(1)
        log("wait out of polls: the call failed")
    finally:
        st.close()
    tb.set_options(split_spin_limit=0)
    x1, e1, _ = run(scan)
    np.testing.assert_array_equal(x1, x0)
    np.testing.assert_array_equal(e1, e0)


def test_abort_while_idle_changes_nothing(scan, opts):
    x0, e0, _ = run(scan)
    x1, e1, _ = run(scan, abort_between=True)
    np.testing.assert_array_equal(x1, x0)
    np.testing.assert_array_equal(e1, e0)
