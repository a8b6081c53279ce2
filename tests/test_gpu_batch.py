"""Many reconstructions at once (timbir_amd.batch) on the GPU: concurrent
runs on separate streams give the same objects, bit for bit, as one at a
time and as reconstruct() (src/include/mbir4d.h:34) itself."""
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
