"""tb_init_error_sino_static: the allocation-free variant of
tb_init_error_sino (its scratch is a static per-GPU array), through the C
ABI, against the allocating entry point on the same stage: bit for bit, on
two calls in a row (the second on a smaller stage), and a clear error for a
stage whose scratch would not fit."""
import ctypes as C
import sys

import numpy as np
import pytest

import timbir_amd as tb
from timbir_amd import datagen

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def static_init(st):
    L = tb.lib()
    L.tb_init_error_sino_static.argtypes = [C.c_void_p]
    L.tb_init_error_sino_static.restype = C.c_int
    return L.tb_init_error_sino_static(st.h)


def random_object(st, seed):
    i = st.info
    shape = (i["N_time"], i["N_z"], i["N_y"], i["N_x"])
    return (np.random.default_rng(seed).random(shape) * 0.02).astype(np.float32)


@pytest.mark.parametrize("order", ["large_then_small"])
def test_static_scratch_matches_the_allocating_entry_point(order):
    scans = [datagen.tiny(N=64, nz=12, n_frames=3, n_views=48, rings=0.01, remove_rings=1,
                          seed=31),
             datagen.tiny(N=32, nz=4, n_frames=2, n_views=24, seed=32)]
    for k, scan in enumerate(scans):
        log(f"static scratch call {k}: N={scan['proj_cols']} nz={scan['proj_rows']} "
            f"frames={scan['recon_num']}")
        a = tb.Stage(scan)
        b = tb.Stage(scan)
        x = random_object(a, 40 + k)
        d = np.random.default_rng(50 + k).normal(0, 0.01, (a.info["N_r"], a.info["N_t"]))
        for st in (a, b):
            st.set_object(x)
            st.set_offset(d.astype(np.float32))
        a.init_error_sino()
        assert static_init(b) == 0, tb.lib().tb_last_error().decode()
        np.testing.assert_array_equal(b.error_sino(), a.error_sino())
        a.close()
        b.close()


def test_static_scratch_rejects_a_stage_that_does_not_fit():
    # 16 frames x 512 x 512 x 8 slices: 128 MiB of interleaved object
    n, nz, nf, nv = 512, 8, 16, 32
    scan = dict(projections=np.zeros((nv, nz, n), np.float32),
                weights=np.ones((nv, nz, n), np.float32),
                proj_angles=np.linspace(0, np.pi, nv, endpoint=False).astype(np.float32),
                proj_times=np.arange(nv, dtype=np.float32),
                recon_times=np.linspace(0, nv, nf + 1).astype(np.float32),
                proj_rows=nz, proj_cols=n, proj_num=nv, recon_num=nf, vox_wid=1.0,
                rot_center=n / 2, sig_s=0.01, sig_t=1.0, c_s=1e-2, c_t=1e-2, convg_thresh=1.0,
                remove_rings=0, quad_convex=1, huber_delta=0.5, huber_T=4.0)
    log("static scratch: oversize stage")
    st = tb.Stage(scan)
    assert static_init(st) == -1
    msg = tb.lib().tb_last_error().decode()
    assert "static scratch" in msg and "tb_init_error_sino" in msg, msg
    # the allocating entry point still serves it
    st.init_error_sino()
    st.close()
