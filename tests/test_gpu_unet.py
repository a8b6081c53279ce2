"""The learned initialiser's kernels (tb_unet.hip, through the C ABI) against
the CPU oracle (oracle/unet_ref.py): the matrix-core convolution, its input
and weight gradients, pooling and upsampling, and the whole network's output
and parameter gradients (live, and against the committed fixture
tests/golden/unet_grads.npz).  bf16 storage rounds every activation to 8
significant bits; fp32 sums in another order can move a stored value by one
bf16 step, so network-level results are compared in relative L2 norm."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

import timbir_amd as tb
from oracle import unet_ref as U
from timbir_amd.devmem import DevArray, synchronize

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def nn():
    from timbir_amd import learned
    return learned


def dev(a):
    return DevArray.from_numpy(np.ascontiguousarray(a, np.float32))


def bf(a):
    return DevArray.from_numpy(np.asarray(a, np.float32), "bf16")


def empty(shape, dtype="f32"):
    return DevArray(shape, dtype)


def host(t):
    return t.numpy().astype(np.float32)


def ptr(t):
    return t.ptr if t is not None else None


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("shape,cin,cout", [((2, 4, 6, 70), 32, 64), ((1, 3, 5, 17), 96, 32),
                                            ((1, 2, 4, 130), 192, 128), ((2, 8, 8, 8), 64, 64)])
def test_conv3d_matches_oracle(nn, shape, cin, cout):
    """Forward (fp32 output: the fp32 sums of bf16 products, any order),
    input gradient through the flipped kernel, and weight gradient."""
    L = nn.nn_lib()
    rng = np.random.default_rng(cin + cout)
    x = U.bf16(rng.standard_normal(shape + (cin,)).astype(np.float32))
    w = U.bf16((rng.standard_normal((27, cout, cin)) / np.sqrt(27 * cin)).astype(np.float32))
    b = rng.standard_normal(cout).astype(np.float32)
    g = U.bf16(rng.standard_normal(shape + (cout,)).astype(np.float32))
    log(f"conv3d {shape} {cin}->{cout}")
    s = None
    xt, wt, bt, gt = bf(x), bf(w), dev(b), bf(g)
    y = empty(shape + (cout,))
    assert L.tb_nn_conv3d(ptr(xt), ptr(wt), ptr(bt), ptr(y), *shape, cin, cout, 0, 1, s) == 0
    ref = U.conv3d(x.astype(np.float64), w.astype(np.float64), b)
    err = np.abs(host(y) - ref).max() / np.abs(ref).max()
    log(f"  forward: max error {err:.2e} of max")
    assert err < 5e-5
    # ReLU and bf16 output
    yb = empty(shape + (cout,), "bf16")
    assert L.tb_nn_conv3d(ptr(xt), ptr(wt), ptr(bt), ptr(yb), *shape, cin, cout, 1, 0, s) == 0
    got = host(yb)
    want = U.bf16(np.maximum(ref, 0).astype(np.float32))
    assert np.mean(got != want) < 2e-3 and np.abs(got - want).max() <= 1e-2 * np.abs(want).max()
    # input gradient: the forward kernel with the flipped, transposed weights
    wf = empty((27, cin, cout), "bf16")
    assert L.tb_nn_conv3d_flip(ptr(wt), ptr(wf), cin, cout, s) == 0
    gx = empty(shape + (cin,))
    assert L.tb_nn_conv3d(ptr(gt), ptr(wf), None, ptr(gx), *shape, cout, cin, 0, 1, s) == 0
    ref_gx = U.conv3d_bwd_data(g.astype(np.float64), w.astype(np.float64))
    err = np.abs(host(gx) - ref_gx).max() / np.abs(ref_gx).max()
    log(f"  input gradient: max error {err:.2e} of max")
    assert err < 5e-5
    # weight gradient, several slice counts (the partial sums add in order)
    for slices in (1, 7, 64):
        gw = empty((27, cout, cin))
        part = empty((slices * 27 * cin * cout,))
        assert L.tb_nn_conv3d_wgrad(ptr(xt), ptr(gt), ptr(gw), ptr(part), slices, *shape, cin,
                                    cout, s) == 0
        ref_gw = U.conv3d_wgrad(x.astype(np.float64), g.astype(np.float64))
        err = np.abs(host(gw) - ref_gw).max() / np.abs(ref_gw).max()
        log(f"  weight gradient ({slices} slices): max error {err:.2e} of max")
        assert err < 5e-5


def test_pooling_upsampling_and_head_match_oracle(nn):
    L = nn.nn_lib()
    s = None
    rng = np.random.default_rng(3)
    shape = (2, 4, 6, 10)
    x = U.bf16(rng.standard_normal(shape + (32,)).astype(np.float32))
    xt = bf(x)
    p = empty((2, 2, 3, 5, 32), "bf16")
    assert L.tb_nn_avgpool2(ptr(xt), ptr(p), *shape, 32, 0, s) == 0
    gp = empty(shape + (32,), "bf16")
    assert L.tb_nn_avgpool2(ptr(p), ptr(gp), *shape, 32, 1, s) == 0
    a = U.bf16(rng.standard_normal((2, 2, 3, 5, 64)).astype(np.float32))
    cat = empty(shape + (96,), "bf16")
    assert L.tb_nn_up2_concat(ptr(bf(a)), ptr(xt), ptr(cat), *shape, 64, 32, s) == 0
    ga = empty((2, 2, 3, 5, 64), "bf16")
    gb = empty(shape + (32,), "bf16")
    assert L.tb_nn_up2_concat_bwd(ptr(cat), ptr(ga), ptr(gb), *shape, 64, 32, s) == 0
    # the mean of 8 bf16 values in fp32, rounded once: equal to the oracle's
    # up to one rounding step where the sums are formed in another order
    want = U.bf16(U.avgpool2(x.astype(np.float64)).astype(np.float32))
    assert np.abs(host(p) - want).max() <= 2 ** -7 * np.abs(want).max()
    np.testing.assert_array_equal(host(gp), U.bf16(U.avgpool2_bwd(host(p))))
    np.testing.assert_array_equal(host(cat), np.concatenate([U.up2(a), x], axis=-1))
    np.testing.assert_array_equal(host(gb), x)
    np.testing.assert_array_equal(host(ga), U.bf16(U.up2_bwd(U.up2(a).astype(np.float64))
                                                   .astype(np.float32)))


def run_network(nn, params, fbp, mean, target):
    net = nn.UNet(params=params)
    f, m, t = dev(fbp), dev(mean), dev(target)
    out = net.forward(f, m, keep=True)
    loss, g = net.loss_and_grad(out, t)
    net.backward(g)
    grads = net.grads()
    return host(out), loss, grads


def test_network_gradients_match_the_oracle(nn):
    """Output, loss and every parameter gradient of the GPU network against
    the oracle with the same bf16 rounding points, live and against the
    committed fixture (norms and seeded probe projections)."""
    sys.path.insert(0, GOLD)
    import make_unet_fixture as fx
    params, fbp, mean, target = fx.case()
    log(f"network gradients {fx.SHAPE}")
    out, loss, grads = run_network(nn, params, fbp, mean, target)
    o_out, o_loss, o_grads = fx.oracle_grads(params, fbp, mean, target)
    gold = np.load(os.path.join(GOLD, "unet_grads.npz"))
    assert abs(o_loss - float(gold["loss"])) <= 1e-12 * o_loss   # oracle reproduces its fixture
    r_out = rel(out - fbp, o_out - fbp)
    log(f"  output correction rel err {r_out:.2e}; loss {loss:.6e} vs {o_loss:.6e}")
    assert r_out < 1e-2
    assert abs(loss - o_loss) <= 1e-2 * o_loss
    worst = 0.0
    for name, (gw, gb) in o_grads.items():
        for kind, want in (("w", gw), ("b", gb)):
            got = grads[(name, kind)]
            r = rel(got, want)
            key = f"{name}.{kind}"
            pr = fx.probes(got.shape, key).reshape(fx.N_PROBES, -1) @ got.reshape(-1).astype(np.float64)
            rp = rel(pr, gold[key + ".proj"])
            rn = abs(np.linalg.norm(got) - float(gold[key + ".norm"])) / float(gold[key + ".norm"])
            worst = max(worst, r)
            log(f"  {key:6s} rel err {r:.2e} (fixture: norm {rn:.1e}, probes {rp:.1e})")
            assert r < 3e-2, key
            assert rn < 3e-2 and rp < 5e-2, key
    log(f"  worst gradient rel err {worst:.2e}")


def test_training_steps_reduce_the_loss(nn):
    """Adam steps on one sample reduce the squared error (the kernels'
    gradients point downhill)."""
    sys.path.insert(0, GOLD)
    import make_unet_fixture as fx
    params, fbp, mean, target = fx.case()
    net = nn.UNet(params=params)
    f, m, t = dev(fbp), dev(mean), dev(target)
    losses = [net.train_step(f, m, t, lr=1e-3) for _ in range(30)]
    log(f"  losses {losses[0]:.4e} -> {losses[-1]:.4e}")
    assert losses[-1] < 0.5 * losses[0]
