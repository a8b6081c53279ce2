"""Pins of the learned initialiser's CPU oracle (oracle/unet_ref.py), CPU
only.  Each primitive is checked against torch's CPU implementation (a
library that shares none of the oracle's code): conv3d against
torch.nn.functional.conv3d, its input and weight gradients against torch
autograd, pooling and upsampling against avg_pool3d / interpolate.  The whole
network's backward pass is checked against torch autograd of the same
network assembled from torch primitives, and against central finite
differences, in float64 with the bf16 rounding switched off."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
import torch.nn.functional as F  # noqa: E402

from oracle import unet_ref as U  # noqa: E402


def to_torch_w(w):
    """[27][Cout][Cin] -> torch's [Cout][Cin][3][3][3] (cross-correlation)."""
    return torch.from_numpy(np.ascontiguousarray(w.reshape(3, 3, 3, *w.shape[1:])
                                                 .transpose(3, 4, 0, 1, 2)))


def to_torch_x(x):
    return torch.from_numpy(np.ascontiguousarray(x.transpose(0, 4, 1, 2, 3)))


def from_torch_x(t):
    return t.detach().numpy().transpose(0, 2, 3, 4, 1)


@pytest.mark.parametrize("shape,cin,cout", [((2, 4, 6, 10), 32, 64), ((1, 3, 5, 7), 64, 32)])
def test_conv3d_and_its_gradients_match_torch(shape, cin, cout):
    rng = np.random.default_rng(1)
    x = rng.standard_normal(shape + (cin,))
    w = rng.standard_normal((27, cout, cin))
    b = rng.standard_normal(cout)
    g = rng.standard_normal(shape + (cout,))
    xt = to_torch_x(x).requires_grad_(True)
    wt = to_torch_w(w).requires_grad_(True)
    yt = F.conv3d(xt, wt, torch.from_numpy(b), padding=1)
    np.testing.assert_allclose(U.conv3d(x, w, b), from_torch_x(yt), rtol=1e-10, atol=1e-10)
    (yt * to_torch_x(g)).sum().backward()
    np.testing.assert_allclose(U.conv3d_bwd_data(g, w), from_torch_x(xt.grad), rtol=1e-10,
                               atol=1e-10)
    gw = wt.grad.numpy().transpose(2, 3, 4, 0, 1).reshape(27, cout, cin)
    np.testing.assert_allclose(U.conv3d_wgrad(x, g), gw, rtol=1e-10, atol=1e-9)


def test_pooling_and_upsampling_match_torch():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((2, 4, 6, 8, 3))
    np.testing.assert_allclose(U.avgpool2(x), from_torch_x(F.avg_pool3d(to_torch_x(x), 2)),
                               rtol=1e-12)
    c = rng.standard_normal((2, 2, 3, 4, 3))
    np.testing.assert_array_equal(
        U.up2(c), from_torch_x(F.interpolate(to_torch_x(c), scale_factor=2, mode="nearest")))
    # the backward passes are the adjoints: <A x, y> = <x, A^T y>
    y = rng.standard_normal((2, 2, 3, 4, 3))
    assert abs(np.sum(U.avgpool2(x) * y) - np.sum(x * U.avgpool2_bwd(y))) < 1e-10
    z = rng.standard_normal((2, 4, 6, 8, 3))
    assert abs(np.sum(U.up2(c) * z) - np.sum(c * U.up2_bwd(z))) < 1e-10


def test_bf16_rounding_is_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5e-3, 0.0],
                    np.float32)
    want = torch.from_numpy(vals).to(torch.bfloat16).float().numpy()
    np.testing.assert_array_equal(U.bf16(vals), want)
    r = np.random.default_rng(3).standard_normal(10000).astype(np.float32) * 30
    np.testing.assert_array_equal(U.bf16(r), torch.from_numpy(r).to(torch.bfloat16).float().numpy())


def small_params(seed, scale=0.5):
    rng = np.random.default_rng(seed)
    p = U.init_params(seed)
    out = {}
    for k, (w, b) in p.items():
        out[k] = (w.astype(np.float64), rng.standard_normal(b.shape) * 0.1)
    out["head"] = (rng.standard_normal(32) * scale, np.array([0.05]))
    return out


def torch_network(params, fbp, mean, scale):
    """The same network from torch primitives (float64, channels first)."""
    P = {k: (torch.from_numpy(to_torch_w(w).numpy() if k != "head" else w).requires_grad_(True),
             torch.from_numpy(b).requires_grad_(True)) for k, (w, b) in params.items()}
    x = np.zeros(fbp.shape + (32,))
    x[..., 0] = scale * fbp
    x[..., 1] = scale * mean
    acts = {"x": to_torch_x(x)}
    for name, src, _, _ in U.ARCH:
        if src.startswith("pool:"):
            xin = F.avg_pool3d(acts[src[5:]], 2)
        elif src.startswith("cat:"):
            a, b = src[4:].split(",")
            xin = torch.cat([F.interpolate(acts[a], scale_factor=2, mode="nearest"), acts[b]], 1)
        else:
            xin = acts[src]
        acts[name] = F.relu(F.conv3d(xin, P[name][0], P[name][1], padding=1))
    h = acts["d1"]
    out = torch.from_numpy(fbp) + (torch.einsum("ncdhw,c->ndhw", h, P["head"][0])
                                   + P["head"][1][0]) / scale
    return out, P


def test_network_gradients_match_torch_autograd():
    rng = np.random.default_rng(4)
    shape = (2, 4, 8, 8)
    fbp = rng.random(shape) * 0.02
    mean = fbp.mean(axis=0)
    params = small_params(5)
    out, acts = U.forward(params, fbp, mean, 50.0, rnd=U.exact)
    g = rng.standard_normal(shape)
    grads = U.backward(params, acts, g, 50.0, rnd=U.exact)
    tout, P = torch_network(params, fbp, mean, 50.0)
    np.testing.assert_allclose(out, tout.detach().numpy(), rtol=1e-10, atol=1e-12)
    (tout * torch.from_numpy(g)).sum().backward()
    for name in [a[0] for a in U.ARCH]:
        tw = P[name][0].grad.numpy().transpose(2, 3, 4, 0, 1).reshape(grads[name][0].shape)
        np.testing.assert_allclose(grads[name][0], tw, rtol=1e-8, atol=1e-10 * np.abs(tw).max())
        np.testing.assert_allclose(grads[name][1], P[name][1].grad.numpy(), rtol=1e-8,
                                   atol=1e-12)
    np.testing.assert_allclose(grads["head"][0], P["head"][0].grad.numpy(), rtol=1e-10)
    np.testing.assert_allclose(grads["head"][1], P["head"][1].grad.numpy(), rtol=1e-10)


def test_network_gradients_match_finite_differences():
    rng = np.random.default_rng(6)
    shape = (1, 4, 8, 8)
    fbp = rng.random(shape) * 0.02
    mean = fbp[0]
    params = small_params(7)
    g = rng.standard_normal(shape)

    def f(p):
        return float(np.sum(U.forward(p, fbp, mean, 50.0, rnd=U.exact)[0] * g))
    _, acts = U.forward(params, fbp, mean, 50.0, rnd=U.exact)
    grads = U.backward(params, acts, g, 50.0, rnd=U.exact)
    for name, idx in (("e1a", (13, 5, 1)), ("b2", (4, 100, 7)), ("d2a", (26, 3, 150)),
                      ("d1", (0, 31, 2))):
        w, b = params[name]
        h = 1e-6
        wp, wm = w.copy(), w.copy()
        wp[idx] += h
        wm[idx] -= h
        fd = (f(dict(params, **{name: (wp, b)})) - f(dict(params, **{name: (wm, b)}))) / (2 * h)
        assert abs(fd - grads[name][0][idx]) <= 1e-6 * max(1.0, abs(fd)), (name, fd,
                                                                             grads[name][0][idx])


def test_gradient_fixture_is_the_oracle_output():
    """tests/golden/unet_grads.npz holds what the oracle computes today for
    the fixture's seeded case (numpy only: deterministic)."""
    import os
    import sys
    gold_dir = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    sys.path.insert(0, gold_dir)
    import make_unet_fixture as fx
    params, fbp, mean, target = fx.case()
    out, loss, grads = fx.oracle_grads(params, fbp, mean, target)
    gold = np.load(os.path.join(gold_dir, "unet_grads.npz"))
    assert abs(loss - float(gold["loss"])) <= 1e-12 * loss
    s = fx.summarise(grads)
    for k, v in s.items():
        np.testing.assert_allclose(v, gold[k], rtol=1e-9)
    np.testing.assert_allclose(grads["head"][0], gold["head_w"], rtol=1e-9)
