"""CPU oracle of the learned initialiser's network (numpy, fp32 arithmetic).

TEST INFRASTRUCTURE ONLY: imported by tests/ to check the GPU kernels of
timbir_amd/csrc/tb_unet.hip and the network built from them
(timbir_amd/learned.py), never by the product.

The reference has no network (its initial object is zero or a file read
without interpolation, XT_ICD_update.c:549-571); this oracle restates the
network the GPU trains, with the same rounding points: weights, activations
and activation gradients are stored as bf16 (round to nearest even),
products are accumulated in fp32.  tests/test_unet_oracle.py pins it against
torch's CPU conv3d and autograd, which share none of this code, and against
finite differences.

Layouts as on the GPU: activations [N][D][H][W][C] (channels last), conv
weights [27][Cout][Cin] with tap = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1).
"""
from __future__ import annotations

import numpy as np

TAPS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def bf16(x) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).reshape(a.shape)


def _shift(x, dz, dy, dx):
    """y[n, z, y, x] = x[n, z + dz, y + dy, x + dx], zero outside."""
    N, D, H, W, C = x.shape
    out = np.zeros_like(x)
    zs = slice(max(0, -dz), D - max(0, dz))
    ys = slice(max(0, -dy), H - max(0, dy))
    xs = slice(max(0, -dx), W - max(0, dx))
    zi = slice(max(0, dz), D - max(0, -dz))
    yi = slice(max(0, dy), H - max(0, -dy))
    xi = slice(max(0, dx), W - max(0, -dx))
    out[:, zs, ys, xs] = x[:, zi, yi, xi]
    return out


def conv3d(x, w, b=None, relu=False):
    """out[v][co] = b[co] + sum_tap sum_ci x[v + off(tap)][ci] w[tap][co][ci]."""
    out = np.zeros(x.shape[:4] + (w.shape[1],), np.result_type(x, w))
    for t, (dz, dy, dx) in enumerate(TAPS):
        out += _shift(x, dz, dy, dx) @ w[t].T
    if b is not None:
        out += b
    return np.maximum(out, 0) if relu else out


def conv3d_bwd_data(g, w):
    """d/dx of sum(g * conv3d(x, w)): sum_tap g[v - off(tap)] w[tap]."""
    out = np.zeros(g.shape[:4] + (w.shape[2],), np.result_type(g, w))
    for t, (dz, dy, dx) in enumerate(TAPS):
        out += _shift(g, -dz, -dy, -dx) @ w[t]
    return out


def conv3d_wgrad(x, g):
    """gw[tap][co][ci] = sum_v g[v][co] x[v + off(tap)][ci]."""
    C_o, C_i = g.shape[-1], x.shape[-1]
    gf = g.reshape(-1, C_o)
    return np.stack([gf.T @ _shift(x, dz, dy, dx).reshape(-1, C_i) for dz, dy, dx in TAPS])


def avgpool2(x):
    N, D, H, W, C = x.shape
    return x.reshape(N, D // 2, 2, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4, 6))


def avgpool2_bwd(g):
    return np.repeat(np.repeat(np.repeat(g, 2, 1), 2, 2), 2, 3) / 8.0


def up2(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 1), 2, 2), 2, 3)


def up2_bwd(g):
    N, D, H, W, C = g.shape
    return g.reshape(N, D // 2, 2, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4, 6))


# the network: (name, input, Cin, Cout) of each 3x3x3 conv + ReLU, in order;
# 'pool:a' / 'cat:a,b' name the other inputs
ARCH = [
    ("e1a", "x", 32, 32), ("e1", "e1a", 32, 32),
    ("e2a", "pool:e1", 32, 64), ("e2", "e2a", 64, 64),
    ("b1", "pool:e2", 64, 128), ("b2", "b1", 128, 128),
    ("d2a", "cat:b2,e2", 192, 64), ("d2", "d2a", 64, 64),
    ("d1a", "cat:d2,e1", 96, 32), ("d1", "d1a", 32, 32),
]


def exact(x):
    """No rounding: the same network in the arithmetic of its inputs (used
    with float64 inputs for finite-difference checks)."""
    return np.asarray(x)


def make_input(fbp, mean, scale, rnd=bf16):
    """[N][D][H][W] fbp and [D][H][W] mean -> [N][D][H][W][32] input."""
    x = np.zeros(fbp.shape + (32,), fbp.dtype)
    x[..., 0] = scale * fbp
    if mean is not None:
        x[..., 1] = scale * np.broadcast_to(mean, fbp.shape)
    return rnd(x)


def forward(params, fbp, mean, scale, rnd=bf16):
    """params: dict name -> (w [27][Cout][Cin], b [Cout]) plus 'head' ->
    (w [32], b [1]).  rnd is the rounding of stored weights and activations
    (bf16 as on the GPU, or `exact`).  Returns (out, acts)."""
    acts = {"x": make_input(fbp, mean, scale, rnd)}
    for name, src, cin, cout in ARCH:
        if src.startswith("pool:"):
            x = rnd(avgpool2(acts[src[5:]]))
            acts["pool:" + src[5:]] = x
        elif src.startswith("cat:"):
            a, b = src[4:].split(",")
            x = np.concatenate([up2(acts[a]), acts[b]], axis=-1)
            acts[src] = x
        else:
            x = acts[src]
        w, b = params[name]
        acts[name] = rnd(conv3d(x, rnd(w), b, relu=True))
    hw, hb = params["head"]
    h = acts["d1"]
    out = fbp + (h @ hw + hb[0]) / scale
    return out, acts


def backward(params, acts, gout, scale, rnd=bf16):
    """Gradients of sum(gout * out) w.r.t. every parameter, with the GPU's
    rounding of stored activation gradients (rnd)."""
    grads = {}
    hw, hb = params["head"]
    h = acts["d1"]
    gs = gout / scale
    grads["head"] = (np.tensordot(gs, h, axes=(list(range(gs.ndim)), list(range(gs.ndim)))),
                     np.array([gs.sum()]))
    g = {"d1": rnd(gs[..., None] * hw)}
    for name, src, cin, cout in reversed(ARCH):
        gy = g.pop(name)
        gy = np.where(acts[name] > 0, gy, 0).astype(gy.dtype)
        key = src if src.startswith(("pool:", "cat:")) else src
        x = acts[key]
        w, _ = params[name]
        grads[name] = (conv3d_wgrad(x, gy), gy.reshape(-1, cout).sum(axis=0))
        if src == "x":
            continue
        gx = rnd(conv3d_bwd_data(gy, rnd(w)))
        if src.startswith("pool:"):
            tgt = src[5:]
            gx = rnd(avgpool2_bwd(gx))
            g[tgt] = rnd(g.get(tgt, 0) + gx)
        elif src.startswith("cat:"):
            a, b = src[4:].split(",")
            ca = acts[a].shape[-1]
            ga = rnd(up2_bwd(gx[..., :ca]))
            gb = gx[..., ca:]
            g[a] = rnd(g.get(a, 0) + ga)
            g[b] = rnd(g.get(b, 0) + gb)
        else:
            g[src] = rnd(g.get(src, 0) + gx)
    return grads


def init_params(seed=0):
    """He-normal conv weights, zero biases, zero head (the network starts as
    the identity on the back-projection)."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, _, cin, cout in ARCH:
        std = np.sqrt(2.0 / (27 * cin))
        w = (rng.standard_normal((27, cout, cin)) * std).astype(np.float32)
        if name == "e1a":
            w[:, :, 2:] = 0.0   # padding channels carry nothing
        p[name] = (w, np.zeros(cout, np.float32))
    p["head"] = (np.zeros(32, np.float32), np.zeros(1, np.float32))
    return p
