"""Learned initialiser: a 3D U-Net that maps a filtered back-projection to a
starting object for ICD.

The reference starts its first stage from zero (XT_ICD_update.c:549-553)
and has a path that starts a stage from an object read without
interpolation (initICD = 1, XT_ICD_update.c:558-571).  Here the finest stage
starts from the network's output instead, written into the stage's object in
HBM (tb_stage_object_device).  The MAP problem is the same whatever the
start, so ICD convergence is the judge of the network's error; zero or the
back-projection remain the fallback (tb_options.init_fbp).

Network (3 levels, 32 base channels, 3x3x3 convolutions, bf16 weights and
activations, fp32 accumulation on the matrix cores):
    x (2 channels: the frame's back-projection and the mean over frames,
       padded to 32) -> e1a, e1 (32) -> pool -> e2a, e2 (64) -> pool ->
    b1, b2 (128) -> up + [e2] -> d2a, d2 (64) -> up + [e1] -> d1a, d1 (32)
    -> out = fbp + (head . d1 + b) / scale
Every kernel (convolution forward, weight and input gradients, pooling,
upsampling, loss, Adam) is the library's own (tb_nn_*, tb_unet.hip); the
buffers come from its own allocator (timbir_amd.devmem), and everything runs
on its runtime's default stream.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib as _lib
from .devmem import DevArray, synchronize

# (name, input, Cin, Cout): 3x3x3 convolution + ReLU, in order
ARCH = [
    ("e1a", "x", 32, 32), ("e1", "e1a", 32, 32),
    ("e2a", "pool:e1", 32, 64), ("e2", "e2a", 64, 64),
    ("b1", "pool:e2", 64, 128), ("b2", "b1", 128, 128),
    ("d2a", "cat:b2,e2", 192, 64), ("d2", "d2a", 64, 64),
    ("d1a", "cat:d2,e1", 96, 32), ("d1", "d1a", 32, 32),
]
SCALE = 50.0      # attenuation per pixel (0-0.02) to O(1) network inputs
SLICES = 256      # partial-sum slices of the weight-gradient reductions

_BOUND = False


def nn_lib():
    """The library with the tb_nn_* signatures bound."""
    global _BOUND
    L = _lib()
    if _BOUND:
        return L
    vp, i32, i64, f32 = C.c_void_p, C.c_int32, C.c_int64, C.c_float
    sig = {
        "tb_nn_conv3d": [vp, vp, vp, vp] + [i32] * 8 + [vp],
        "tb_nn_conv3d_flip": [vp, vp, i32, i32, vp],
        "tb_nn_conv3d_wgrad": [vp, vp, vp, vp] + [i32] * 7 + [vp],
        "tb_nn_channel_sum": [vp, vp, vp, i32, i64, i32, vp],
        "tb_nn_relu_bwd": [vp, vp, i64, vp],
        "tb_nn_avgpool2": [vp, vp] + [i32] * 6 + [vp],
        "tb_nn_up2_concat": [vp, vp, vp] + [i32] * 6 + [vp],
        "tb_nn_up2_concat_bwd": [vp, vp, vp] + [i32] * 6 + [vp],
        "tb_nn_add": [vp, vp, i64, vp],
        "tb_nn_make_input": [vp, vp, vp, i64, i64, f32, vp],
        "tb_nn_frame_mean": [vp, vp, i64, i64, vp],
        "tb_nn_head": [vp, vp, vp, vp, vp, i64, i32, f32, vp],
        "tb_nn_head_bwd": [vp, vp, vp, vp, vp, i32, i64, i32, f32, vp],
        "tb_nn_mse": [vp, vp, vp, vp, i32, i64, vp],
        "tb_nn_sum_parts": [vp, vp, i64, i32, vp],
        "tb_nn_adam": [vp, vp, vp, vp, vp, i64, f32, f32, f32, f32, i32, vp],
        "tb_nn_to_bf16": [vp, vp, i64, vp],
        "tb_nn_add": [vp, vp, i64, vp],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    L.tb_nn_last_error.restype = C.c_char_p
    _BOUND = True
    return L


def _ok(rc):
    if rc != 0:
        from . import TimbirError
        raise TimbirError(nn_lib().tb_nn_last_error().decode())


def _p(t):
    return t.ptr if t is not None else None


S0 = None   # the runtime's default stream


class UNet:
    """The network's parameters and optimiser state on the GPU, and its
    forward and backward passes built from the tb_nn_* kernels."""

    def __init__(self, params: dict | None = None, seed: int = 0):
        if params is None:
            params = init_params(seed)
        self.master, self.wbf, self.m, self.v, self.grad = {}, {}, {}, {}, {}
        for name, (w, b) in params.items():
            for k, arr in (("w", w), ("b", b)):
                t = DevArray.from_numpy(np.ascontiguousarray(arr, np.float32))
                self.master[(name, k)] = t
                self.m[(name, k)] = DevArray(t.shape, "f32", zero=True)
                self.v[(name, k)] = DevArray(t.shape, "f32", zero=True)
                self.grad[(name, k)] = DevArray(t.shape, "f32", zero=True)
            if name != "head":
                wb = DevArray(self.master[(name, "w")].shape, "bf16")
                _ok(nn_lib().tb_nn_to_bf16(_p(self.master[(name, "w")]), _p(wb), wb.numel(), S0))
                self.wbf[name] = wb
        self.step_count = 0
        self.acts: dict = {}

    # -- helpers ---------------------------------------------------------------
    def params(self) -> dict:
        """Host copies of the master parameters, {name: (w, b)}."""
        out = {}
        for name, _, _, _ in ARCH + [("head", None, None, None)]:
            out[name] = (self.master[(name, "w")].numpy(), self.master[(name, "b")].numpy())
        return out

    def grads(self) -> dict:
        """Host copies of the last backward pass's gradients, {(name, 'w'|'b'): array}."""
        return {k: v.numpy() for k, v in self.grad.items()}

    # -- forward ---------------------------------------------------------------
    def forward(self, fbp, mean=None, keep: bool = False, out=None):
        """fbp: float32 [N][D][H][W] on the device (D, H, W divisible by 4),
        a DevArray or (device address, shape); mean: a float32 DevArray
        [D][H][W] (shared by the N samples) or [N][D][H][W], or None.
        Returns the float32 [N][D][H][W] output as a DevArray, written to
        `out` (a DevArray or a device address; it may be fbp's own memory)
        when given."""
        L, s = nn_lib(), S0
        if isinstance(fbp, tuple):
            f_addr, (N, D, H, W) = fbp
        else:
            f_addr, (N, D, H, W) = fbp.addr, tuple(fbp.shape)
        if D % 4 or H % 4 or W % 4:
            raise ValueError(f"volume {D}x{H}x{W}: every side must be divisible by 4")
        nvox = N * D * H * W
        per = D * H * W
        if mean is not None and mean.numel() == nvox:
            per = nvox
        x = DevArray((N, D, H, W, 32), "bf16")
        _ok(L.tb_nn_make_input(C.c_void_p(f_addr), _p(mean), _p(x), nvox, per, SCALE, s))
        acts = {"x": (x, (N, D, H, W))}
        for name, src, cin, cout in ARCH:
            if src.startswith("pool:"):
                a, (n_, d_, h_, w_) = acts[src[5:]]
                xin = DevArray((n_, d_ // 2, h_ // 2, w_ // 2, cin), "bf16")
                _ok(L.tb_nn_avgpool2(_p(a), _p(xin), n_, d_, h_, w_, cin, 0, s))
                shp = (n_, d_ // 2, h_ // 2, w_ // 2)
                acts[src] = (xin, shp)
            elif src.startswith("cat:"):
                a, b = src[4:].split(",")
                ta, _ = acts[a]
                tb_, shp = acts[b]
                ca, cb = ta.shape[-1], tb_.shape[-1]
                xin = DevArray(shp + (ca + cb,), "bf16")
                _ok(L.tb_nn_up2_concat(_p(ta), _p(tb_), _p(xin), *shp, ca, cb, s))
                acts[src] = (xin, shp)
            else:
                xin, shp = acts[src]
            y = DevArray(shp + (cout,), "bf16")
            _ok(L.tb_nn_conv3d(_p(xin), _p(self.wbf[name]), _p(self.master[(name, "b")]),
                               _p(y), *shp, cin, cout, 1, 0, s))
            acts[name] = (y, shp)
            if not keep:
                # free what no later layer reads
                for k in [k for k in acts if k not in _LIVE[name]]:
                    del acts[k]
        if out is None:
            out = DevArray((N, D, H, W), "f32")
        o_addr = out if isinstance(out, int) else out.addr
        _ok(L.tb_nn_head(_p(acts["d1"][0]), _p(self.master[("head", "w")]),
                         _p(self.master[("head", "b")]), C.c_void_p(f_addr), C.c_void_p(o_addr),
                         nvox, 32, SCALE, s))
        self.acts = acts if keep else {}
        return out

    def predict_into_stage(self, stage, frames_per_pass: int = 4):
        """Replace a stage's object (its filtered back-projection, from
        tb_init_fbp) by the network's output, in HBM: the temporal mean of
        the back-projection, then the network over a few frames at a time,
        written in place.  Returns when the GPU has finished."""
        stage.synchronize()
        addr, _ = stage.object_device()
        i = stage.info
        F, per = i["N_time"], i["N_z"] * i["N_y"] * i["N_x"]
        shape = (i["N_z"], i["N_y"], i["N_x"])
        mean = DevArray(shape, "f32")
        _ok(nn_lib().tb_nn_frame_mean(C.c_void_p(addr), _p(mean), F, per, S0))
        for f0 in range(0, F, frames_per_pass):
            n = min(frames_per_pass, F - f0)
            a = addr + f0 * per * 4
            self.forward((a, (n,) + shape), mean, out=a)
        synchronize()

    # -- backward --------------------------------------------------------------
    def backward(self, gout):
        """Gradients of sum(gout * forward(...)) into self.grad (the forward
        must have run with keep=True)."""
        L, s = nn_lib(), S0
        acts = self.acts
        h, shp = acts["d1"]
        nvox = int(np.prod(shp))
        part = DevArray((SLICES, 33), "f32")
        gh = DevArray(shp + (32,), "bf16")
        _ok(L.tb_nn_head_bwd(_p(h), _p(self.master[("head", "w")]), _p(gout), _p(gh), _p(part),
                             SLICES, nvox, 32, SCALE, s))
        hb = DevArray((33,), "f32")
        _ok(L.tb_nn_sum_parts(_p(part), _p(hb), 33, SLICES, s))
        synchronize()
        both = hb.numpy()
        self.grad[("head", "w")].upload(both[:32])
        self.grad[("head", "b")].upload(both[32:])
        g = {"d1": gh}
        wpart = None
        for name, src, cin, cout in reversed(ARCH):
            gy = g.pop(name)
            y, shp = acts[name]
            _ok(L.tb_nn_relu_bwd(_p(y), _p(gy), gy.numel(), s))
            xin, _ = acts[src]
            nv = int(np.prod(shp))
            need = SLICES * 27 * cin * cout
            if wpart is None or wpart.numel() < need:
                wpart = DevArray((need,), "f32")
            _ok(L.tb_nn_conv3d_wgrad(_p(xin), _p(gy), _p(self.grad[(name, "w")]), _p(wpart),
                                     SLICES, *shp, cin, cout, s))
            _ok(L.tb_nn_channel_sum(_p(gy), _p(self.grad[(name, "b")]), _p(wpart), SLICES, nv,
                                    cout, s))
            if src == "x":
                continue
            wt = DevArray((27, cin, cout), "bf16")
            _ok(L.tb_nn_conv3d_flip(_p(self.wbf[name]), _p(wt), cin, cout, s))
            gx = DevArray(shp + (cin,), "bf16")
            _ok(L.tb_nn_conv3d(_p(gy), _p(wt), None, _p(gx), *shp, cout, cin, 0, 0, s))
            if src.startswith("pool:"):
                tgt = src[5:]
                _, tshp = acts[tgt]
                gt = DevArray(tshp + (cin,), "bf16")
                _ok(L.tb_nn_avgpool2(_p(gx), _p(gt), *tshp, cin, 1, s))
                self._accumulate(g, tgt, gt)
            elif src.startswith("cat:"):
                a, b = src[4:].split(",")
                ta, ashp = acts[a]
                ca, cb = ta.shape[-1], cin - ta.shape[-1]
                ga = DevArray(ashp + (ca,), "bf16")
                gb = DevArray(shp + (cb,), "bf16")
                _ok(L.tb_nn_up2_concat_bwd(_p(gx), _p(ga), _p(gb), *shp, ca, cb, s))
                self._accumulate(g, a, ga)
                self._accumulate(g, b, gb)
            else:
                self._accumulate(g, src, gx)

    def _accumulate(self, g, key, t):
        if key in g:
            _ok(nn_lib().tb_nn_add(_p(g[key]), _p(t), t.numel(), S0))
        else:
            g[key] = t

    # -- training --------------------------------------------------------------
    def adam(self, lr: float, b1=0.9, b2=0.999, eps=1e-8):
        L = nn_lib()
        self.step_count += 1
        for key, p in self.master.items():
            name, kind = key
            wb = self.wbf[name] if (kind == "w" and name != "head") else None
            _ok(L.tb_nn_adam(_p(p), _p(self.grad[key]), _p(self.m[key]), _p(self.v[key]),
                             _p(wb), p.numel(), lr, b1, b2, eps, self.step_count, S0))

    def loss_and_grad(self, out, target):
        """Mean squared error (a float, after a wait) and its gradient."""
        L = nn_lib()
        g = DevArray(out.shape, "f32")
        blocks = 256
        part = DevArray((blocks,), "f64")
        _ok(L.tb_nn_mse(_p(out), _p(target), _p(g), _p(part), blocks, out.numel(), S0))
        return float(part.numpy().sum()) / out.numel(), g

    def train_step(self, fbp, mean, target, lr):
        out = self.forward(fbp, mean, keep=True)
        loss, g = self.loss_and_grad(out, target)
        self.backward(g)
        self.adam(lr)
        self.acts = {}
        return loss

    # -- persistence -----------------------------------------------------------
    def save(self, path: str, **meta):
        arrays = {}
        for name, (w, b) in self.params().items():
            arrays[name + ".w"] = w
            arrays[name + ".b"] = b
        for k, v in meta.items():
            arrays["meta." + k] = np.asarray(v)
        np.savez_compressed(path, **arrays)

    @staticmethod
    def load_params(path: str) -> dict:
        """Parameters saved by save(), or with the convolution weights as
        bf16 bit patterns (uint16), the values the kernels use."""
        from .devmem import from_bf16_bits
        z = np.load(path)
        out = {}
        for name, _, _, _ in ARCH + [("head", None, None, None)]:
            w = z[name + ".w"]
            w = from_bf16_bits(w) if w.dtype == np.uint16 else w.astype(np.float32)
            out[name] = (np.ascontiguousarray(w), z[name + ".b"].astype(np.float32))
        return out


def _live_sets() -> dict:
    """For each layer, the activations some later layer (or the head) still
    reads once it has run: what an inference pass keeps."""
    order = [a[0] for a in ARCH]
    need = {}
    for i, name in enumerate(order):
        later = set()
        for name2, src, _, _ in ARCH[i + 1:]:
            if src.startswith("pool:"):
                later.add(src[5:])
            elif src.startswith("cat:"):
                later.update(src[4:].split(","))
            else:
                later.add(src)
        later.add(name)
        later.add("d1")
        need[name] = later
    return need


_LIVE = _live_sets()


def init_params(seed: int = 0) -> dict:
    """He-normal convolution weights, zero biases and a zero head: the
    untrained network returns the back-projection unchanged."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, _, cin, cout in ARCH:
        std = np.sqrt(2.0 / (27 * cin))
        w = (rng.standard_normal((27, cout, cin)) * std).astype(np.float32)
        if name == "e1a":
            w[:, :, 2:] = 0.0   # padding channels of the input carry nothing
        p[name] = (w, np.zeros(cout, np.float32))
    p["head"] = (np.zeros(32, np.float32), np.zeros(1, np.float32))
    return p
