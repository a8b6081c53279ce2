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
upsampling, loss, Adam) is the library's own (tb_nn_*, tb_unet.hip).  torch
only allocates the device buffers and supplies the stream.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import lib as _lib

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
    return C.c_void_p(t.data_ptr())


class UNet:
    """The network's parameters and optimiser state on the GPU, and its
    forward and backward passes built from the tb_nn_* kernels."""

    def __init__(self, params: dict | None = None, seed: int = 0, device: str = "cuda"):
        import torch
        self.torch = torch
        self.dev = torch.device(device)
        if params is None:
            params = init_params(seed)
        self.master, self.wbf, self.m, self.v, self.grad = {}, {}, {}, {}, {}
        for name, (w, b) in params.items():
            for k, arr in (("w", w), ("b", b)):
                t = torch.as_tensor(np.ascontiguousarray(arr, np.float32)).to(self.dev)
                self.master[(name, k)] = t
                self.m[(name, k)] = torch.zeros_like(t)
                self.v[(name, k)] = torch.zeros_like(t)
                self.grad[(name, k)] = torch.zeros_like(t)
            if name != "head":
                self.wbf[name] = self.master[(name, "w")].to(torch.bfloat16)
        self.step_count = 0
        self.acts: dict = {}

    # -- helpers ---------------------------------------------------------------
    def _stream(self):
        return C.c_void_p(self.torch.cuda.current_stream(self.dev).cuda_stream)

    def _bf(self, shape):
        return self.torch.empty(shape, dtype=self.torch.bfloat16, device=self.dev)

    def params(self) -> dict:
        """Host copies of the master parameters, {name: (w, b)}."""
        out = {}
        for name, _, _, _ in ARCH + [("head", None, None, None)]:
            out[name] = (self.master[(name, "w")].cpu().numpy(),
                         self.master[(name, "b")].cpu().numpy())
        return out

    # -- forward ---------------------------------------------------------------
    def forward(self, fbp, mean=None, keep: bool = False, out=None):
        """fbp: float32 [N][D][H][W] on the device (D, H, W divisible by 4),
        a torch tensor or (device address, shape); mean: float32 [D][H][W]
        (shared by the N samples) or [N][D][H][W], or None.  Returns the
        float32 [N][D][H][W] output, written to `out` (a tensor or a device
        address; it may be fbp's own memory) when given."""
        L, s = nn_lib(), self._stream()
        if isinstance(fbp, tuple):
            f_addr, (N, D, H, W) = fbp
        else:
            f_addr, (N, D, H, W) = fbp.data_ptr(), tuple(fbp.shape)
        if D % 4 or H % 4 or W % 4:
            raise ValueError(f"volume {D}x{H}x{W}: every side must be divisible by 4")
        nvox = N * D * H * W
        per = D * H * W
        if mean is not None and mean.numel() == nvox:
            per = nvox
        x = self._bf((N, D, H, W, 32))
        _ok(L.tb_nn_make_input(C.c_void_p(f_addr), _p(mean) if mean is not None else None,
                               _p(x), nvox, per, SCALE, s))
        acts = {"x": (x, (N, D, H, W))}
        for name, src, cin, cout in ARCH:
            if src.startswith("pool:"):
                a, (n_, d_, h_, w_) = acts[src[5:]]
                xin = self._bf((n_, d_ // 2, h_ // 2, w_ // 2, cin))
                _ok(L.tb_nn_avgpool2(_p(a), _p(xin), n_, d_, h_, w_, cin, 0, s))
                shp = (n_, d_ // 2, h_ // 2, w_ // 2)
                acts[src] = (xin, shp)
            elif src.startswith("cat:"):
                a, b = src[4:].split(",")
                ta, _ = acts[a]
                tb_, shp = acts[b]
                ca, cb = ta.shape[-1], tb_.shape[-1]
                xin = self._bf(shp + (ca + cb,))
                _ok(L.tb_nn_up2_concat(_p(ta), _p(tb_), _p(xin), *shp, ca, cb, s))
                acts[src] = (xin, shp)
            else:
                xin, shp = acts[src]
            y = self._bf(shp + (cout,))
            _ok(L.tb_nn_conv3d(_p(xin), _p(self.wbf[name]), _p(self.master[(name, "b")]),
                               _p(y), *shp, cin, cout, 1, 0, s))
            acts[name] = (y, shp)
        if out is None:
            out = self.torch.empty((N, D, H, W), dtype=self.torch.float32, device=self.dev)
        o_addr = out if isinstance(out, int) else out.data_ptr()
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
        torch = self.torch
        stage.synchronize()
        addr, _ = stage.object_device()
        i = stage.info
        F, per = i["N_time"], i["N_z"] * i["N_y"] * i["N_x"]
        shape = (i["N_z"], i["N_y"], i["N_x"])
        mean = torch.empty(shape, dtype=torch.float32, device=self.dev)
        _ok(nn_lib().tb_nn_frame_mean(C.c_void_p(addr), _p(mean), F, per, self._stream()))
        for f0 in range(0, F, frames_per_pass):
            n = min(frames_per_pass, F - f0)
            a = addr + f0 * per * 4
            self.forward((a, (n,) + shape), mean, out=a)
        torch.cuda.synchronize(self.dev)

    # -- backward --------------------------------------------------------------
    def backward(self, gout):
        """Gradients of sum(gout * forward(...)) into self.grad (the forward
        must have run with keep=True)."""
        L, s = nn_lib(), self._stream()
        acts = self.acts
        torch = self.torch
        h, shp = acts["d1"]
        nvox = int(np.prod(shp))
        part = torch.empty((SLICES, 33), dtype=torch.float32, device=self.dev)
        gh = self._bf(shp + (32,))
        _ok(L.tb_nn_head_bwd(_p(h), _p(self.master[("head", "w")]), _p(gout), _p(gh), _p(part),
                             SLICES, nvox, 32, SCALE, s))
        hb = torch.empty(33, dtype=torch.float32, device=self.dev)
        _ok(L.tb_nn_sum_parts(_p(part), _p(hb), 33, SLICES, s))
        self.grad[("head", "w")].copy_(hb[:32])
        self.grad[("head", "b")].copy_(hb[32:])
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
                wpart = torch.empty(need, dtype=torch.float32, device=self.dev)
            _ok(L.tb_nn_conv3d_wgrad(_p(xin), _p(gy), _p(self.grad[(name, "w")]), _p(wpart),
                                     SLICES, *shp, cin, cout, s))
            _ok(L.tb_nn_channel_sum(_p(gy), _p(self.grad[(name, "b")]), _p(wpart), SLICES, nv,
                                    cout, s))
            if src == "x":
                continue
            wt = self._bf((27, cin, cout))
            _ok(L.tb_nn_conv3d_flip(_p(self.wbf[name]), _p(wt), cin, cout, s))
            gx = self._bf(shp + (cin,))
            _ok(L.tb_nn_conv3d(_p(gy), _p(wt), None, _p(gx), *shp, cout, cin, 0, 0, s))
            if src.startswith("pool:"):
                tgt = src[5:]
                _, tshp = acts[tgt]
                gt = self._bf(tshp + (cin,))
                _ok(L.tb_nn_avgpool2(_p(gx), _p(gt), *tshp, cin, 1, s))
                self._accumulate(g, tgt, gt)
            elif src.startswith("cat:"):
                a, b = src[4:].split(",")
                ta, ashp = acts[a]
                ca, cb = ta.shape[-1], cin - ta.shape[-1]
                ga = self._bf(ashp + (ca,))
                gb = self._bf(shp + (cb,))
                _ok(L.tb_nn_up2_concat_bwd(_p(gx), _p(ga), _p(gb), *shp, ca, cb, s))
                self._accumulate(g, a, ga)
                self._accumulate(g, b, gb)
            else:
                self._accumulate(g, src, gx)

    def _accumulate(self, g, key, t):
        if key in g:
            _ok(nn_lib().tb_nn_add(_p(g[key]), _p(t), t.numel(), self._stream()))
        else:
            g[key] = t

    # -- training --------------------------------------------------------------
    def adam(self, lr: float, b1=0.9, b2=0.999, eps=1e-8):
        L, s = nn_lib(), self._stream()
        self.step_count += 1
        for key, p in self.master.items():
            name, kind = key
            wb = self.wbf[name] if (kind == "w" and name != "head") else None
            _ok(L.tb_nn_adam(_p(p), _p(self.grad[key]), _p(self.m[key]), _p(self.v[key]),
                             _p(wb) if wb is not None else None, p.numel(), lr, b1, b2, eps,
                             self.step_count, s))

    def loss_and_grad(self, out, target):
        """Mean squared error and its gradient (fp32)."""
        L, s = nn_lib(), self._stream()
        torch = self.torch
        g = torch.empty_like(out)
        blocks = 256
        part = torch.empty(blocks, dtype=torch.float64, device=self.dev)
        _ok(L.tb_nn_mse(_p(out), _p(target), _p(g), _p(part), blocks, out.numel(), s))
        return part.sum().item() / out.numel(), g

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
        z = np.load(path)
        out = {}
        for name, _, _, _ in ARCH + [("head", None, None, None)]:
            out[name] = (z[name + ".w"], z[name + ".b"])
        return out


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
