"""Device arrays for the learned initialiser's kernels, in the library's own
HIP runtime (tb_dev_malloc / tb_dev_copy / tb_dev_free).

The kernels (tb_nn_*) and these copies all run on that runtime's default
stream, so they are ordered with each other; tb_dev_synchronize waits for
all of them.  bf16 data crosses to the host as float32 (rounded to nearest
even on the way in).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import TimbirError, lib as _lib

_ITEM = {"f32": 4, "bf16": 2, "f64": 8, "i32": 4}
_BOUND = False


def _L():
    global _BOUND
    L = _lib()
    if not _BOUND:
        L.tb_dev_malloc.argtypes = [C.c_uint64, C.POINTER(C.c_void_p)]
        L.tb_dev_free.argtypes = [C.c_void_p]
        L.tb_dev_copy.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_int32]
        L.tb_dev_zero.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p]
        L.tb_dev_synchronize.argtypes = []
        L.tb_nn_last_error.restype = C.c_char_p
        _BOUND = True
    return L


def _ok(rc):
    if rc != 0:
        raise TimbirError(_L().tb_nn_last_error().decode())


def to_bf16_bits(a) -> np.ndarray:
    """float32 values -> bfloat16 bit patterns (uint16), nearest even."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    return (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF).astype(np.uint16)


def from_bf16_bits(b) -> np.ndarray:
    return (np.asarray(b, np.uint16).astype(np.uint32) << 16).view(np.float32)


def synchronize():
    _ok(_L().tb_dev_synchronize())


# Blocks freed by DevArray, by size, for reuse.  Every kernel and copy on
# them runs on the runtime's default stream, so a block handed out again is
# written only after the work that read it (stream order): no wait needed,
# and no hipFree (which would wait for the device) on the hot path.
_POOL: dict = {}


def release_pool():
    """Return every cached block to the runtime."""
    for blocks in _POOL.values():
        for a in blocks:
            _L().tb_dev_free(C.c_void_p(a))
    _POOL.clear()


class DevArray:
    """A device buffer of `shape` and `dtype` ('f32', 'bf16', 'f64', 'i32'),
    or a view of memory owned elsewhere (addr given)."""

    def __init__(self, shape, dtype: str = "f32", addr: int | None = None, zero: bool = False):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = dtype
        self.nbytes = int(np.prod(self.shape)) * _ITEM[dtype]
        self._own = addr is None
        if addr is None:
            blocks = _POOL.get(self.nbytes)
            if blocks:
                addr = blocks.pop()
            else:
                p = C.c_void_p()
                _ok(_L().tb_dev_malloc(self.nbytes, C.byref(p)))
                addr = p.value
            if zero:
                _ok(_L().tb_dev_zero(C.c_void_p(addr), self.nbytes, None))
        self.addr = addr

    @property
    def ptr(self):
        return C.c_void_p(self.addr)

    def numel(self) -> int:
        return int(np.prod(self.shape))

    def __del__(self):
        try:
            if self._own and self.addr:
                _POOL.setdefault(self.nbytes, []).append(self.addr)
                self.addr = None
        except Exception:
            pass

    @classmethod
    def from_numpy(cls, a, dtype: str | None = None) -> "DevArray":
        a = np.asarray(a)
        if dtype is None:
            dtype = {np.dtype(np.float32): "f32", np.dtype(np.float64): "f64",
                     np.dtype(np.int32): "i32"}[a.dtype]
        out = cls(a.shape, dtype)
        out.upload(a)
        return out

    def upload(self, a):
        if self.dtype == "bf16":
            h = to_bf16_bits(a)
        else:
            h = np.ascontiguousarray(a, {"f32": np.float32, "f64": np.float64,
                                         "i32": np.int32}[self.dtype])
        if h.nbytes != self.nbytes:
            raise ValueError(f"upload of {h.nbytes} bytes into {self.nbytes}")
        _ok(_L().tb_dev_copy(self.ptr, C.c_void_p(h.ctypes.data), self.nbytes, 0))

    def numpy(self) -> np.ndarray:
        """A host copy (bf16 as float32); waits for the device."""
        synchronize()
        if self.dtype == "bf16":
            h = np.empty(self.shape, np.uint16)
        else:
            h = np.empty(self.shape, {"f32": np.float32, "f64": np.float64,
                                      "i32": np.int32}[self.dtype])
        _ok(_L().tb_dev_copy(C.c_void_p(h.ctypes.data), self.ptr, self.nbytes, 1))
        return from_bf16_bits(h).reshape(self.shape) if self.dtype == "bf16" else h

    def copy_from(self, other: "DevArray"):
        if other.nbytes != self.nbytes:
            raise ValueError("copy between arrays of different sizes")
        _ok(_L().tb_dev_copy(self.ptr, other.ptr, self.nbytes, 2))

    def zero(self):
        _ok(_L().tb_dev_zero(self.ptr, self.nbytes, None))
