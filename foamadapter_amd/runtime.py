"""Device context and device arrays over libfvhip (fv_ctx_*, fv_device_*).

``Context`` is the executor (createExecutor("GPU"), src/auxiliary/setup.cpp:51-80):
one HIP stream, an optional HBM budget.  ``DeviceArray`` is caller-owned device
memory (NeoN::Vector storage on the GPU executor).
"""
from __future__ import annotations

import atexit
import ctypes as C
import weakref

import numpy as np

from ._lib import FV_F32, FV_F64, call, lib

_DT = {np.dtype(np.float64): FV_F64, np.dtype(np.float32): FV_F32}


def prec_of(dtype) -> int:
    return _DT[np.dtype(dtype)]


_LIVE = weakref.WeakSet()


@atexit.register
def _close_all():
    # at interpreter exit, release every context's handles while they are all
    # still alive (finalizers would otherwise run in arbitrary order)
    for ctx in list(_LIVE):
        try:
            ctx.close()
        except Exception:
            pass


class Context:
    def __init__(self, device: int = 0, hbm_budget: int = 0):
        h = C.c_void_p()
        call("fv_ctx_create", device, hbm_budget, C.byref(h))
        self.handle = h
        self.device = device
        self._children = weakref.WeakSet()
        _LIVE.add(self)

    def adopt(self, obj):
        self._children.add(obj)
        return obj

    def sync(self):
        call("fv_sync", self.handle)

    def record(self, slot: int):
        """Record HIP event `slot` on the context's stream."""
        call("fv_event_record", self.handle, slot)

    def elapsed_ms(self, a: int, b: int) -> float:
        ms = C.c_float()
        call("fv_event_elapsed", self.handle, a, b, C.byref(ms))
        return float(ms.value)

    @property
    def bytes_in_use(self) -> int:
        return int(lib.fv_ctx_bytes_in_use(self.handle))

    def empty(self, shape, dtype=np.float64) -> "DeviceArray":
        return DeviceArray(self, shape, dtype)

    def upload(self, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = DeviceArray(self, a.shape, a.dtype)
        if a.nbytes:
            call("fv_memcpy_h2d", self.handle, d.ptr, C.c_void_p(a.ctypes.data), a.nbytes)
        return d

    def zeros(self, shape, dtype=np.float64) -> "DeviceArray":
        d = DeviceArray(self, shape, dtype)
        d.fill(0.0)
        return d

    def uniform(self, n: int, lo: float, hi: float, seed: int, stream: int,
                dtype=np.float64) -> "DeviceArray":
        d = DeviceArray(self, (n,), dtype)
        call("fv_fill_uniform", self.handle, prec_of(dtype), d.ptr, n, lo, hi, seed, stream)
        return d

    def close(self):
        if self.handle:
            for ch in list(self._children):
                try:
                    ch.close()
                except Exception:
                    pass
            lib.fv_ctx_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    def __init__(self, ctx: Context, shape, dtype=np.float64):
        self.ctx = ctx
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = C.c_void_p()
        call("fv_device_alloc", ctx.handle, self.nbytes, C.byref(p))
        self._ptr = p
        ctx.adopt(self)

    @property
    def ptr(self) -> C.c_void_p:
        return self._ptr

    @property
    def size(self) -> int:
        return int(np.prod(self.shape, dtype=np.int64))

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            call("fv_memcpy_d2h", self.ctx.handle, C.c_void_p(out.ctypes.data), self.ptr, self.nbytes)
        return out

    def copy_from(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        call("fv_memcpy_h2d", self.ctx.handle, self.ptr, C.c_void_p(a.ctypes.data), a.nbytes)

    def fill(self, value: float):
        call("fv_fill", self.ctx.handle, prec_of(self.dtype), self.ptr, self.size, float(value))

    def astype(self, dtype) -> "DeviceArray":
        d = DeviceArray(self.ctx, self.shape, dtype)
        call("fv_convert", self.ctx.handle, prec_of(dtype), d.ptr, prec_of(self.dtype), self.ptr,
             self.size)
        return d

    def close(self):
        if self._ptr and self._ptr.value and self.ctx.handle:
            lib.fv_device_free(self.ctx.handle, self._ptr)
        self._ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
