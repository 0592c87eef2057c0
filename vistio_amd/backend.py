"""GpuBackend (crates/vistio-gpu/src/backend.rs:19-46) on the MI355X.

ComputeBuffer (buffers.rs) holds f32 data on the host; here a buffer is a
contiguous float32 numpy array. Each primitive copies its operands to the
device, runs one kernel and copies the result back.
"""
import ctypes as C

import numpy as np

from . import _native as nat
from .solver import GpuError, _check


class ComputeBuffer:
    """buffers.rs:10-56: a length-n f32 buffer."""

    def __init__(self, data):
        self.data = np.ascontiguousarray(data, dtype=np.float32)

    @staticmethod
    def zeros(n: int) -> "ComputeBuffer":
        return ComputeBuffer(np.zeros(n, dtype=np.float32))

    @staticmethod
    def from_data(values) -> "ComputeBuffer":
        return ComputeBuffer(np.asarray(values, dtype=np.float32))

    def __len__(self) -> int:
        return int(self.data.size)

    def as_slice(self) -> np.ndarray:
        return self.data

    def copy_from_slice(self, src) -> None:
        src = np.asarray(src, dtype=np.float32)
        self.data[: src.size] = src


class HipBackend:
    """The trait's HIP implementation (what CpuFallback is on the CPU)."""

    def __init__(self, device: int = 0):
        self._device = device
        self._L = None
        self._ctx = C.c_void_p()

    def init(self) -> None:
        self._L = nat.load()
        _check(self._L.vh_create(self._device, C.byref(self._ctx)), what="vh_create")

    def name(self) -> str:
        return "hip_gfx950"

    def is_gpu(self) -> bool:
        return True

    def _need(self):
        if not self._ctx:
            raise GpuError("backend not initialized: call init()")

    def axpy(self, a: ComputeBuffer, b: ComputeBuffer, scale: float, out: ComputeBuffer) -> None:
        self._need()
        _check(self._L.vh_backend_axpy(self._ctx, nat.ptr(a.data, C.c_float), len(a), nat.ptr(b.data, C.c_float),
                                       len(b), float(scale), nat.ptr(out.data, C.c_float), len(out)),
               self._ctx, "axpy", err_cls=GpuError)

    def dot(self, a: ComputeBuffer, b: ComputeBuffer) -> float:
        self._need()
        r = C.c_double(0.0)
        _check(self._L.vh_backend_dot(self._ctx, nat.ptr(a.data, C.c_float), len(a), nat.ptr(b.data, C.c_float),
                                      len(b), C.byref(r)), self._ctx, "dot", err_cls=GpuError)
        return r.value

    def fill(self, buf: ComputeBuffer, value: float) -> None:
        self._need()
        _check(self._L.vh_backend_fill(self._ctx, nat.ptr(buf.data, C.c_float), len(buf), float(value)),
               self._ctx, "fill", err_cls=GpuError)

    def close(self) -> None:
        if self._ctx:
            self._L.vh_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
