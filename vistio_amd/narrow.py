"""NarrowPhase (crates/vistio-contact/src/narrow.rs:21-35) on the MI355X.

HipVertexTriangleTest is VertexTriangleTest (vertex_triangle.rs:17-82) through
vh_narrow_vertex_triangle: for each broad-phase candidate (a = vertex vi,
b = vertex vj) it tests vi against every triangle that contains vj and not
vi, in the reference's f32 arithmetic, and returns the contacts in the
reference's order. A candidate is a pair of vertex indices (CandidatePair,
broad.rs:10-18); a contact is a ContactPair (contact.rs:22-44).
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as nat
from .solver import _check

# One ContactPair per record; contact_type 0 is ContactType::VertexTriangle.
CONTACT_DTYPE = np.dtype([("indices", np.uint32, 4), ("distance", np.float32), ("normal", np.float32, 3),
                          ("barycentric", np.float32, 3), ("contact_type", np.int32), ("is_self", np.int32)])
assert CONTACT_DTYPE.itemsize == C.sizeof(nat.ContactPairC)


@dataclass
class CandidatePair:
    """broad.rs:10-18: two primitive indices (vertices here) and whether both
    belong to the same mesh."""
    a: int
    b: int
    is_self: bool = True


class HipVertexTriangleTest:
    """The trait's HIP implementation of VertexTriangleTest."""

    def __init__(self, device: int = 0):
        self._L = nat.load()
        self._ctx = C.c_void_p()
        _check(self._L.vh_create(device, C.byref(self._ctx)), what="vh_create")

    def name(self) -> str:
        return "vertex_triangle_test"

    def detect(self, candidates, positions, indices, thickness: float) -> np.ndarray:
        """candidates: CandidatePair list or (k, 2) vertex indices; positions:
        (3, n) f32 (SimulationState's pos_x, pos_y, pos_z); indices: the
        mesh's triangle indices (3t). Returns a CONTACT_DTYPE array."""
        if len(candidates) and isinstance(candidates[0], CandidatePair):
            cand = np.array([(c.a, c.b) for c in candidates], dtype=np.uint32)
        else:
            cand = np.asarray(candidates, dtype=np.uint32).reshape(-1, 2)
        ca, cb = np.ascontiguousarray(cand[:, 0]), np.ascontiguousarray(cand[:, 1])
        pos = np.ascontiguousarray(np.asarray(positions, dtype=np.float32).reshape(3, -1))
        px, py, pz = (np.ascontiguousarray(pos[k]) for k in range(3))
        tri = np.ascontiguousarray(np.asarray(indices, dtype=np.uint32).ravel())
        n, t = px.size, tri.size // 3
        cnt = C.c_uint64(0)
        args = [self._ctx, n, nat.ptr(px, C.c_float), nat.ptr(py, C.c_float), nat.ptr(pz, C.c_float), t,
                nat.ptr(tri, C.c_uint32), ca.size, nat.ptr(ca, C.c_uint32), nat.ptr(cb, C.c_uint32),
                float(thickness)]
        out = np.zeros(max(4 * ca.size, 16), dtype=CONTACT_DTYPE)
        _check(self._L.vh_narrow_vertex_triangle(*args, out.ctypes.data_as(C.POINTER(nat.ContactPairC)),
                                                  out.size, C.byref(cnt)), self._ctx, "detect")
        k = int(cnt.value)
        if k > out.size:  # more contacts than the first guess: ask again with room for all
            out = np.zeros(k, dtype=CONTACT_DTYPE)
            _check(self._L.vh_narrow_vertex_triangle(*args, out.ctypes.data_as(C.POINTER(nat.ContactPairC)),
                                                      out.size, C.byref(cnt)), self._ctx, "detect")
        return out[:k]

    def close(self) -> None:
        if self._ctx:
            self._L.vh_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
