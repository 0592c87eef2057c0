"""ctypes front end of the CPU oracle (oracle/cpu_ref.cpp).

TEST INFRASTRUCTURE ONLY. Imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of bench.py, as the checker or the CPU baseline.
The product package (vistio_amd) never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libvh_oracle.so")
_lib = None

F64 = 0      # reference algorithm re-evaluated in f64
F32_REF = 1  # reference precision split: f32 state/elements, f64 solve


class OrParams(C.Structure):
    _fields_ = [
        ("model", C.c_int),
        ("k_membrane", C.c_double),
        ("k_bending", C.c_double),
        ("density", C.c_double),
        ("dt_matrix", C.c_double),
        ("gravity", C.c_double * 3),
        ("damping", C.c_double),
        ("rayleigh", C.c_double),
        ("max_iterations", C.c_int),
        ("tolerance", C.c_double),
        ("has_ground", C.c_int),
        ("ground_height", C.c_double),
    ]


class OrStepResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("final_residual", C.c_double),
                ("converged", C.c_int), ("wall_time", C.c_double)]


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp, dp, fp, up = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_float), C.POINTER(C.c_uint32)
        L.or_create.restype = vp
        L.or_create.argtypes = [C.c_int, C.c_int, fp, fp, fp, C.c_int, up, C.c_int, up,
                                C.POINTER(C.c_uint8), C.POINTER(OrParams)]
        L.or_destroy.argtypes = [vp]
        L.or_step.argtypes = [vp, C.c_double, C.c_int, C.POINTER(OrStepResult)]
        L.or_get.argtypes = [vp, C.c_int, dp]
        L.or_set.argtypes = [vp, C.c_int, dp]
        L.or_masses.argtypes = [vp, dp]
        L.or_solve_time.restype = C.c_double
        L.or_solve_time.argtypes = [vp]
        L.or_matrix_nnz.restype = C.c_int
        L.or_matrix_nnz.argtypes = [vp]
        L.or_matrix.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), dp]
        L.or_factor_nnz.restype = C.c_longlong
        L.or_factor_nnz.argtypes = [vp]
        L.or_solve.argtypes = [vp, dp, dp]
        L.or_local_rhs.argtypes = [vp, dp, dp, dp]
        L.or_elements.argtypes = [vp, dp, dp, dp]
        L.or_hinges.argtypes = [vp, dp, dp]
        L.or_polar.argtypes = [C.c_int, dp, dp]
        L.or_target_f.argtypes = [C.c_int, C.c_int, dp, dp]
        L.or_dm_inv.argtypes = [C.c_int, dp, dp, dp]
        L.or_dihedral.restype = C.c_double
        L.or_dihedral.argtypes = [C.c_int, dp]
        L.or_bend_project.argtypes = [C.c_int, dp, C.c_double, dp]
        L.or_membrane_project.argtypes = [C.c_int, C.c_int, dp, dp, dp]
        _lib = L
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def params(model=1, k_membrane=1e3, k_bending=1e-4, density=0.15, dt_matrix=1.0 / 60.0,
           gravity=(0.0, -9.81, 0.0), damping=0.0, rayleigh=0.01, max_iterations=10,
           tolerance=0.0, ground_height=None) -> OrParams:
    p = OrParams()
    p.model = model
    p.k_membrane, p.k_bending, p.density, p.dt_matrix = k_membrane, k_bending, density, dt_matrix
    for k in range(3):
        p.gravity[k] = gravity[k]
    p.damping, p.rayleigh = damping, rayleigh
    p.max_iterations, p.tolerance = max_iterations, tolerance
    p.has_ground = 0 if ground_height is None else 1
    p.ground_height = 0.0 if ground_height is None else ground_height
    return p


class Oracle:
    """One oracle solver instance (init + step), mirroring ProjectiveDynamicsSolver."""

    def __init__(self, mode, px, py, pz, indices, hinges, pinned, p: OrParams):
        L = lib()
        self.n = int(px.size)
        self._keep = [np.ascontiguousarray(a, dtype=np.float32) for a in (px, py, pz)]
        tri = np.ascontiguousarray(indices, dtype=np.uint32)
        hg = np.ascontiguousarray(hinges, dtype=np.uint32).reshape(-1)
        pin = np.ascontiguousarray(pinned, dtype=np.uint8)
        self.t, self.h = tri.size // 3, hg.size // 4
        self.p = p
        self._h = L.or_create(mode, self.n, _p(self._keep[0], C.c_float), _p(self._keep[1], C.c_float),
                              _p(self._keep[2], C.c_float), self.t, _p(tri, C.c_uint32), self.h,
                              _p(hg, C.c_uint32), _p(pin, C.c_uint8), C.byref(p))
        if not self._h:
            raise RuntimeError("oracle init failed (Cholesky factorization failed or bad input)")

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_destroy(self._h)
            self._h = None

    def step(self, dt=1.0 / 60.0, nsteps=1) -> OrStepResult:
        r = OrStepResult()
        lib().or_step(self._h, dt, nsteps, C.byref(r))
        return r

    def get(self, which=0) -> np.ndarray:
        out = np.empty(3 * self.n)
        lib().or_get(self._h, which, _p(out, C.c_double))
        return out.reshape(3, self.n)

    def set(self, which, arr):
        a = _d(arr).reshape(-1)
        lib().or_set(self._h, which, _p(a, C.c_double))

    def positions(self):
        return self.get(0)

    def masses(self):
        out = np.empty(self.n)
        lib().or_masses(self._h, _p(out, C.c_double))
        return out

    def matrix(self):
        import scipy.sparse as sp
        nnz = lib().or_matrix_nnz(self._h)
        Ap = np.empty(self.n + 1, dtype=np.int32)
        Ai = np.empty(nnz, dtype=np.int32)
        Ax = np.empty(nnz)
        lib().or_matrix(self._h, _p(Ap, C.c_int), _p(Ai, C.c_int), _p(Ax, C.c_double))
        return sp.csr_matrix((Ax, Ai, Ap), shape=(self.n, self.n))

    def factor_nnz(self) -> int:
        return int(lib().or_factor_nnz(self._h))

    def solve(self, b):
        b = _d(b)
        x = np.empty_like(b)
        lib().or_solve(self._h, _p(b, C.c_double), _p(x, C.c_double))
        return x

    def solve_time(self) -> float:
        return float(lib().or_solve_time(self._h))

    def elements(self):
        dm, area, w = np.empty(4 * self.t), np.empty(self.t), np.empty(self.t)
        lib().or_elements(self._h, _p(dm, C.c_double), _p(area, C.c_double), _p(w, C.c_double))
        return dm.reshape(self.t, 4), area, w

    def hinges(self):
        rest, w = np.empty(self.h), np.empty(self.h)
        lib().or_hinges(self._h, _p(rest, C.c_double), _p(w, C.c_double))
        return rest, w

    def local_rhs(self):
        rhs = np.empty(3 * self.n)
        tt = np.empty(9 * self.t)
        ht = np.empty(12 * self.h)
        lib().or_local_rhs(self._h, _p(rhs, C.c_double), _p(tt, C.c_double), _p(ht, C.c_double))
        return rhs.reshape(3, self.n), tt.reshape(self.t, 3, 3), ht.reshape(self.h, 4, 3)


def polar(F, mode=F64):
    F = _d(F).reshape(6)
    out = np.empty(16)
    lib().or_polar(mode, _p(F, C.c_double), _p(out, C.c_double))
    return dict(rot=out[0:6].reshape(2, 3), stretch=out[6:10], s=out[10:12], v0=out[12:14], v1=out[14:16])


def target_f(F, model=1, mode=F64):
    F = _d(F).reshape(6)
    out = np.empty(6)
    lib().or_target_f(mode, model, _p(F, C.c_double), _p(out, C.c_double))
    return out.reshape(2, 3)


def dm_inv(e1, e2, mode=F64):
    e1, e2 = _d(e1), _d(e2)
    out = np.empty(4)
    lib().or_dm_inv(mode, _p(e1, C.c_double), _p(e2, C.c_double), _p(out, C.c_double))
    return out


def dihedral(p4, mode=F64) -> float:
    p = _d(p4).reshape(12)
    return float(lib().or_dihedral(mode, _p(p, C.c_double)))


def bend_project(p4, rest, mode=F64):
    p = _d(p4).reshape(12)
    out = np.empty(12)
    lib().or_bend_project(mode, _p(p, C.c_double), rest, _p(out, C.c_double))
    return out.reshape(4, 3)


def membrane_project(p3, dm, model=1, mode=F64):
    p = _d(p3).reshape(9)
    d = _d(dm).reshape(4)
    out = np.empty(9)
    lib().or_membrane_project(mode, model, _p(p, C.c_double), _p(d, C.c_double), _p(out, C.c_double))
    return out.reshape(3, 3)
