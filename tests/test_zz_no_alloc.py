"""The allocation-free dot product (vh_backend_dot_noalloc) through the C ABI:
its result equals vh_backend_dot's bit for bit (the same kernels, the same
block count, so the same summation order), on two calls in a row, the second
with a smaller input; sizes beyond its maximum and length mismatches are
refused with a clear error."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_ELEMS = 1 << 21


def _lib():
    from vistio_amd import _native as nat
    L = nat.load()
    fp, dp = C.POINTER(C.c_float), C.POINTER(C.c_double)
    for nm in ("vh_backend_dot", "vh_backend_dot_noalloc"):
        f = getattr(L, nm)
        f.argtypes = [C.c_void_p, fp, C.c_uint64, fp, C.c_uint64, dp]
        f.restype = C.c_int
    L.vh_create.argtypes = [C.c_int32, C.POINTER(C.c_void_p)]
    L.vh_create.restype = C.c_int
    L.vh_destroy.argtypes = [C.c_void_p]
    L.vh_last_error.argtypes = [C.c_void_p]
    L.vh_last_error.restype = C.c_char_p
    return L


def _dot(L, fn, ctx, a, b):
    out = C.c_double(-1.0)
    fp = C.POINTER(C.c_float)
    st = getattr(L, fn)(ctx, a.ctypes.data_as(fp), a.size, b.ctypes.data_as(fp), b.size, C.byref(out))
    return st, out.value


def test_noalloc_dot_matches_dot():
    L = _lib()
    ctx = C.c_void_p()
    assert L.vh_create(0, C.byref(ctx)) == 0
    rng = np.random.default_rng(2024)
    try:
        for n in (MAX_ELEMS, 300_001, 7):  # the full size, then smaller inputs over the same array
            a = rng.standard_normal(n).astype(np.float32)
            b = rng.standard_normal(n).astype(np.float32)
            st0, want = _dot(L, "vh_backend_dot", ctx, a, b)
            st1, got = _dot(L, "vh_backend_dot_noalloc", ctx, a, b)
            exact = float(np.dot(a.astype(np.float64), b.astype(np.float64)))
            print(f"n={n}: noalloc {got!r} dot {want!r} numpy-f64 {exact!r}", flush=True)
            assert st0 == 0 and st1 == 0
            assert got == want
            assert abs(got - exact) <= 1e-12 * np.abs(a.astype(np.float64) * b).sum()
        big = np.ones(MAX_ELEMS + 1, np.float32)
        st, _ = _dot(L, "vh_backend_dot_noalloc", ctx, big, big)
        assert st == 8 and b"exceed its maximum" in L.vh_last_error(ctx)
        st, _ = _dot(L, "vh_backend_dot_noalloc", ctx, big[:5], big[:6])
        assert st == 8 and b"length mismatch" in L.vh_last_error(ctx)
        st, v = _dot(L, "vh_backend_dot_noalloc", ctx, big[:0], big[:0])
        assert st == 0 and v == 0.0
    finally:
        L.vh_destroy(ctx)
