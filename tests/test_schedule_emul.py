"""CPU replay of the GPU solve schedule against the oracle's direct solve.

tests/native/setup_emul.cpp links the product's host set-up
(vistio_amd/csrc/vh_setup.cpp) and replays its forward, root and backward
task lists in FP64, with the same push-slot data flow as the HIP kernels.
This checks the nested-dissection ordering, the supernode blocks, the root
inverse, the band schedule and the slot maps without a GPU; the GPU parity
tests then check the kernels that execute the same lists.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from oracle import mesh, ref

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def replay(L, c, leaf=16, max_band_levels=0, seed=0):
    px, py, pz, idx = mesh.quad_grid(c, c, 1.0, 1.0)
    hg = np.ascontiguousarray(mesh.hinges4(idx, px.size)).ravel()
    pin = mesh.pinned_top_row(c, c)
    n = px.size
    b = np.random.default_rng(seed).standard_normal(3 * n)
    x = np.empty(3 * n)
    st = np.zeros(8, dtype=np.int64)
    rc = L.emul_solve(n, _p(px, C.c_float), _p(py, C.c_float), _p(pz, C.c_float), idx.size // 3,
                      _p(idx, C.c_uint32), hg.size // 4, _p(hg, C.c_uint32), _p(pin, C.c_uint8), C.c_double(1e3),
                      C.c_double(1e-4), C.c_double(0.15), C.c_double(1.0 / 60.0), leaf, max_band_levels,
                      _p(b, C.c_double), _p(x, C.c_double), _p(st, C.c_longlong))
    assert rc == 0, f"schedule replay failed with code {rc}"
    o = ref.Oracle(ref.F64, px, py, pz, idx, mesh.hinges4(idx, px.size), pin, ref.params())
    B = b.reshape(3, n)
    xo = np.stack([o.solve(B[k]) for k in range(3)])
    return x.reshape(3, n), xo, dict(zip(["supernodes", "bands", "nnz_factor", "nnz_a", "max_front",
                                          "fwd_tasks", "bwd_tasks", "band_starts"], st.tolist()))


@pytest.mark.parametrize("c", [1, 4, 8, 20, 48])
def test_replay_matches_direct_solve(emul, c):
    x, xo, st = replay(emul, c)
    assert np.abs(x - xo).max() <= 1e-11 * np.abs(xo).max()


@pytest.mark.parametrize("mbl", [1, 2, 3, 6])
def test_band_schedules(emul, mbl):
    """Every band grouping (one dissection level per band up to six) solves
    the same system; more levels per band means fewer bands."""
    x, xo, st = replay(emul, 40, max_band_levels=mbl)
    assert np.abs(x - xo).max() <= 1e-11 * np.abs(xo).max()
    assert st["bands"] >= 1


def test_single_band_root_only(emul):
    """A mesh small enough for one supernode runs only root tasks: no
    backward tasks remain."""
    x, xo, st = replay(emul, 4)
    assert st["supernodes"] == 1 and st["bands"] == 1 and st["bwd_tasks"] == 0
    assert np.abs(x - xo).max() <= 1e-12 * np.abs(xo).max()


@pytest.mark.parametrize("leaf", [4, 64])
def test_leaf_sizes(emul, leaf):
    x, xo, _ = replay(emul, 24, leaf=leaf)
    assert np.abs(x - xo).max() <= 1e-11 * np.abs(xo).max()


def test_device_index_limits(emul):
    """The set-up refuses a system whose block, root-inverse, front-slot or
    RHS-slot totals would overflow the kernels' 32-bit offsets, instead of
    letting the device index out of bounds."""
    f = emul.emul_index_fits
    f.argtypes = [C.c_longlong] * 4
    lim = 2 ** 31 - 1
    assert f(lim, lim, lim, lim) == 1
    assert f(0, 0, 0, 0) == 1
    for k in range(4):
        args = [1000] * 4
        args[k] = lim + 1
        assert f(*args) == 0, k


@pytest.mark.parametrize("c,leaf", [(1, 16), (7, 16), (20, 16), (64, 16), (64, 64), (130, 16)])
def test_local_slot_plan(emul, c, leaf):
    """The local step sums each block's corner contributions per vertex and
    writes one RHS slot per (block, vertex): replayed on the CPU, the plan
    sums every corner once into a slot of its own vertex, writes every slot
    once, and gives each vertex the same total as the per-corner sums. The
    sizes cover a chunk that is a single element, partial last chunks
    (20x20: 800 triangles, 1160 hinges) and many chunks."""
    px, py, pz, idx = mesh.quad_grid(c, c, 1.0, 1.0)
    hg = np.ascontiguousarray(mesh.hinges4(idx, px.size)).ravel()
    pin = mesh.pinned_top_row(c, c)
    st = np.zeros(3, dtype=np.int64)
    f = emul.emul_local_plan
    rc = f(px.size, _p(px, C.c_float), _p(py, C.c_float), _p(pz, C.c_float), idx.size // 3, _p(idx, C.c_uint32),
           hg.size // 4, _p(hg, C.c_uint32), _p(pin, C.c_uint8), leaf, _p(st, C.c_longlong))
    assert rc == 0, f"slot plan check failed with code {rc}"
    slots, corners, most = st.tolist()
    print(f"{c}x{c}: {slots} slots for {corners} corners, at most {most} vertices per block", flush=True)
    assert slots <= corners
    if c >= 64:
        assert slots * 3 < corners  # per-block sums pay off on a sheet of several blocks
