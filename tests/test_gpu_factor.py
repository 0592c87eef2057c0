"""The device factorization (vistio_amd/csrc/vh_factor.hpp, MFMA FP64) against
a host copy of the same factorization (tests/native/host_factor.cpp) run on
the same symbolic set-up: the partitioned-inverse blocks M, M^T and the
roots' dense inverses agree to 1e-12 relative (FP64) and to FP32 rounding
(FP32-IR); the factor behind them is FaerSolver::factorize
(crates/vistio-math/src/faer_solver.rs:67-93) of the PD system matrix.
The solve parity tests (test_gpu_parity.py) hold the same blocks to the
oracle's own Cholesky through the solves they drive."""
import ctypes as C

import numpy as np
import pytest

from conftest import gpu_solver, sheet

pytestmark = pytest.mark.gpu


def _p(a, t):
    return None if a is None else np.ascontiguousarray(a).ctypes.data_as(C.POINTER(t))


def host_blocks(emul, px, py, pz, idx, hg, pin, value_bytes=8, k_mem=1e3, k_bend=1e-4, density=0.15, km=None,
                kb=None, dens=None, max_band_levels=0):
    f = emul.emul_factor
    sizes = (C.c_longlong * 2)()
    args = [px.size, _p(px.astype(np.float32), C.c_float), _p(py.astype(np.float32), C.c_float),
            _p(pz.astype(np.float32), C.c_float), idx.size // 3, _p(idx.astype(np.uint32), C.c_uint32),
            hg.shape[0], _p(np.ascontiguousarray(hg, np.uint32), C.c_uint32), _p(pin.astype(np.uint8), C.c_uint8),
            C.c_double(k_mem), C.c_double(k_bend), C.c_double(density),
            _p(None if km is None else np.asarray(km, np.float64), C.c_double),
            _p(None if kb is None else np.asarray(kb, np.float64), C.c_double),
            _p(None if dens is None else np.asarray(dens, np.float64), C.c_double), value_bytes, max_band_levels]
    assert f(*args, None, None, None, sizes) == 0
    M, Mt, Sr = np.empty(sizes[0]), np.empty(sizes[0]), np.empty(sizes[1])
    assert f(*args, _p(M, C.c_double), _p(Mt, C.c_double), _p(Sr, C.c_double), sizes) == 0
    return M, Mt, Sr


def _compare(g, ref_blocks, precision):
    errs = []
    for which, want in enumerate(ref_blocks):
        got = g.blocks(which)
        assert got.shape == want.shape
        if precision == "f64":
            err = float(np.abs(got - want).max() / np.abs(want).max())
            assert err <= 1e-12, (which, err)
        else:  # the FP64 device values rounded to float: the host values rounded, within one float ulp
            err = float(np.abs(got - want.astype(np.float32).astype(np.float64)).max() / np.abs(want).max())
            assert err <= 2 ** -23, (which, err)
        errs.append(err)
    return errs


@pytest.mark.parametrize("c,precision", [(8, "f64"), (20, "f64"), (64, "f64"), (256, "f64"), (20, "f32_ir"),
                                         (256, "f32_ir")])
def test_device_factor_matches_host(emul, c, precision):
    px, py, pz, idx, hg, pin = sheet(c)
    # the device's default band depth: 3 levels in FP64, 4 in FP32-IR
    mbl = 3 if precision == "f64" else 4
    ref_blocks = host_blocks(emul, px, py, pz, idx, hg, pin, value_bytes=8 if precision == "f64" else 4,
                             max_band_levels=mbl)
    g = gpu_solver(px, py, pz, idx, hg, pin, precision=precision)
    info = g.info()
    errs = _compare(g, ref_blocks, precision)
    print(f"{c}x{c} {precision}: {info['supernodes']} supernodes, {info['levels']} bands, host set-up "
          f"{info['setup_seconds']:.3f} s, device factor {info['factor_seconds'] * 1e3:.2f} ms; "
          f"rel err M {errs[0]:.2e} Mt {errs[1]:.2e} Sroot {errs[2]:.2e}", flush=True)
    g.close()


def test_device_factor_batch_matches_host(emul):
    """Five sheets of different sizes and materials in one system (a forest
    of five dissection trees, five roots)."""
    import vistio_amd as va
    from vistio_amd.batch import SheetSpec, build_batch
    b = build_batch([SheetSpec(8, 0.15, 1e3, 1e-4), SheetSpec(12, 0.30, 4e2, 8e-4), SheetSpec(5, 0.09, 2.5e3, 1e-5),
                     SheetSpec(16, 0.22, 1.5e3, 3e-4), SheetSpec(3, 0.40, 7e2, 5e-5)])
    m = b.mesh
    ref_blocks = host_blocks(emul, m.pos_x, m.pos_y, m.pos_z, m.indices, b.topology.interior_edges[:, :4], b.pinned,
                             km=b.k_membrane, kb=b.k_bending, dens=b.density)
    cfg = va.SolverConfig(max_iterations=10, tolerance=0.0, gravity=(0.0, -9.81, 0.0), damping=0.0,
                          rayleigh_mass_damping=0.01, precision="f64")
    g = va.HipPdSolver()
    g.init_raw(m, b.topology, cfg, b.pinned, "co_rotational", b.k_membrane, b.k_bending, b.density)
    errs = _compare(g, ref_blocks, "f64")
    print(f"5-sheet batch: rel err {errs}", flush=True)
    g.close()


def test_device_factor_refuses_a_matrix_that_is_not_spd():
    """A negative membrane weight makes A = M/h^2 + sum w G^T G indefinite: the
    device factorization meets a non-positive pivot and init returns
    VH_E_NOT_SPD, which the binding raises as InvalidConfig
    ("Cholesky factorization failed", pd_solver.rs:135-137)."""
    import vistio_amd as va
    px, py, pz, idx, hg, pin = sheet(8)
    with pytest.raises(va.InvalidConfig, match="positive definite"):
        gpu_solver(px, py, pz, idx, hg, pin, k_membrane=-1e6)
