import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


DT = 1.0 / 60.0


def sheet(c, rotate_xz=False):
    """quad_grid(c, c, 1, 1) with row 0 pinned (the hanging sheets); rotate_xz
    gives the reference's Scenario::hanging_sheet pose (scenarios.rs:91-96)."""
    from oracle import mesh
    px, py, pz, idx = mesh.quad_grid(c, c, 1.0, 1.0)
    if rotate_xz:
        pz = py.copy()
        py = np.full_like(px, np.float32(1.0))
    hg = mesh.hinges4(idx, px.size)
    pin = mesh.pinned_top_row(c, c)
    return px, py, pz, idx, hg, pin


PHYS = dict(model=1, k_membrane=1e3, k_bending=1e-4, density=0.15, dt_matrix=DT,
            gravity=(0.0, -9.81, 0.0), damping=0.0, rayleigh=0.01, max_iterations=10, tolerance=0.0)


def gpu_solver(px, py, pz, idx, hg, pin, precision="f64", ir_sweeps=1, ir_final_sweeps=3, leaf=16, max_band_levels=0,
               task_entries=0, **over):
    """HIP solver with the same physics as the oracle's params(**PHYS)."""
    import vistio_amd as va
    ph = dict(PHYS)
    ph.update(over)
    m = va.TriangleMesh(px, py, pz, idx)
    topo = va.Topology(np.concatenate([hg, np.zeros((hg.shape[0], 2), np.uint32)], 1))
    cfg = va.SolverConfig(max_iterations=ph["max_iterations"], tolerance=ph["tolerance"], gravity=ph["gravity"],
                          damping=ph["damping"], rayleigh_mass_damping=ph["rayleigh"], precision=precision,
                          ir_sweeps=ir_sweeps, ir_final_sweeps=ir_final_sweeps, leaf_size=leaf,
                          max_band_levels=max_band_levels, task_entries=task_entries)
    s = va.HipPdSolver()
    s.init_raw(m, topo, cfg, pin, "co_rotational" if ph["model"] == 1 else "arap", ph["k_membrane"],
               ph["k_bending"], ph["density"])
    return s


def oracle_solver(px, py, pz, idx, hg, pin, mode=0, **over):
    from oracle import ref
    ph = dict(PHYS)
    ph.update(over)
    return ref.Oracle(mode, px, py, pz, idx, hg, pin, ref.params(**ph))


def max_vertex_err(a, b):
    return float(np.sqrt(((np.asarray(a) - np.asarray(b)) ** 2).sum(0)).max())


@pytest.fixture(scope="session")
def emul(tmp_path_factory):
    """tests/native: the product's host set-up (vistio_amd/csrc/vh_setup.cpp)
    with a CPU replay of the solve schedule and a host copy of the
    factorization, built with g++ for this test session."""
    out = tmp_path_factory.mktemp("emul") / "libemul.so"
    nat = os.path.join(ROOT, "tests", "native")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "vistio_amd", "csrc"), os.path.join(nat, "setup_emul.cpp"),
                    os.path.join(nat, "host_factor.cpp"), os.path.join(ROOT, "vistio_amd", "csrc", "vh_setup.cpp"),
                    "-o", str(out)], check=True)
    return ctypes.CDLL(str(out))
