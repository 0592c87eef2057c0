"""The C ABI boundary on the CPU: the library loads, exports every entry
point include/vistio_hip.h declares, the host-side helpers and argument
checks behave, and nothing numerical runs without a GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "vistio_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vh_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from vistio_amd import _native as nat
    L = nat.load()
    names = declared_symbols()
    assert len(names) >= 15
    missing = [s for s in names if not hasattr(L, s)]
    assert not missing, missing
    assert set(names) == set(nat.EXPORTS)


def test_null_arguments_are_rejected_without_a_gpu():
    from vistio_amd import _native as nat
    L = nat.load()
    assert L.vh_create(0, None) == 8                      # VH_E_INVALID_ARG
    assert L.vh_destroy(None) == 8
    assert L.vh_pd_step_resident(None, 1.0 / 60, 1, None) == 8
    assert L.vh_pd_solve(None, None, None, 1) == 8
    assert L.vh_last_error(None) == b"null context"


def test_create_fails_cleanly_without_gpu():
    import torch  # noqa: F401  (only to ask whether a device is visible)
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from vistio_amd import _native as nat
    ctx = C.c_void_p()
    assert nat.load().vh_create(0, C.byref(ctx)) == 6    # VH_E_GPU, no fallback
    import vistio_amd as va
    with pytest.raises(va.GpuError):
        va.HipPdSolver()


@pytest.mark.parametrize("c", [1, 3, 20, 64])
def test_mesh_helpers_match_oracle_restatement(c):
    import vistio_amd as va
    from oracle import mesh
    m = va.quad_grid(c, c, 1.0, 1.0)
    px, py, pz, idx = mesh.quad_grid(c, c, 1.0, 1.0)
    assert np.array_equal(m.pos_x, px) and np.array_equal(m.pos_y, py) and np.array_equal(m.pos_z, pz)
    assert np.array_equal(m.indices, idx)
    assert np.array_equal(va.Topology.build(m).interior_edges, mesh.interior_edges(idx, px.size))


def test_mesh_helper_rejects_bad_indices():
    from vistio_amd import _native as nat
    idx = np.array([0, 1, 7], dtype=np.uint32)
    cnt = C.c_uint32(0)
    rc = nat.load().vh_mesh_interior_edges(3, 1, nat.ptr(idx, C.c_uint32), None, 0, C.byref(cnt))
    assert rc == 5  # VH_E_INVALID_MESH


def test_state_constructor_errors():
    import vistio_amd as va
    m = va.quad_grid(2, 2, 1.0, 1.0)
    with pytest.raises(va.InvalidMesh):
        va.SimulationState.from_mesh_with_masses(m, np.ones(9), np.zeros(4))   # state.rs:92-98
    s = va.SimulationState.from_mesh_with_masses(m, np.ones(9), [1, 0, 0, 0, 0, 0, 0, 0, 0])
    assert s.inv_mass[0] == 0.0 and s.inv_mass[1] == 1.0 and s.mass[0] == np.finfo(np.float32).max


def test_fabric_properties_mapping():
    import vistio_amd as va
    p = va.FabricProperties(density=150.0, stretch_stiffness_warp=100, stretch_stiffness_weft=100,
                            bending_stiffness_warp=1e-6, bending_stiffness_weft=1e-6)
    assert p.avg_stretch_stiffness() * 10.0 == 1e3        # element.rs:106
    assert abs(p.avg_bending_stiffness() * 100.0 - 1e-4) < 1e-18  # bending.rs:160
    assert not p.is_anisotropic()
