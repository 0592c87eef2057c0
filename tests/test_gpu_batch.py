"""Several sheets stepped as one system (block-diagonal matrix, one
dissection tree per sheet, one launch sequence): each sheet's positions
follow its own oracle run with its own material."""
import numpy as np
import pytest

from conftest import DT, max_vertex_err

pytestmark = pytest.mark.gpu


def _batch():
    from vistio_amd.batch import SheetSpec, build_batch
    specs = [SheetSpec(8, 0.15, 1e3, 1e-4), SheetSpec(12, 0.30, 4e2, 8e-4), SheetSpec(5, 0.09, 2.5e3, 1e-5),
             SheetSpec(16, 0.22, 1.5e3, 3e-4), SheetSpec(3, 0.40, 7e2, 5e-5)]
    return build_batch(specs)


def _oracle_sheet(b, s):
    from oracle import ref
    sl = b.sheet_slice(s)
    v0 = b.vert_off[s]
    t0, t1 = b.tri_off[s], b.tri_off[s + 1]
    h0, h1 = b.hinge_off[s], b.hinge_off[s + 1]
    idx = (b.mesh.indices.reshape(-1, 3)[t0:t1] - v0).astype(np.uint32).ravel()
    hg = (b.topology.interior_edges[h0:h1, :4].astype(np.int64) - v0).astype(np.uint32)
    sp = b.specs[s]
    return ref.Oracle(ref.F64, b.mesh.pos_x[sl], b.mesh.pos_y[sl], b.mesh.pos_z[sl], idx, hg, b.pinned[sl],
                      ref.params(k_membrane=sp.k_membrane, k_bending=sp.k_bending, density=sp.density))


@pytest.mark.parametrize("precision,tol", [("f64", 1e-9), ("f32_ir", 1e-6)])
def test_batch_matches_per_sheet_oracles(precision, tol):
    import vistio_amd as va
    b = _batch()
    cfg = va.SolverConfig(max_iterations=10, tolerance=0.0, gravity=(0.0, -9.81, 0.0), damping=0.0,
                          rayleigh_mass_damping=0.01, precision=precision, ir_sweeps=1, ir_final_sweeps=2)
    g = va.HipPdSolver()
    g.init_raw(b.mesh, b.topology, cfg, b.pinned, "co_rotational", b.k_membrane, b.k_bending, b.density)
    info = g.info()
    print(f"{precision}: {info['n']} verts in {len(b.specs)} sheets, {info['levels']} bands", flush=True)
    oracles = [_oracle_sheet(b, s) for s in range(len(b.specs))]
    # lumped masses per sheet equal each sheet's own init
    for s, o in enumerate(oracles):
        assert np.allclose(g.lumped_masses()[b.sheet_slice(s)], o.masses(), rtol=1e-6)
    for steps in (1, 30):
        g.step_resident(DT, steps)
        x = g.positions()
        for s, o in enumerate(oracles):
            o.step(DT, steps)
            err = max_vertex_err(x[:, b.sheet_slice(s)], o.positions())
            print(f"  sheet {s} ({b.specs[s].cells}^2) after +{steps}: {err:.2e} m", flush=True)
            assert err <= tol
    g.close()


@pytest.mark.parametrize("precision,tol", [("f64", 1e-9), ("f32_ir", 1e-6)])
def test_batch_with_bands_over_4096_tasks(precision, tol):
    """Eight 128 x 128 sheets: the setup cuts forward band 1 into 5304 tasks
    and backward bands 2 and 3 into 5080 and 4304, above the 4096 beyond
    which a band launch gives up the one contiguous run of tasks per XCD
    and runs task b on block b. Every sheet still follows its own oracle."""
    import vistio_amd as va
    from vistio_amd.batch import SheetSpec, build_batch
    b = build_batch([SheetSpec(128, 0.15, 1e3, 1e-4) if s % 2 == 0 else SheetSpec(128, 0.3, 4e2, 8e-4)
                     for s in range(8)])
    cfg = va.SolverConfig(max_iterations=10, tolerance=0.0, gravity=(0.0, -9.81, 0.0), damping=0.0,
                          rayleigh_mass_damping=0.01, precision=precision, ir_sweeps=1, ir_final_sweeps=1)
    g = va.HipPdSolver()
    g.init_raw(b.mesh, b.topology, cfg, b.pinned, "co_rotational", b.k_membrane, b.k_bending, b.density)
    info = g.info()
    print(f"{precision}: {info['n']} verts, {info['supernodes']} supernodes, {info['levels']} bands", flush=True)
    g.step_resident(DT, 8)
    x = g.positions()
    for s in range(len(b.specs)):
        o = _oracle_sheet(b, s)
        o.step(DT, 8)
        err = max_vertex_err(x[:, b.sheet_slice(s)], o.positions())
        print(f"  sheet {s}: {err:.2e} m", flush=True)
        assert err <= tol
    g.close()


def test_autotune_picks_a_candidate_that_keeps_parity():
    import vistio_amd as va
    from vistio_amd.autotune import DEFAULTS, KNOBS, autotune
    from conftest import sheet
    from oracle import ref
    px, py, pz, idx, hg, pin = sheet(20, rotate_xz=True)
    m = va.TriangleMesh(px, py, pz, idx)
    topo = va.Topology(np.concatenate([hg, np.zeros((hg.shape[0], 2), np.uint32)], 1))
    cfg = va.SolverConfig(max_iterations=10, tolerance=0.0, gravity=(0.0, -9.81, 0.0), damping=0.0,
                          rayleigh_mass_damping=0.01, precision="f64")
    best, table = autotune(m, topo, cfg, pin, "co_rotational", 1e3, 1e-4, 0.15, DT, steps=3)
    print(len(table), "candidates;", {k: getattr(best, k) for k in DEFAULTS}, flush=True)
    assert len(table) >= 4 and all("steps_per_s" in r or "error" in r for r in table)
    for name, values in KNOBS:
        assert getattr(best, name) in values
    g = va.HipPdSolver()
    g.init_raw(m, topo, best, pin, "co_rotational", 1e3, 1e-4, 0.15)
    o = ref.Oracle(ref.F64, px, py, pz, idx, hg, pin, ref.params())
    g.step_resident(DT, 30)
    o.step(DT, 30)
    assert max_vertex_err(g.positions(), o.positions()) <= 1e-9
    g.close()
