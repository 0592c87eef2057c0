"""GPU parity of the other membrane models and of Discrete Shells bending
(init_with_material_tier3) against the oracle: the reference's hanging
sheet pose (XZ plane, bending active), 60 steps."""
import numpy as np
import pytest

from conftest import DT, max_vertex_err, sheet

pytestmark = pytest.mark.gpu

# (model name, oracle model, bending, warp, weft, bending warp, bending weft)
CASES = [
    ("co_rotational", 1, "discrete_shells", 0.0, 0.0, 0.0, 0.0),
    ("anisotropic_corotational", 4, "discrete_shells", 0.9, 0.6, 0.5, 0.3),
    ("isotropic_linear", 2, "dihedral", 0.0, 0.0, 0.0, 0.0),
    ("orthotropic_linear", 3, "dihedral", 0.98, 0.5, 0.0, 0.0),
    ("orthotropic_linear", 3, "discrete_shells", 0.5, 0.9, 0.85, 0.4),
]


@pytest.mark.parametrize("precision,tol", [("f64", 1e-9), ("f32_ir", 1e-6)])
@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}-{c[2]}-{i}" for i, c in enumerate(CASES)])
def test_models_match_oracle(case, precision, tol):
    import vistio_amd as va
    from oracle import ref
    name, omodel, bending, warp, weft, bw, bf = case
    px, py, pz, idx, hg, pin = sheet(20, rotate_xz=True)
    k_mem, k_b, dens = 9.0, 47.5, 0.2
    m = va.TriangleMesh(px, py, pz, idx)
    topo = va.Topology(np.concatenate([hg, np.zeros((hg.shape[0], 2), np.uint32)], 1))
    cfg = va.SolverConfig(max_iterations=10, tolerance=0.0, gravity=(0.0, -9.81, 0.0), damping=0.0,
                          rayleigh_mass_damping=0.01, precision=precision, ir_sweeps=1, ir_final_sweeps=2)
    g = va.HipPdSolver()
    g.init_raw(m, topo, cfg, pin, name, k_mem, k_b, dens, bending=bending, warp=warp, weft=weft,
               bend_warp=bw, bend_weft=bf)
    o = ref.Oracle(ref.F64, px, py, pz, idx, hg, pin,
                   ref.params(model=omodel, k_membrane=k_mem, k_bending=k_b, density=dens,
                              bending_kind=2 if bending == "discrete_shells" else 1, warp=warp, weft=weft,
                              bend_warp=bw, bend_weft=bf))
    for steps in (1, 20, 39):
        g.step_resident(DT, steps)
        o.step(DT, steps)
        err = max_vertex_err(g.positions(), o.positions())
        print(f"{name}/{bending} {precision} +{steps}: {err:.2e} m", flush=True)
        assert err <= tol
    # the motion is not trivial: the sheet swings and bends
    assert np.abs(o.positions() - np.stack([px, py, pz])).max() > 0.05
    g.close()


def test_tier3_init_from_material_database():
    """init_with_material_tier3 with the reference's presets
    (MaterialDatabase::with_defaults): every preset steps, and the
    anisotropic ones take the warp/weft bending modulation (their Discrete
    Shells weights differ from the isotropic init's)."""
    import vistio_amd as va
    db = va.MaterialDatabase.with_defaults()
    m = va.quad_grid(5, 5, 1.0, 1.0)
    topo = va.Topology.build(m)
    pin = np.zeros(m.vertex_count(), dtype=np.uint8)
    cfg = va.SolverConfig(max_iterations=3)
    for name in ("cotton_twill", "silk_charmeuse", "denim_14oz", "jersey_knit", "chiffon"):
        props = db.get(name)
        for model in ("co_rotational", "anisotropic_corotational"):
            s = va.HipPdSolver()
            s.init_with_material_tier3(m, topo, cfg, props, model, pin)
            st = va.SimulationState.from_mesh_with_masses(m, s.lumped_masses(), pin)
            r = s.step(st, 1.0 / 60.0)
            assert r.iterations > 0 and np.all(np.isfinite(st.pos_y))
            s.close()
