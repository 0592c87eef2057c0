"""GPU parity: the HIP path through the C ABI against the CPU oracle."""
import numpy as np
import pytest

from conftest import DT, gpu_solver, max_vertex_err, oracle_solver, sheet

pytestmark = pytest.mark.gpu


def _perturbed_state(px, py, pz, seed, amp=2e-3):
    rng = np.random.default_rng(seed)
    x = np.stack([px, py, pz]).astype(np.float64)
    x = x + amp * rng.standard_normal(x.shape) * np.array([[1.0], [1.0], [0.5]])
    pred = x + 1e-3 * rng.standard_normal(x.shape)
    return x, pred


@pytest.mark.parametrize("c", [8, 20, 64])
def test_solve_f64_matches_oracle(c):
    px, py, pz, idx, hg, pin = sheet(c)
    o = oracle_solver(px, py, pz, idx, hg, pin)
    g = gpu_solver(px, py, pz, idx, hg, pin)
    b = np.random.default_rng(c).standard_normal((3, px.size))
    xo = np.stack([o.solve(b[k]) for k in range(3)])
    xg = g.solve(b)
    rel = np.abs(xg - xo).max() / np.abs(xo).max()
    print(f"c={c} solve f64 rel err {rel:.3e}", flush=True)
    assert rel < 1e-11
    A = o.matrix()
    res = max(np.abs(A @ xg[k] - b[k]).max() for k in range(3)) / np.abs(b).max()
    assert res < 1e-11


@pytest.mark.parametrize("c", [8, 20])
def test_solve_f32_ir_converges(c):
    px, py, pz, idx, hg, pin = sheet(c)
    o = oracle_solver(px, py, pz, idx, hg, pin)
    g = gpu_solver(px, py, pz, idx, hg, pin, precision="f32_ir")
    b = np.random.default_rng(c).standard_normal((3, px.size))
    xo = np.stack([o.solve(b[k]) for k in range(3)])
    errs = []
    for sweeps in (1, 2, 3):
        xg = g.solve(b, sweeps=sweeps)
        errs.append(np.abs(xg - xo).max() / np.abs(xo).max())
    print(f"c={c} f32 IR rel err by sweeps {errs}", flush=True)
    assert errs[0] < 1e-3
    assert errs[2] < 1e-10
    assert errs[1] < errs[0] and errs[2] < errs[1]


@pytest.mark.parametrize("precision,tol", [("f64", 1e-11), ("f32_ir", 1e-6)])
def test_residual_matches_oracle(precision, tol):
    px, py, pz, idx, hg, pin = sheet(20, rotate_xz=True)
    o = oracle_solver(px, py, pz, idx, hg, pin)
    g = gpu_solver(px, py, pz, idx, hg, pin, precision=precision)
    x, pred = _perturbed_state(px, py, pz, 7)
    o.set(0, x); o.set(3, pred)
    g.set_state(0, x); g.set_state(3, pred)
    b, _, _ = o.local_rhs()
    A = o.matrix()
    r_o = np.stack([b[k] - A @ x[k] for k in range(3)])
    r_g = g.residual(DT)
    scale = np.abs(b).max()
    rel = np.abs(r_g - r_o).max() / scale
    print(f"{precision} residual rel err {rel:.3e} (|b| {scale:.3e})", flush=True)
    assert rel < tol


@pytest.mark.parametrize("c,nsteps", [(20, 120)])
def test_step_f64_matches_oracle(c, nsteps):
    px, py, pz, idx, hg, pin = sheet(c)
    o = oracle_solver(px, py, pz, idx, hg, pin)
    g = gpu_solver(px, py, pz, idx, hg, pin)
    o.step(DT, nsteps)
    r = g.step_resident(DT, nsteps)
    err = max_vertex_err(g.positions(), o.positions())
    print(f"c={c} f64 steps={nsteps} max vertex err {err:.3e} m, iters {r.iterations}", flush=True)
    assert r.iterations == 10
    assert err < 1e-9


def test_step_f64_rotated_sheet_matches_oracle():
    px, py, pz, idx, hg, pin = sheet(20, rotate_xz=True)
    o = oracle_solver(px, py, pz, idx, hg, pin)
    g = gpu_solver(px, py, pz, idx, hg, pin)
    o.step(DT, 60)
    g.step_resident(DT, 60)
    err = max_vertex_err(g.positions(), o.positions())
    print(f"rotated sheet f64 60 steps max vertex err {err:.3e} m", flush=True)
    assert err < 1e-9


@pytest.mark.parametrize("c", [20])
def test_step_f32_ir_within_1e6(c):
    px, py, pz, idx, hg, pin = sheet(c)
    o = oracle_solver(px, py, pz, idx, hg, pin)
    g = gpu_solver(px, py, pz, idx, hg, pin, precision="f32_ir", ir_sweeps=2, ir_final_sweeps=3)
    o.step(DT, 120)
    g.step_resident(DT, 120)
    err = max_vertex_err(g.positions(), o.positions())
    print(f"c={c} f32_ir 120 steps max vertex err {err:.3e} m", flush=True)
    assert err < 1e-6


@pytest.mark.parametrize("max_band_levels", [1, 2, 0])
def test_band_schedules_solve_to_oracle(max_band_levels):
    """One band per dissection level (1), pairs (2) and the cost-model choice
    (0) are different block structures of the same factor: all must solve to
    the oracle's precision."""
    px, py, pz, idx, hg, pin = sheet(64)
    o = oracle_solver(px, py, pz, idx, hg, pin)
    b = np.random.default_rng(11).standard_normal((3, px.size))
    xo = np.stack([o.solve(b[k]) for k in range(3)])
    g = gpu_solver(px, py, pz, idx, hg, pin, max_band_levels=max_band_levels)
    xg = g.solve(b)
    rel = np.abs(xg - xo).max() / np.abs(xo).max()
    print(f"max_band_levels={max_band_levels} levels={g.info()['levels']} rel err {rel:.3e}", flush=True)
    assert rel < 1e-11


def test_tolerance_early_exit_matches_oracle():
    """tolerance > 0: the device-side convergence flag stops the iteration
    at the same count as the oracle (pd_solver.rs:442-444)."""
    px, py, pz, idx, hg, pin = sheet(20)
    o = oracle_solver(px, py, pz, idx, hg, pin, tolerance=1e-6, max_iterations=30)
    g = gpu_solver(px, py, pz, idx, hg, pin, tolerance=1e-6, max_iterations=30)
    for step in range(5):
        ro = o.step(DT, 1)
        rg = g.step_resident(DT, 1)
        assert rg.iterations == ro.iterations, (step, rg.iterations, ro.iterations)
        assert abs(rg.final_residual - ro.final_residual) <= 1e-6 * ro.final_residual + 1e-15
    assert max_vertex_err(g.positions(), o.positions()) < 1e-9


def test_step_with_ground_matches_oracle():
    """SimulationState.ground_height in the plain step (state.rs:160-212,
    pd_solver.rs:437-451): the hanging sheet swings onto a ground at
    y = 0.6; resident and host-state stepping both follow the oracle."""
    import vistio_amd as va
    px, py, pz, idx, hg, pin = sheet(20, rotate_xz=True)
    gh = float(np.float32(0.6))  # SimulationState.ground_height is an f32
    o = oracle_solver(px, py, pz, idx, hg, pin, ground_height=gh)
    g = gpu_solver(px, py, pz, idx, hg, pin)
    g.set_ground_height(0.6)
    o.step(DT, 60)
    g.step_resident(DT, 60)
    err = max_vertex_err(g.positions(), o.positions())
    low = float(g.positions()[1].min())
    print(f"ground 0.6, 60 steps: max vertex err {err:.3e} m, lowest y {low:.6f}", flush=True)
    assert err < 1e-9 and 0.6 <= low < 0.61  # resting on the ground (without it the sheet swings to y < 0)
    h = gpu_solver(px, py, pz, idx, hg, pin)
    st = va.SimulationState.from_mesh_with_masses(va.TriangleMesh(px, py, pz, idx), h.lumped_masses(), pin)
    st.ground_height = 0.6
    for _ in range(60):
        h.step(st, DT)
    x = np.stack([st.pos_x, st.pos_y, st.pos_z]).astype(np.float64)
    err = max_vertex_err(x, o.positions())
    print(f"host-state stepping with ground: max vertex err {err:.3e} m (f32 state)", flush=True)
    assert err < 1e-5 and float(st.pos_y.min()) >= np.float32(0.6)
    g.close()
    h.close()
