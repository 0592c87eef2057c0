"""GPU parity of the IPC contact path against the oracle (FP64 restatement).

Broad phase: the candidate triangle pairs equal the oracle's all-pairs set
exactly. Detection and CCD: FP64 agreement to rounding. The whole
step_with_ipc: positions and velocities after several steps of a sheet
dropped on the ground, a sheet draped on a sphere and a sheet folded onto
itself.
"""
import numpy as np
import pytest

from conftest import DT, PHYS, max_vertex_err

pytestmark = pytest.mark.gpu


def _mesh(cols, rows, w, h):
    from oracle import mesh
    px, py, pz, idx = mesh.quad_grid(cols, rows, w, h)
    return px, py, pz, idx, mesh.hinges4(idx, px.size)


def _solver(px, py, pz, idx, hg, pin, ipc_cfg, precision="f64", **phys):
    import vistio_amd as va
    ph = dict(PHYS)
    ph.update(phys)
    m = va.TriangleMesh(px, py, pz, idx)
    topo = va.Topology(np.concatenate([hg, np.zeros((hg.shape[0], 2), np.uint32)], 1))
    cfg = va.SolverConfig(max_iterations=ph["max_iterations"], tolerance=ph["tolerance"], gravity=ph["gravity"],
                          damping=ph["damping"], rayleigh_mass_damping=ph["rayleigh"], precision=precision,
                          ir_sweeps=2, ir_final_sweeps=3)
    s = va.HipPdSolver()
    s.init_raw(m, topo, cfg, pin, "co_rotational", ph["k_membrane"], ph["k_bending"], ph["density"])
    s.init_ipc(ipc_cfg)
    return s


def _oracle(px, py, pz, idx, hg, pin, **phys):
    from oracle import ref
    ph = dict(PHYS)
    ph.update(phys)
    return ref.Oracle(ref.F64, px, py, pz, idx, hg, pin, ref.params(**ph))


def _cfgs(ground=None, sphere=None, self_collision=False, d_hat=1e-3, kappa=1e3, thickness=None, al=5):
    import vistio_amd as va
    from oracle import ref
    th = 1.5 * d_hat if thickness is None else thickness
    g = va.IpcConfig(ground_height=ground, sphere=sphere, self_collision=self_collision, barrier_d_hat=d_hat,
                     barrier_kappa=kappa, thickness=th, al_max_iterations=al, compliant_contact=False)
    o = ref.ipc_params(ground=ground, sphere=sphere, self_collision=self_collision, d_hat=d_hat, kappa=kappa,
                       thickness=th, al_max_iterations=al, compliant_contact=False)
    return g, o


def _f32_exact(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def _folded_sheet(c=8, gap=0.02, seed=0):
    """A 2c x c strip folded in half along x = 0.5: the upper half lies gap
    above the lower one (a connected mesh with a sharp fold), plus noise."""
    px, py, pz, idx, hg = _mesh(2 * c, c, 1.0, 0.5)
    x = px.astype(np.float64)
    y = np.zeros_like(x)
    z = py.astype(np.float64)
    up = x > 0.0
    x[up] = -x[up]
    y[up] = gap
    rng = np.random.default_rng(seed)
    y = y + rng.normal(size=y.size) * gap * 0.05
    pos = np.stack([x, y, z])
    return pos.astype(np.float32), idx, hg


def test_bvh_pairs_match_all_pairs():
    from oracle import ref
    # 64 and 100 cells: 8192 and 20000 triangles, several tiles of the
    # Morton sort (2048 keys per tile) and of the pair-offset scan. 24 cells
    # at margin 0.1: about 95 pairs per triangle, so some leaves have more
    # pairs than the single walk keeps (64) and are walked again, in the same
    # detection as leaves whose kept pairs are copied.
    for c, margin, seed in ((12, 0.01, 0), (16, 0.05, 1), (10, 0.0, 2), (64, 0.004, 3), (100, 0.002, 4),
                            (24, 0.1, 6)):
        px, py, pz, idx, hg = _mesh(c, c, 1.0, 1.0)
        rng = np.random.default_rng(seed)
        pz = (rng.normal(size=px.size) * 0.05).astype(np.float32)
        pin = np.zeros(px.size, dtype=np.uint8)
        g, _ = _cfgs(self_collision=True)
        s = _solver(px, py, pz, idx, hg, pin, g)
        pos = np.stack([px, py, pz]).astype(np.float64)
        got = s.bvh_pairs(pos, margin)
        want = ref.tri_pairs(px, py, pz, idx, margin)
        print(f"c={c} margin={margin}: {len(got)} pairs", flush=True)
        assert got.shape == want.shape and np.array_equal(got.astype(np.int64), want.astype(np.int64))
        s.close()
    pos, idx, hg = _folded_sheet()
    pin = np.zeros(pos.shape[1], dtype=np.uint8)
    g, _ = _cfgs(self_collision=True)
    s = _solver(pos[0], pos[1], pos[2], idx, hg, pin, g)
    got = s.bvh_pairs(pos.astype(np.float64), 0.03)
    want = ref.tri_pairs(pos[0], pos[1], pos[2], idx, 0.03)
    assert np.array_equal(got.astype(np.int64), want.astype(np.int64)) and len(got) > 0
    s.close()


def test_detect_matches_oracle():
    from oracle import ref
    pos, idx, hg = _folded_sheet(gap=0.015)
    n = pos.shape[1]
    pin = np.zeros(n, dtype=np.uint8)
    pos64 = pos.astype(np.float64)
    pos64[1] += 0.012  # lower layer 12 mm above the ground at y = 0
    for case in ("self", "ground", "sphere", "all"):
        kw = dict(self_collision=case in ("self", "all"), ground=0.0 if case in ("ground", "all") else None,
                  sphere=((0.0, -0.2, 0.25), 0.215) if case in ("sphere", "all") else None)
        g, o = _cfgs(d_hat=4e-4, kappa=1e3, **kw)
        s = _solver(pos[0], pos[1], pos[2], idx, hg, pin, g)
        for dh, kappa in ((4e-4, 1e3), (1e-3, 37.5)):
            got = s.detect_contacts(pos64, dh, kappa)
            want = ref.detect(pos64, idx, o, dh, kappa)
            print(f"{case} dh={dh}: active {got['active']} self {got['self_contacts']} pairs {got['pairs']}", flush=True)
            assert got["active"] == want["active"] and got["self_contacts"] == want["self_contacts"]
            assert np.array_equal(got["in_contact"], want["in_contact"])
            scale = max(1.0, np.abs(want["grad"]).max())
            assert np.abs(got["grad"] - want["grad"]).max() <= 1e-10 * scale
            assert np.abs(got["normal"] - want["normal"]).max() <= 1e-9
            hs = max(1.0, np.abs(want["hessian_diag"]).max())
            assert np.abs(got["hessian_diag"] - want["hessian_diag"]).max() <= 1e-10 * hs
            assert got["max_violation"] == pytest.approx(want["max_violation"], abs=1e-15)
        s.close()


def test_ccd_step_matches_oracle():
    from oracle import ref
    pos, idx, hg = _folded_sheet(gap=0.02)
    n = pos.shape[1]
    pin = np.zeros(n, dtype=np.uint8)
    p0 = pos.astype(np.float64)
    p0[1] += 0.03
    rng = np.random.default_rng(5)
    moves = []
    d = np.zeros_like(p0); d[1] = -0.05; moves.append(d)                       # everything through the ground
    d = np.zeros_like(p0); d[1, p0[1] > 0.04] = -0.03; moves.append(d)         # upper layer through the lower
    moves.append(rng.normal(size=p0.shape) * 0.01)
    # the sheet as a piece, 22 cm sideways with 1e-7 m of motion of its own:
    # the pair cull bounds the relative motion, so every self pair is culled
    d = np.array([[0.2], [0.0], [0.1]]) + rng.normal(size=p0.shape) * 1e-7; moves.append(d)
    # both layers 5 cm sideways, the upper one also down through the lower
    d = np.zeros_like(p0); d[0] = 0.05; d[1, p0[1] > 0.04] = -0.03; moves.append(d)
    for case in ("self", "ground", "sphere", "all"):
        kw = dict(self_collision=case in ("self", "all"), ground=0.0 if case in ("ground", "all") else None,
                  sphere=((0.0, -0.25, 0.25), 0.26) if case in ("sphere", "all") else None)
        g, o = _cfgs(**kw)
        s = _solver(pos[0], pos[1], pos[2], idx, hg, pin, g)
        for d in moves:
            got = s.compute_ccd_step(p0, p0 + d)
            want = ref.ccd_step(p0, p0 + d, idx, o)
            print(f"{case}: alpha gpu {got} oracle {want}", flush=True)
            assert got == pytest.approx(want, rel=1e-12, abs=1e-15)
        s.close()


def _run_pair(px, py, pz, idx, hg, pin, g, o, steps, **phys):
    s = _solver(px, py, pz, idx, hg, pin, g, **phys)
    orc = _oracle(px, py, pz, idx, hg, pin, **phys)
    errs = []
    for k in range(steps):
        r, st = s.step_ipc_resident(DT, 1)
        ro, sto = orc.step_ipc(o, DT)
        assert r.iterations == ro.iterations, (k, r.iterations, ro.iterations)
        assert st["al_iterations"] == sto.al_iterations, (k, st, sto.al_iterations)
        errs.append(max_vertex_err(s.positions(), orc.positions()))
    v = np.abs(s.get_state(1) - orc.get(1)).max()
    lam = np.abs(s.get_lambda() - orc.get_lambda()).max()
    s.close()
    return errs, v, lam, st


def test_step_ipc_sheet_on_ground():
    c = 8
    px, py, pz, idx, hg = _mesh(c, c, 0.5, 0.5)
    pz2, py2 = py.copy(), np.full_like(py, 0.03)
    pin = np.zeros(px.size, dtype=np.uint8)
    g, o = _cfgs(ground=0.0, d_hat=1e-3, kappa=1e2)
    errs, v, lam, st = _run_pair(px, py2, pz2, idx, hg, pin, g, o, 20)
    print("ground: max err per step", ["%.1e" % e for e in errs], "v", v, "lambda", lam, st, flush=True)
    assert max(errs) <= 1e-9 and v <= 1e-7 and lam <= 1e-6


def test_step_ipc_sphere_drape():
    c = 12
    px, py, pz, idx, hg = _mesh(c, c, 0.6, 0.6)
    y = np.full_like(py, 0.33)
    pin = np.zeros(px.size, dtype=np.uint8)
    g, o = _cfgs(sphere=((0.0, 0.0, 0.0), 0.3), ground=-0.3, d_hat=1e-3, kappa=1e2)
    errs, v, lam, st = _run_pair(px, y, py.copy(), idx, hg, pin, g, o, 15)
    print("sphere: max err per step", ["%.1e" % e for e in errs], "v", v, "lambda", lam, st, flush=True)
    assert max(errs) <= 1e-8 and v <= 1e-6


def test_step_ipc_self_fold():
    pos, idx, hg = _folded_sheet(c=6, gap=0.03)
    pin = np.zeros(pos.shape[1], dtype=np.uint8)
    y = pos[1] + np.float32(0.002)
    g, o = _cfgs(self_collision=True, ground=0.0, d_hat=4e-4, kappa=1e2, thickness=0.01)
    errs, v, lam, st = _run_pair(pos[0], y, pos[2], idx, hg, pin, g, o, 12)
    print("self fold: max err per step", ["%.1e" % e for e in errs], "v", v, "lambda", lam, st, flush=True)
    assert st["self_contacts"] >= 0
    assert max(errs) <= 1e-8 and v <= 1e-6


def test_step_with_ipc_host_state():
    """The drop-in entry point on f32 host state agrees with the resident
    step to the f32 rounding of the state; the multipliers come back."""
    import vistio_amd as va
    c = 6
    px, py, pz, idx, hg = _mesh(c, c, 0.5, 0.5)
    pz2, py2 = py.copy(), np.full_like(py, 0.02)
    pin = np.zeros(px.size, dtype=np.uint8)
    g, _ = _cfgs(ground=0.0, d_hat=1e-3, kappa=1e2)
    a = _solver(px, py2, pz2, idx, hg, pin, g)
    b = _solver(px, py2, pz2, idx, hg, pin, g)
    m = va.TriangleMesh(px, py2, pz2, idx)
    state = va.SimulationState.from_mesh_with_masses(m, a.lumped_masses(), pin)
    for _ in range(6):
        a.step_with_ipc(state, DT)
        b.step_ipc_resident(DT, 1)
    assert np.all(np.isfinite(state.pos_y)) and state.pos_y.min() > 0.0
    err = np.abs(b.positions() - np.stack([state.pos_x, state.pos_y, state.pos_z])).max()
    lam = np.abs(b.get_lambda() - np.stack([state.al_lambda_x, state.al_lambda_y, state.al_lambda_z])).max()
    print("host vs resident:", err, lam, flush=True)
    assert err < 1e-5 and lam <= 1e-3 * max(1.0, np.abs(b.get_lambda()).max())
    a.close()
    b.close()


def test_crossings_match_oracle():
    """The intersection-freeness check: two crumpled sheets through each
    other (one mesh) give the oracle's all-pairs count; one sheet gives 0."""
    from oracle import mesh, ref
    for c, amp, seed in ((8, 0.02, 0), (12, 0.05, 1)):
        px, py, pz, idx = mesh.quad_grid(c, c, 1.0, 1.0)
        rng = np.random.default_rng(seed)
        one = np.stack([px, py, pz]).astype(np.float64)
        one[2] = rng.normal(size=px.size) * amp
        two = one.copy()
        two[0] += 0.37 / c
        two[1] += 0.23 / c
        two[2] = rng.normal(size=px.size) * amp
        pos = np.concatenate([one, two], 1)
        idx2 = np.concatenate([idx, idx + px.size]).astype(np.uint32)
        hg = mesh.hinges4(idx2, pos.shape[1])
        pin = np.zeros(pos.shape[1], dtype=np.uint8)
        g, _ = _cfgs(self_collision=True)
        p32 = pos.astype(np.float32)
        s = _solver(p32[0], p32[1], p32[2], idx2, hg, pin, g)
        want = ref.crossings(pos, idx2)
        got = s.count_crossings(pos)
        print(f"c={c}: crossings gpu {got} oracle {want}", flush=True)
        assert got == want and want > 10
        assert s.count_crossings(np.concatenate([one, one + np.array([[0.0], [0.0], [0.5]])], 1)) == 0
        s.close()


def _ribbon(cols, rows):
    """A cols x rows ribbon, 0.2 m x 2.0 m, rotated 45 degrees about its short
    (x) axis, lowest point 0.5 m above the ground y = 0."""
    px, py, pz, idx, hg = _mesh(cols, rows, 0.2, 2.0)
    c = np.sqrt(0.5)
    y = py.astype(np.float64) * c
    z = py.astype(np.float64) * c
    y += 0.5 - y.min()
    return px, y.astype(np.float32), z.astype(np.float32), idx, hg


def test_ribbon_drop_parity_and_intersection_free():
    """A ribbon dropped at 45 degrees onto the ground folds onto itself:
    FP64 positions follow the oracle while the contact set builds up, and the
    FP32-refined run stays intersection-free (no crossing triangle pair) at
    the end of every step."""
    import vistio_amd as va
    from oracle import ref
    px, py, pz, idx, hg = _ribbon(4, 40)
    n = px.size
    pin = np.zeros(n, dtype=np.uint8)
    dh = 5e-4
    g = va.IpcConfig(ground_height=0.0, self_collision=True, barrier_d_hat=dh, barrier_kappa=1e4,
                     thickness=1.5 * dh, al_max_iterations=8, compliant_contact=False)
    o = ref.ipc_params(ground=0.0, self_collision=True, d_hat=dh, kappa=1e4, thickness=1.5 * dh,
                       al_max_iterations=8, compliant_contact=False)
    errs, v, lam, st = _run_pair(px, py, pz, idx, hg, pin, g, o, 75)
    # Once the ribbon lands (step ~40) the motion is chaotic: the oracle
    # itself, restarted from positions perturbed by 1e-13 m, drifts by 1e-5 m
    # within 30 steps. Before that the bar is 1e-8 m; after it, the GPU may
    # drift no more than 10x the oracle's own drift.
    a, b = _oracle(px, py, pz, idx, hg, pin), _oracle(px, py, pz, idx, hg, pin)
    b.set(0, b.get(0) + np.random.default_rng(0).normal(size=(3, n)) * 1e-13)
    drift = []
    for _ in range(75):
        a.step_ipc(o, DT)
        b.step_ipc(o, DT)
        drift.append(max_vertex_err(a.positions(), b.positions()))
    print("ribbon f64: max err every 5 steps", ["%.1e" % e for e in errs[::5]], "v", v, st, flush=True)
    print("oracle drift from 1e-13 m:       ", ["%.1e" % e for e in drift[::5]], flush=True)
    assert max(errs[:40]) <= 1e-8 and st["self_contacts"] > 0
    assert max(errs) <= 10 * max(drift)
    s = _solver(px, py, pz, idx, hg, pin, g, precision="f32_ir")
    most = 0
    for k in range(150):
        r, st = s.step_ipc_resident(DT, 1)
        most = max(most, st["self_contacts"])
        assert s.count_crossings() == 0, k
    x = s.positions()
    print("ribbon f32_ir: most self contacts", most, "final", st, "y range", x[1].min(), x[1].max(), flush=True)
    assert most > 0 and x[1].min() > -1e-3
    assert ref.crossings(x, idx) == 0
    s.close()


def test_pair_buffer_grows_on_demand():
    """A candidate-pair buffer that starts far too small (4 pairs) grows
    inside every entry point: pairs, detection, CCD, crossings and the IPC
    step give the oracle's results, as the reference's growable pair list
    does."""
    from dataclasses import replace
    from oracle import ref
    pos, idx, hg = _folded_sheet(c=6, gap=0.015)
    n = pos.shape[1]
    pin = np.zeros(n, dtype=np.uint8)
    pos64 = pos.astype(np.float64)
    pos64[1] += 0.012
    g, o = _cfgs(self_collision=True, ground=0.0, d_hat=4e-4, kappa=1e3)
    s = _solver(pos[0], pos[1], pos[2], idx, hg, pin, replace(g, max_pairs=4))
    got = s.bvh_pairs(pos.astype(np.float64), 0.03)
    want = ref.tri_pairs(pos[0], pos[1], pos[2], idx, 0.03)
    print(f"pairs: {len(got)}", flush=True)
    assert len(want) > 4 and np.array_equal(got.astype(np.int64), want.astype(np.int64))
    s.close()
    for what in ("detect", "ccd"):
        s = _solver(pos[0], pos[1], pos[2], idx, hg, pin, replace(g, max_pairs=4))
        if what == "detect":
            d = s.detect_contacts(pos64, 4e-4, 1e3)
            w = ref.detect(pos64, idx, o, 4e-4, 1e3)
            print(f"detect: pairs {d['pairs']} self {d['self_contacts']}", flush=True)
            assert d["pairs"] > 4 and d["self_contacts"] == w["self_contacts"] > 0
            assert np.abs(d["grad"] - w["grad"]).max() <= 1e-10 * max(1.0, np.abs(w["grad"]).max())
        elif what == "ccd":
            dl = np.zeros_like(pos64)
            dl[1, pos64[1] > 0.02] = -0.03
            a = s.compute_ccd_step(pos64, pos64 + dl)
            print(f"ccd: alpha {a}", flush=True)
            assert a == pytest.approx(ref.ccd_step(pos64, pos64 + dl, idx, o), rel=1e-12, abs=1e-15) and a < 1
        s.close()
    from oracle import mesh
    px, py, pz, idx1 = mesh.quad_grid(8, 8, 1.0, 1.0)
    rng = np.random.default_rng(0)
    one = np.stack([px, py, pz]).astype(np.float64)
    one[2] = rng.normal(size=px.size) * 0.02
    two = one + np.array([[0.37 / 8], [0.23 / 8], [0.0]])
    two[2] = rng.normal(size=px.size) * 0.02
    both = np.concatenate([one, two], 1)
    idx2 = np.concatenate([idx1, idx1 + px.size]).astype(np.uint32)
    p32 = both.astype(np.float32)
    s = _solver(p32[0], p32[1], p32[2], idx2, mesh.hinges4(idx2, both.shape[1]), np.zeros(both.shape[1], np.uint8),
                replace(g, max_pairs=4))
    c = s.count_crossings(both)
    print(f"crossings: {c}", flush=True)
    assert c == ref.crossings(both, idx2) > 10
    s.close()
    pos, idx, hg = _folded_sheet(c=6, gap=0.03)
    y = pos[1] + np.float32(0.002)
    g, o = _cfgs(self_collision=True, ground=0.0, d_hat=4e-4, kappa=1e2, thickness=0.01)
    errs, v, lam, st = _run_pair(pos[0], y, pos[2], idx, hg, pin, replace(g, max_pairs=4), o, 12)
    print("grown buffer self fold: max err per step", ["%.1e" % e for e in errs], "v", v, st, flush=True)
    assert st["pairs"] > 4 and max(errs) <= 1e-8 and v <= 1e-6


def test_pair_buffer_grows_at_an_al_update():
    """Growth at the detection that closes an AL iteration: a solver whose
    buffer holds exactly the step-start pairs of a step whose candidate set
    then widens takes that step, and the next ones, as a solver with room to
    spare does from the same state."""
    from dataclasses import replace
    pos, idx, hg = _folded_sheet(c=6, gap=0.03)
    pin = np.zeros(pos.shape[1], dtype=np.uint8)
    y = pos[1] + np.float32(0.002)
    g, _ = _cfgs(self_collision=True, ground=0.0, d_hat=4e-4, kappa=1e2, thickness=0.01)
    a = _solver(pos[0], y, pos[2], idx, hg, pin, g)
    found = None
    for k in range(30):
        state = a.get_state(0), a.get_state(1), a.get_lambda()
        r, st = a.step_ipc_resident(DT, 1)
        start = len(a.bvh_pairs(a.get_state(3), 0.01))  # the step-start query's positions and margin
        if st["pairs"] > start:
            found = k, state, start, st["pairs"]
            break
    a.close()
    assert found, "no step widened its candidate set"
    k, (x0, v0, l0), start, end = found
    print(f"step {k}: {start} pairs at the start, {end} at the last AL update", flush=True)
    runs = []
    for cap in (0, start):
        s = _solver(pos[0], y, pos[2], idx, hg, pin, replace(g, max_pairs=cap))
        s.set_state(0, x0)
        s.set_state(1, v0)
        s.set_lambda(l0)
        r, st = s.step_ipc_resident(DT, 1)
        first = (s.positions(), st["pairs"], r.iterations, st["al_iterations"])
        s.step_ipc_resident(DT, 5)
        runs.append((first, s.positions(), s.get_lambda()))
        s.close()
    (f0, x5a, la), (f1, x5b, lb) = runs
    print("after the step:", max_vertex_err(f0[0], f1[0]), "after 5 more:", max_vertex_err(x5a, x5b), flush=True)
    assert f1[1] == f0[1] == end and f1[2:] == f0[2:]
    assert max_vertex_err(f0[0], f1[0]) <= 1e-10 and max_vertex_err(x5a, x5b) <= 1e-8
    assert np.abs(la - lb).max() <= 1e-6 * max(1.0, np.abs(la).max())


def test_edge_edge_ccd_matches_oracle():
    """IpcConfig.ccd_edge_edge adds edge_edge_ccd (ccd.rs:138-229) on the nine
    edge pairs of every candidate triangle pair to the CCD step. Two
    triangles whose edges cross like ccd_tests.rs:54-66 (edge (a, b) sweeps
    down through edge (c, d) at t = 0.5), while no vertex ever crosses the
    other triangle's plane: vertex-triangle CCD allows the full step, the
    edge-edge tests stop it at 0.9 t = 0.45. On the folded sheet's moves
    the step fraction equals the oracle's too (same FP64 formula), to 1e-12."""
    from oracle import ref
    P = np.array([[-1.0, 0.0, 0.05], [1.0, 0.0, 0.05], [0.0, 0.0, 1.0],
                  [0.0, -1.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, -1.0]]).T.astype(np.float32)
    idx = np.arange(6, dtype=np.uint32)
    hg = np.zeros((0, 4), np.uint32)
    d = np.zeros((3, 6))
    d[2, :3] = -0.1
    pos, fidx, fhg = _folded_sheet(gap=0.02)
    fp0 = pos.astype(np.float64)
    fp0[1] += 0.03
    rng = np.random.default_rng(9)
    fmoves = [rng.normal(size=fp0.shape) * 0.01, np.where(fp0[1:2] > 0.04, [[0.0], [-0.03], [0.0]], 0.0)]
    got = {}
    for ee in (False, True):
        g, o = _cfgs(self_collision=True, thickness=0.1)
        g.ccd_edge_edge = ee
        o.ccd_edge_edge = int(ee)
        s = _solver(P[0], P[1], P[2], idx, hg, np.zeros(6, np.uint8), g)
        p0 = P.astype(np.float64)
        got[ee] = s.compute_ccd_step(p0, p0 + d)
        want = ref.ccd_step(p0, p0 + d, idx, o)
        print(f"crossing edges, edge_edge={ee}: alpha gpu {got[ee]} oracle {want}", flush=True)
        assert got[ee] == pytest.approx(want, rel=1e-12, abs=1e-15)
        s.close()
        g, o = _cfgs(self_collision=True)
        g.ccd_edge_edge = ee
        o.ccd_edge_edge = int(ee)
        s = _solver(pos[0], pos[1], pos[2], fidx, fhg, np.zeros(pos.shape[1], np.uint8), g)
        for mv in fmoves:
            a = s.compute_ccd_step(fp0, fp0 + mv)
            b = ref.ccd_step(fp0, fp0 + mv, fidx, o)
            print(f"folded sheet, edge_edge={ee}: alpha gpu {a} oracle {b}", flush=True)
            assert a == pytest.approx(b, rel=1e-12, abs=1e-15)
        s.close()
    assert got[False] == 1.0
    assert got[True] == pytest.approx(0.45, rel=1e-6)  # the corners are f32 (0.05 -> 0.0500000007)


def test_edge_edge_ccd_step_matches_oracle():
    """A contact step with edge-edge CCD on: the folded sheet falls onto
    itself and the ground for 6 steps; positions follow the FP64 oracle with
    the same option, and the AL and PD iteration counts agree."""
    pos, idx, hg = _folded_sheet(gap=0.02)
    n = pos.shape[1]
    pin = np.zeros(n, np.uint8)
    pos = pos.copy()
    pos[1] += np.float32(0.01)
    g, o = _cfgs(ground=0.0, self_collision=True, kappa=1e2, d_hat=1e-5)
    g.ccd_edge_edge = True
    o.ccd_edge_edge = 1
    s = _solver(pos[0], pos[1], pos[2], idx, hg, pin, g)
    orc = _oracle(pos[0], pos[1], pos[2], idx, hg, pin)
    for k in range(6):
        r, st = s.step_ipc_resident(DT, 1)
        ro, so = orc.step_ipc(o)
        assert (st["al_iterations"], r.iterations) == (so.al_iterations, ro.iterations), k
    err = max_vertex_err(s.positions(), orc.positions())
    print(f"edge-edge CCD step: 6 steps, max vertex err {err:.3e} m, last alpha {st['last_alpha']}", flush=True)
    assert err < 1e-8
    s.close()


def test_contact_steps_are_deterministic():
    """The self-contact gradients are summed per vertex in a fixed order (no
    floating-point atomics): two runs of the same folded-sheet contact steps
    give bitwise-identical positions, multipliers and diagnostics."""
    pos, idx, hg = _folded_sheet(gap=0.004)
    n = pos.shape[1]
    pin = np.zeros(n, np.uint8)
    runs = []
    for _ in range(2):
        g, _ = _cfgs(ground=0.0, self_collision=True, kappa=1e2, d_hat=1e-4)
        s = _solver(pos[0], pos[1], pos[2], idx, hg, pin, g, precision="f32_ir")
        sts = []
        for k in range(15):
            _, st = s.step_ipc_resident(DT, 1)
            sts.append((st["al_iterations"], st["active"], st["self_contacts"], st["pairs"]))
        runs.append((s.positions(), s.get_lambda(), sts))
        s.close()
    (x0, l0, s0), (x1, l1, s1) = runs
    print(f"self contacts per step: {[t[2] for t in s0]}", flush=True)
    assert max(t[2] for t in s0) > 0
    assert s0 == s1
    assert np.array_equal(x0, x1) and np.array_equal(l0, l1)


@pytest.mark.parametrize("precision", ["f32_ir", "f64"])
def test_contact_al_graph_replays_match_eager_launches(precision, monkeypatch):
    """The head of a contact step (begin, first detection: boxes, Morton
    sort, LBVH, pair walks and scans, self contacts) and each AL iteration
    (PD loop, detection, multiplier update) are captured once into HIP graphs
    and replayed, with host reads of the pair counts and AL flag between the
    replays. The same folded-sheet steps with graphs disabled
    (VISTIO_HIP_GRAPH=0: every kernel launched directly) give bitwise the
    same positions, multipliers and diagnostics at every step."""
    pos, idx, hg = _folded_sheet(gap=0.004)
    pin = np.zeros(pos.shape[1], np.uint8)
    runs = []
    for graph in ("1", "0"):
        monkeypatch.setenv("VISTIO_HIP_GRAPH", graph)
        g, _ = _cfgs(ground=0.0, self_collision=True, kappa=1e2, d_hat=1e-4)
        s = _solver(pos[0], pos[1], pos[2], idx, hg, pin, g, precision=precision)
        per_step = []
        for k in range(12):
            _, st = s.step_ipc_resident(DT, 1)
            per_step.append((s.positions().copy(), st["al_iterations"], st["active"], st["self_contacts"]))
        runs.append((per_step, s.get_lambda()))
        s.close()
    (a, la), (b, lb) = runs
    assert max(t[3] for t in a) > 0
    for k, (sa, sb) in enumerate(zip(a, b)):
        assert sa[1:] == sb[1:], (k, sa[1:], sb[1:])
        assert np.array_equal(sa[0], sb[0]), k
    assert np.array_equal(la, lb)
