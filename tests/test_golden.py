"""Golden fixtures (tests/golden/, written by tests/golden/make_golden.py from
the oracle): the oracle must keep reproducing them (CPU), they satisfy the
invariants the physics fixes, and the HIP path matches them (GPU) without
running the oracle."""
import os

import numpy as np
import pytest

from conftest import DT, max_vertex_err

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(GOLD, name))


def _case(d, name):
    pos = d[f"{name}_pos"]
    return pos[0], pos[1], pos[2], d[f"{name}_tri"], d[f"{name}_hinges"], d[f"{name}_pinned"]


# ---- CPU: the oracle against its fixtures ------------------------------------

def test_oracle_reproduces_golden_steps():
    from oracle import ref
    d = _load("sheet20_steps.npz")
    for name in ("curtain", "rotated"):
        px, py, pz, idx, hg, pin = _case(d, name)
        for tag, mode in (("f64", ref.F64), ("f32ref", ref.F32_REF)):
            o = ref.Oracle(mode, px, py, pz, idx, hg, pin, ref.params())
            done = 0
            for m in d[f"{name}_marks"]:
                o.step(DT, int(m) - done)
                done = int(m)
                assert np.array_equal(o.positions(), d[f"{name}_{tag}_x{m}"]), (name, tag, m)


def test_golden_steps_invariants():
    d = _load("sheet20_steps.npz")
    for name in ("curtain", "rotated"):
        px, py, pz, idx, hg, pin = _case(d, name)
        rest = np.stack([px, py, pz]).astype(np.float64)
        last = int(d[f"{name}_marks"][-1])
        for tag in ("f64", "f32ref"):
            x = d[f"{name}_{tag}_x{last}"]
            assert np.array_equal(x[:, pin == 1], rest[:, pin == 1])       # pinned vertices never move
            assert x[1, pin == 0].mean() < rest[1, pin == 0].mean()         # the sheet sags
        if name == "curtain":  # the in-plane curtain stays in z = 0 up to rounding of the solve
            assert np.abs(d[f"{name}_f64_x{last}"][2]).max() < 1e-12
        # the reference's precision split drifts from the FP64 algorithm
        drift = max_vertex_err(d[f"{name}_f64_x{last}"], d[f"{name}_f32ref_x{last}"])
        assert 0.0 < drift < 1e-2


def test_oracle_reproduces_golden_local_step():
    from oracle import ref
    d = _load("sheet20_local.npz")
    s = _load("sheet20_steps.npz")
    px, py, pz, idx, hg, pin = _case(s, "rotated")
    o = ref.Oracle(ref.F64, px, py, pz, idx, hg, pin, ref.params())
    o.set(0, d["x"])
    o.set(3, d["pred"])
    rhs, tt, ht = o.local_rhs()
    assert np.array_equal(rhs, d["rhs"]) and np.array_equal(tt, d["tri_targets"])
    assert np.array_equal(ht, d["hinge_targets"])
    # targets are rigid-ish images of the current elements: hinge edge
    # vertices are unchanged (bending.rs:182-227) and the RHS sums the
    # inertia term and the element terms, whose forces sum to zero
    x = d["x"]
    assert np.allclose(ht[:, 0, :], x[:, hg[:, 0]].T) and np.allclose(ht[:, 1, :], x[:, hg[:, 1]].T)
    m = o.masses()
    inertia = (m / DT ** 2) * d["pred"]
    assert np.allclose((d["rhs"] - inertia).sum(1), 0.0, atol=1e-8 * np.abs(d["rhs"]).max())


def test_oracle_reproduces_golden_pairs_and_ccd():
    from oracle import ref
    p = _load("folded_pairs.npz")
    pos, tri = p["pos"], p["tri"]
    for k, m in enumerate(p["margins"]):
        got = ref.tri_pairs(pos[0], pos[1], pos[2], tri, m)
        assert np.array_equal(got, p[f"pairs{k}"])
    c = _load("ccd_table.npz")
    q = ref.ipc_params(self_collision=True, thickness=0.5)
    for case, toi, alpha in zip(c["cases"], c["toi"], c["alpha"]):
        P0, P1 = case
        t = ref.vt_ccd(P0[3], P1[3], P0[0], P1[0], P0[1], P1[1], P0[2], P1[2])
        assert (2.0 if t is None else t) == toi
        assert ref.ccd_step(P0.T, P1.T, c["tri"], q) == alpha
    assert 10 < int((c["toi"] < 2.0).sum()) < 54


# ---- GPU: the HIP path against the fixtures ----------------------------------

def _gpu(px, py, pz, idx, hg, pin, precision):
    from conftest import gpu_solver
    return gpu_solver(px, py, pz, idx, hg, pin, precision=precision, ir_sweeps=1, ir_final_sweeps=2)


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [("f64", 1e-9), ("f32_ir", 1e-6)])
def test_gpu_matches_golden_steps(precision, tol):
    d = _load("sheet20_steps.npz")
    for name in ("curtain", "rotated"):
        px, py, pz, idx, hg, pin = _case(d, name)
        g = _gpu(px, py, pz, idx, hg, pin, precision)
        done = 0
        for m in d[f"{name}_marks"]:
            g.step_resident(DT, int(m) - done)
            done = int(m)
            err = max_vertex_err(g.positions(), d[f"{name}_f64_x{m}"])
            print(f"{name} {precision} step {m}: {err:.2e} m", flush=True)
            assert err <= tol
        g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [("f64", 1e-11), ("f32_ir", 2e-6)])
def test_gpu_matches_golden_residual(precision, tol):
    d = _load("sheet20_local.npz")
    s = _load("sheet20_steps.npz")
    px, py, pz, idx, hg, pin = _case(s, "rotated")
    g = _gpu(px, py, pz, idx, hg, pin, precision)
    g.set_state(0, d["x"])
    g.set_state(3, d["pred"])
    rel = np.abs(g.residual(DT) - d["residual"]).max() / np.abs(d["rhs"]).max()
    print(f"{precision} residual rel err {rel:.2e}", flush=True)
    assert rel < tol
    g.close()


@pytest.mark.gpu
def test_gpu_matches_golden_pairs_and_ccd():
    import vistio_amd as va
    from oracle import mesh
    p = _load("folded_pairs.npz")
    pos, tri = p["pos"], p["tri"]
    hg = mesh.hinges4(tri, pos.shape[1])
    pin = np.zeros(pos.shape[1], dtype=np.uint8)
    g = _gpu(pos[0], pos[1], pos[2], tri, hg, pin, "f64")
    g.init_ipc(va.IpcConfig(self_collision=True, barrier_d_hat=1e-3, thickness=1.5e-3, compliant_contact=False))
    for k, m in enumerate(p["margins"]):
        got = g.bvh_pairs(pos.astype(np.float64), float(m))
        assert np.array_equal(got.astype(np.int64), p[f"pairs{k}"].astype(np.int64)), float(m)
    g.close()
    c = _load("ccd_table.npz")
    ctri = c["tri"]
    nohinge = np.zeros((0, 4), dtype=np.uint32)
    for k, (case, alpha) in enumerate(zip(c["cases"], c["alpha"])):
        P0, P1 = case
        q = P0.T.astype(np.float32)
        s = _gpu(q[0], q[1], q[2], ctri, nohinge, np.zeros(6, dtype=np.uint8), "f64")
        s.init_ipc(va.IpcConfig(self_collision=True, barrier_d_hat=1e-3, thickness=0.5, compliant_contact=False))
        got = s.compute_ccd_step(P0.T, P1.T)
        s.close()
        assert got == pytest.approx(alpha, rel=1e-12, abs=1e-15), k
