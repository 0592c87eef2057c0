"""Pins the oracle's other membrane models and the Discrete Shells bending
(oracle/cpu_ref.cpp: target_of, discrete_shells) on the CPU.

Known answers from the reference's tests
(crates/vistio-material/tests/material_tests.rs:180-402,
crates/vistio-solver/tests/solver_tests.rs:783-944); the rest is fixed by
mathematics: closed-form targets for diagonal deformation gradients, the
models' limits (anisotropic with k = 1 is the co-rotational model,
orthotropic with equal stiffness is ARAP), and the Discrete Shells stencil
as the gradient of the dihedral angle (finite differences)."""
import numpy as np
import pytest

from oracle import mesh, ref

I32 = np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])  # Mat3x2::IDENTITY, columns as rows


def F(c0, c1):
    return np.concatenate([np.asarray(c0, float), np.asarray(c1, float)])


def energy(Fm, model, warp=0.0, weft=0.0):
    T = ref.target_f(Fm, model, warp=warp, weft=weft)
    return 0.5 * float(((Fm.reshape(2, 3) - T) ** 2).sum())


X, Y = np.array([1.0, 0, 0]), np.array([0.0, 1, 0])


def test_isotropic_linear_kats():
    assert np.allclose(ref.target_f(F(X, Y), ref.MODEL_ISO), I32)           # material_tests.rs:180-188
    assert energy(F(X, Y), ref.MODEL_ISO) < 1e-8                            # :190-196
    assert energy(F(1.5 * X, 1.5 * Y), ref.MODEL_ISO) > 0                   # :198-204
    assert energy(F(Y, -X), ref.MODEL_ISO) > 0.1                            # :206-216 (rotation is strain)
    rng = np.random.default_rng(0)
    for _ in range(20):
        assert np.array_equal(ref.target_f(rng.normal(size=6), ref.MODEL_ISO), I32)


def test_orthotropic_kats_and_closed_forms():
    assert energy(F(X, Y), ref.MODEL_ORTHO, 1.0, 1.0) < 1e-6                # :291-300
    e_warp = energy(F(2 * X, Y), ref.MODEL_ORTHO, 10.0, 1.0)                # :303-322
    e_weft = energy(F(X, 2 * Y), ref.MODEL_ORTHO, 10.0, 1.0)
    assert e_weft < e_warp or abs(e_warp - e_weft) < e_warp * 0.5
    assert energy(F(1.5 * X, 1.5 * Y), ref.MODEL_ORTHO, 5.0, 5.0) > 0       # :324-335
    # closed forms: F = diag(2, 1) has sigma = (2, 1) along (X, Y); k = (1, 0.1)
    assert np.allclose(ref.target_f(F(2 * X, Y), ref.MODEL_ORTHO, warp=10.0, weft=1.0), I32)
    assert e_warp == pytest.approx(0.5)
    # sigma = (2, 1) along (Y, X) for F = diag(1, 2): sigma_0 = 2 takes k_warp
    assert np.allclose(ref.target_f(F(X, 2 * Y), ref.MODEL_ORTHO, warp=10.0, weft=1.0), I32)
    # equal stiffness restores both singular values to 1: the ARAP target R
    rng = np.random.default_rng(1)
    for _ in range(50):
        Fm = rng.normal(size=6)
        assert np.allclose(ref.target_f(Fm, ref.MODEL_ORTHO, warp=3.0, weft=3.0), ref.target_f(Fm, 0), atol=1e-12)


def test_anisotropic_kats_and_closed_forms():
    assert energy(F(X, Y), ref.MODEL_ANISO, 1.0, 0.5) < 1e-6                # :346-352
    assert energy(F(0.5 * X, 0.5 * Y), ref.MODEL_ANISO, 1.0, 0.5) < 1e-6    # :355-364
    e_warp = energy(F(2 * X, Y), ref.MODEL_ANISO, 0.9, 0.1)                 # :366-385
    e_weft = energy(F(X, 2 * Y), ref.MODEL_ANISO, 0.9, 0.1)
    assert e_warp > e_weft
    # closed forms: sigma_0 = 2 along X -> k0 = 0.9 -> target 1 + 1 * 0.1
    assert np.allclose(ref.target_f(F(2 * X, Y), ref.MODEL_ANISO, warp=0.9, weft=0.1), [[1.1, 0, 0], [0, 1, 0]])
    assert e_warp == pytest.approx(0.5 * 0.81)
    assert np.allclose(ref.target_f(F(X, 2 * Y), ref.MODEL_ANISO, warp=0.9, weft=0.1), [[1, 0, 0], [0, 1.9, 0]])
    # warp = weft = 1 is the co-rotational model (:387-402); stiffness is clamped to [0, 1]
    rng = np.random.default_rng(2)
    for _ in range(50):
        Fm = rng.normal(size=6)
        tc = ref.target_f(Fm, 1)
        assert np.allclose(ref.target_f(Fm, ref.MODEL_ANISO, warp=1.0, weft=1.0), tc, atol=1e-12)
        assert np.allclose(ref.target_f(Fm, ref.MODEL_ANISO, warp=100.0, weft=7.0), tc, atol=1e-12)
        # warp = weft = 0: no restoration, the target is F itself (rotation times stretch)
        assert np.allclose(ref.target_f(Fm, ref.MODEL_ANISO, warp=0.0, weft=0.0), Fm.reshape(2, 3), atol=1e-9)


def _ds_oracle(c, rng=None, amp=0.0, **kw):
    px, py, pz, idx = mesh.quad_grid(c, c, 1.0, 1.0)
    if rng is not None:
        pz = (pz + rng.normal(size=px.size) * amp).astype(np.float32)
    hg = mesh.hinges4(idx, px.size)
    pin = np.zeros(px.size, dtype=np.uint8)
    o = ref.Oracle(ref.F64, px, py, pz, idx, hg, pin, ref.params(bending_kind=ref.BENDING_DISCRETE_SHELLS, **kw))
    return o, (px, py, pz), hg


def test_discrete_shells_kats():
    o, _, hg = _ds_oracle(4, k_bending=1.0)
    st = o.hinge_stencils()
    assert st.shape[0] == hg.shape[0]                                       # solver_tests.rs:783-789
    o, _, _ = _ds_oracle(5, k_bending=1.0)
    rest, w = o.hinges()
    assert np.all(rest == np.pi)                                            # :791-805
    o, _, _ = _ds_oracle(3, k_bending=1.0)
    st = o.hinge_stencils()
    _, w = o.hinges()
    assert np.all(np.isfinite(st)) and np.all(w > 0)                        # :807-822
    assert np.allclose((st ** 2).sum(1), 1.0)                               # unit norm
    assert np.allclose(st.sum(1), 0.0, atol=1e-12)                          # translation invariance
    # denim (bending 0.85/0.75) weighs more than silk (0.10/0.08), :870-890
    od, _, _ = _ds_oracle(4, k_bending=(0.85 + 0.75) / 2 * 100)
    osk, _, _ = _ds_oracle(4, k_bending=(0.10 + 0.08) / 2 * 100)
    assert od.hinges()[1].mean() > osk.hinges()[1].mean()


def _dihedral_np(p):
    e = p[1] - p[0]
    ax = e / np.linalg.norm(e)
    qa = p[2] - p[0]; qa = qa - ax * (qa @ ax)
    qb = p[3] - p[0]; qb = qb - ax * (qb @ ax)
    return np.arctan2(np.cross(qa, qb) @ ax, qa @ qb)


def _grad_mag(p):
    g = np.zeros(4)
    for k in range(4):
        for a in range(3):
            d = np.zeros((4, 3)); d[k, a] = 1e-7
            g[k] += ((_dihedral_np(p + d) - _dihedral_np(p - d)) / 2e-7) ** 2
    return np.sqrt(g)


def test_discrete_shells_stencil_is_the_dihedral_gradient():
    """The wing coefficients |e| / (2 A) are the magnitudes of the dihedral
    angle's gradient at the wings (Bridson et al.); the edge-vertex ones add
    the two wings' terms as scalars, which is the gradient's magnitude when
    the two normals are parallel (a flat hinge). Finite differences check
    the wing ratio on a crumpled sheet and the whole direction on a flat one."""
    rng = np.random.default_rng(4)
    o, (px, py, pz), hg = _ds_oracle(6, rng, 0.05, k_bending=1.0)
    P = np.stack([px, py, pz], 1).astype(np.float64)
    st = o.hinge_stencils()
    for e in range(0, hg.shape[0], 7):
        g = _grad_mag(P[hg[e]])
        assert st[e, 2] / st[e, 3] == pytest.approx(g[2] / g[3], rel=1e-5)
    o, (px, py, pz), hg = _ds_oracle(6, rng, 1e-5, k_bending=1.0)
    P = np.stack([px, py, pz], 1).astype(np.float64)
    st = o.hinge_stencils()
    for e in range(0, hg.shape[0], 7):
        g = _grad_mag(P[hg[e]])
        assert np.allclose(np.abs(st[e]), g / np.linalg.norm(g), atol=1e-3)


def test_discrete_shells_anisotropic_modulation_and_matrix():
    """from_topology_with_anisotropic_material: in the XZ plane, edges along X
    take k_weft / k_avg, along Z k_warp / k_avg, diagonals the mean; the
    system matrix keeps rows summing to M/h^2 (stencils sum to zero)."""
    px, py, pz, idx = mesh.quad_grid(4, 4, 1.0, 1.0)
    px, py, pz = px, np.zeros_like(px), py.copy()                           # lay it in the XZ plane
    hg = mesh.hinges4(idx, px.size)
    pin = np.zeros(px.size, dtype=np.uint8)
    kw, kf = 0.9, 0.3
    base = ref.Oracle(ref.F64, px, py, pz, idx, hg, pin, ref.params(bending_kind=2, k_bending=60.0))
    mod = ref.Oracle(ref.F64, px, py, pz, idx, hg, pin, ref.params(bending_kind=2, k_bending=60.0,
                                                                      bend_warp=kw, bend_weft=kf))
    ratio = mod.hinges()[1] / base.hinges()[1]
    P = np.stack([px, py, pz], 1).astype(np.float64)
    for e in range(hg.shape[0]):
        d = P[hg[e, 1]] - P[hg[e, 0]]
        d = d / np.linalg.norm(d)
        c2 = d[0] ** 2 / (d[0] ** 2 + d[2] ** 2)
        assert ratio[e] == pytest.approx((kf * c2 + kw * (1 - c2)) / ((kw + kf) / 2), rel=1e-12)
    A = mod.matrix()
    m = mod.masses()
    assert np.allclose(A @ np.ones(px.size), m * 3600.0, rtol=1e-9)
