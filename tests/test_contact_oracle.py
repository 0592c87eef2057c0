"""Pins the oracle's contact path (oracle/contact_ref.hpp) on the CPU.

Known answers come from the reference's own tests
(crates/vistio-contact/src/tests/{barrier,distance_primitives,ccd,
ipc_response,bvh}_tests.rs); the rest is fixed by mathematics: finite
differences of the energies, brute-force minimisation of distances, sign
changes of the coplanarity function, closed-form ground and sphere cases and
the all-pairs AABB test.
"""
import numpy as np
import pytest

from oracle import mesh, ref

V = np.array


# ---- barrier.rs (barrier_tests.rs) -------------------------------------------

def test_barrier_kats():
    assert ref.barrier_energy(0.02, 0.01) == 0.0                     # barrier_tests.rs:6-9
    assert ref.barrier_energy(0.01, 0.01) == 0.0
    assert ref.barrier_gradient(0.02, 0.01) == 0.0
    assert ref.barrier_hessian(0.02, 0.01) == 0.0
    assert ref.barrier_energy(0.005, 0.01) > 0                        # :14-18
    assert ref.barrier_energy(0.001, 0.01) > ref.barrier_energy(0.005, 0.01)  # :22-26
    assert ref.barrier_gradient(0.005, 0.01) < 0                      # :30-35
    assert ref.barrier_hessian(0.005, 0.01) > 0                       # :39-43
    assert ref.barrier_energy(0.01 - 1e-6, 0.01) < 1e-10              # :47-51
    assert ref.barrier_energy(0.0, 0.01) == 0.0                       # :65-69
    assert ref.barrier_gradient(0.0, 0.01) < -5.0
    k1 = ref.estimate_kappa(0.0, 0.001, 9.81, 1e-3)                   # :73-80
    k2 = ref.estimate_kappa(0.0, 0.01, 9.81, 1e-3)
    assert 0 < k1 < k2 < 1e8
    assert ref.estimate_kappa(5000.0, 0.001, 9.81, 1e-3) == 5000.0    # :84-87
    th = 0.001                                                        # :91-105
    assert ref.barrier_energy_thick(th * th, 0.01, th) > 1e5
    assert ref.barrier_energy_thick(th * th + 0.01 + 0.001, 0.01, th) == 0.0
    assert ref.barrier_energy_thick(th * th + 0.005, 0.01, th) > 0
    assert ref.barrier_gradient_thick(th * th + 0.0001, 0.01, th) < 0  # :109-115


def test_barrier_derivatives_match_finite_differences():
    dh = 0.01
    for d in (0.0005, 0.002, 0.005, 0.009):
        h = 1e-7
        fd = (ref.barrier_energy(d + h, dh) - ref.barrier_energy(d - h, dh)) / (2 * h)
        assert abs(ref.barrier_gradient(d, dh) - fd) < 1e-6 * max(1.0, abs(fd))
        fd2 = (ref.barrier_gradient(d + h, dh) - ref.barrier_gradient(d - h, dh)) / (2 * h)
        assert abs(ref.barrier_hessian(d, dh) - fd2) < 1e-5 * max(1.0, abs(fd2))
    # the gradient is clamped below 1e-3 d_hat (barrier.rs:44)
    assert ref.barrier_gradient(1e-9, dh) == ref.barrier_gradient(1e-5, dh)
    # kappa closed form: m g / (2 sqrt(dh/2) |b'(dh/2)|)
    m, g, d = 0.002, 9.81, 1e-3
    want = m * g / (2 * np.sqrt(d / 2) * abs(ref.barrier_gradient(d / 2, d)))
    assert ref.estimate_kappa(0.0, m, g, d) == pytest.approx(want, rel=1e-14)


# ---- distance_primitives.rs ---------------------------------------------------

T0, T1, T2 = V([0.0, 0, 0]), V([1.0, 0, 0]), V([0.0, 1, 0])


def test_point_triangle_kats():
    d, ty = ref.pt_dist2(V([0.25, 0.25, 0.0]), T0, T1, T2)            # distance_primitives_tests.rs:8-17
    assert d < 1e-10 and ty == ref.PT_FACE
    d, ty = ref.pt_dist2(V([0.25, 0.25, 0.5]), T0, T1, T2)            # :20-29
    assert abs(d - 0.25) < 1e-12 and ty == ref.PT_FACE
    d, ty = ref.pt_dist2(V([-1.0, -1.0, 0.0]), T0, T1, T2)            # :32-41
    assert abs(d - 2.0) < 1e-12 and ty == ref.PT_VERT0
    d, ty = ref.pt_dist2(V([0.5, -0.5, 0.0]), T0, T1, T2)             # :44-53
    assert abs(d - 0.25) < 1e-12 and ty == ref.PT_EDGE0
    d, _ = ref.pt_dist2(V([0.5, 1.0, 0.0]), T0, T1, V([0.5, 0, 0]))   # :56-63 degenerate
    assert abs(d - 1.0) < 1e-12


def test_edge_edge_kats():
    d, _ = ref.ee_dist2(V([0, 0, -1.0]), V([0, 0, 1.0]), V([-1.0, 1, 0]), V([1.0, 1, 0]))  # :68-76
    assert abs(d - 1.0) < 1e-12
    d, _ = ref.ee_dist2(V([0, 0, 0.0]), V([1.0, 0, 0]), V([0, 1.0, 0]), V([1.0, 1, 0]))    # :79-87
    assert abs(d - 1.0) < 1e-12
    d, _ = ref.ee_dist2(V([0, 0, 0.0]), V([1.0, 0, 0]), V([0.5, 0, 0]), V([0.5, 1, 0]))    # :90-98
    assert d < 1e-10
    d, ty = ref.ee_dist2(V([0, 0, 0.0]), V([0, 0, 0.0]), V([1.0, 0, 0]), V([1.0, 0, 0]))   # :101-111
    assert abs(d - 1.0) < 1e-12 and ty == ref.DEGENERATE


def _brute_pt(p, a, b, c, m=400):
    u = np.linspace(0, 1, m)
    U, W = np.meshgrid(u, u)
    ok = U + W <= 1
    q = a[None, :] + np.outer(U[ok], b - a) + np.outer(W[ok], c - a)
    return ((q - p) ** 2).sum(1).min()


def test_point_triangle_matches_brute_force_and_gradient():
    rng = np.random.default_rng(1)
    for _ in range(40):
        a, b, c = rng.normal(size=(3, 3))
        p = rng.normal(size=3) * 1.5
        d, ty = ref.pt_dist2(p, a, b, c)
        bf = _brute_pt(p, a, b, c)
        assert d <= bf + 1e-12 and bf - d < 5e-4 * max(1.0, bf)
        if ty == ref.PT_FACE:  # the gradient is the exact one in the face region
            g = ref.pt_grad(p, a, b, c)
            h = 1e-6
            for k, base in enumerate((p, a, b, c)):
                for ax in range(3):
                    args = [p.copy(), a.copy(), b.copy(), c.copy()]
                    args[k][ax] += h
                    dp = ref.pt_dist2(*args)[0]
                    args[k][ax] -= 2 * h
                    dm = ref.pt_dist2(*args)[0]
                    assert abs(g[k, ax] - (dp - dm) / (2 * h)) < 1e-6


def test_edge_edge_matches_brute_force_and_gradient():
    rng = np.random.default_rng(2)
    s = np.linspace(0, 1, 801)
    for _ in range(40):
        a0, a1, b0, b1 = rng.normal(size=(4, 3))
        d, _ = ref.ee_dist2(a0, a1, b0, b1)
        P = a0[None, :] + np.outer(s, a1 - a0)
        Q = b0[None, :] + np.outer(s, b1 - b0)
        bf = ((P[:, None, :] - Q[None, :, :]) ** 2).sum(2).min()
        assert d <= bf + 1e-12 and bf - d < 1e-4 * max(1.0, bf)
    # interior closest points: gradient = FD (edge_edge_gradient_finite_difference, :135-158)
    a0, a1, b0, b1 = V([0, 0, -0.5]), V([0, 0, 0.5]), V([-0.5, 1.0, 0]), V([0.5, 1.0, 0])
    g = ref.ee_grad(a0, a1, b0, b1)
    h = 1e-6
    for k in range(4):
        for ax in range(3):
            args = [a0.copy(), a1.copy(), b0.copy(), b1.copy()]
            args[k][ax] += h
            dp = ref.ee_dist2(*args)[0]
            args[k][ax] -= 2 * h
            dm = ref.ee_dist2(*args)[0]
            assert abs(g[k, ax] - (dp - dm) / (2 * h)) < 1e-6


# ---- ccd.rs (ccd_tests.rs) ----------------------------------------------------

def test_ccd_kats():
    a, b, c = V([0.0, 0, 0]), V([1.0, 0, 0]), V([0.0, 1, 0])
    t = ref.vt_ccd(V([0.25, 0.25, 1.0]), V([0.25, 0.25, -1.0]), a, a, b, b, c, c)  # ccd_tests.rs:6-21
    assert t is not None and abs(t - 0.45) < 1e-12
    assert ref.vt_ccd(V([5.0, 5, 1]), V([5.0, 5, -1]), a, a, b, b, c, c) is None   # :24-36
    assert ref.vt_ccd(V([0.25, 0.25, 1]), V([0.25, 0.25, 2]), a, a, b, b, c, c) is None  # :39-51
    t = ref.ee_ccd(V([-1.0, 0, 0.5]), V([-1.0, 0, -0.5]), V([1.0, 0, 0.5]), V([1.0, 0, -0.5]),  # :54-66
                   V([0.0, -1, -0.5]), V([0.0, -1, 0.5]), V([0.0, 1, -0.5]), V([0.0, 1, 0.5]))
    assert t is not None
    z = V([0.0, 0, 0])
    assert ref.ee_ccd(z, z, V([1.0, 0, 0]), V([1.0, 0, 0]), V([0.0, 5, 0]), V([0.0, 5, 0]),  # :69-81
                      V([1.0, 5, 0]), V([1.0, 5, 0])) is None
    r = ref.cubic_roots(1.0, -1.5, 0.75, -0.125)                      # :84-90
    assert len(r) > 0 and np.all(np.abs(r - 0.5) < 0.01)
    r = np.sort(ref.cubic_roots(1.0, -6.0, 11.0, -6.0))               # :93-101
    assert len(r) == 3 and np.allclose(r, [1, 2, 3], atol=1e-12)


def test_cubic_roots_against_numpy():
    rng = np.random.default_rng(3)
    for _ in range(200):
        co = rng.normal(size=4)
        r = ref.cubic_roots(*co)
        want = np.roots(co)
        want = np.sort(want[np.abs(want.imag) < 1e-9].real)
        for x in r:
            assert abs(np.polyval(co, x)) < 1e-8 * max(1.0, np.abs(co).max() * max(1.0, abs(x)) ** 3)
        if len(want) == 3 and np.min(np.diff(want)) > 1e-3:
            assert len(r) == 3 and np.allclose(np.sort(r), want, atol=1e-9)


def _signed_vol(p, a, b, c):
    return np.dot(a - p, np.cross(b - p, c - p))


def test_vertex_triangle_ccd_matches_first_coplanar_crossing():
    """A vertex moving straight through a moving triangle: the TOI is 0.9 x
    the first sign change of the coplanarity function inside the triangle."""
    rng = np.random.default_rng(4)
    hits = 0
    for _ in range(60):
        a0, b0, c0 = V([0.0, 0, 0]), V([1.0, 0, 0]), V([0.0, 1, 0])
        da, db, dc = rng.normal(size=(3, 3)) * 0.1
        w = rng.dirichlet([2, 2, 2])
        p_on = w[0] * a0 + w[1] * b0 + w[2] * c0
        tc = rng.uniform(0.1, 0.9)
        vel = V([0.0, 0, -1.0]) + rng.normal(size=3) * 0.1
        # vertex path through the triangle's position at time tc
        at, bt, ct = a0 + tc * da, b0 + tc * db, c0 + tc * dc
        hit = w[0] * at + w[1] * bt + w[2] * ct
        p0 = hit - tc * vel
        p1 = p0 + vel
        t = ref.vt_ccd(p0, p1, a0, a0 + da, b0, b0 + db, c0, c0 + dc)
        ts = np.linspace(0, 1, 4001)
        P = p0[None] + ts[:, None] * (p1 - p0)[None]
        A, B, Cc = a0[None] + ts[:, None] * da, b0[None] + ts[:, None] * db, c0[None] + ts[:, None] * dc
        f = np.einsum('ij,ij->i', A - P, np.cross(B - P, Cc - P))
        k = np.nonzero(np.sign(f[1:]) != np.sign(f[:-1]))[0]
        assert len(k) > 0
        first = ts[k[0]]
        assert t is not None
        assert abs(t - 0.9 * first) < 1e-3
        hits += 1
    assert hits == 60


# ---- bvh.rs (bvh_tests.rs) and the exact pair set ----------------------------

def test_tri_pairs_sweep_equals_all_pairs():
    for c, margin, seed in ((6, 0.01, 0), (10, 0.2, 1), (8, 0.0, 2)):
        px, py, pz, idx = mesh.quad_grid(c, c, 1.0, 1.0)
        rng = np.random.default_rng(seed)
        pz = (rng.normal(size=px.size) * 0.05).astype(np.float32)  # crumple
        a = ref.tri_pairs(px, py, pz, idx, margin)
        b = ref.tri_pairs(px, py, pz, idx, margin, brute=True)
        assert a.shape == b.shape and np.array_equal(a, b)
        assert np.all(a[:, 0] < a[:, 1])
    # bvh_broad_phase_sparse_mesh_fewer_pairs (bvh_tests.rs:90-115)
    d = ref.tri_pairs(*mesh.quad_grid(4, 4, 0.1, 0.1), 0.01)
    s = ref.tri_pairs(*mesh.quad_grid(4, 4, 100.0, 100.0), 0.01)
    assert len(s) <= len(d)


# ---- intersection-freeness check ---------------------------------------------

def _cross_np(P, Q):
    """Independent form: an edge of one triangle meets the plane of the other
    strictly between its ends, at a point with positive barycentrics."""
    def seg_hits(p, q, a, b, c):
        n = np.cross(b - a, c - a)
        dp, dq = n @ (p - a), n @ (q - a)
        if dp * dq >= 0:
            return False
        x = p + dp / (dp - dq) * (q - p)
        w = [np.cross(b - x, c - x) @ n, np.cross(c - x, a - x) @ n, np.cross(a - x, b - x) @ n]
        return min(w) > 0
    return any(seg_hits(P[e], P[(e + 1) % 3], *Q) or seg_hits(Q[e], Q[(e + 1) % 3], *P) for e in range(3))


def test_crossings_kats_and_independent_form():
    tri = V([0, 1, 2, 3, 4, 5], dtype=np.uint32)
    base = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0]])
    def count(Q):
        return ref.crossings(np.concatenate([base, Q]).T, tri)
    assert count(np.array([[0.2, 0.2, -0.5], [0.3, 0.2, 0.5], [0.2, 0.4, 0.5]])) == 1   # pierces the face
    assert count(np.array([[0.2, 0.2, 0.1], [0.3, 0.2, 0.5], [0.2, 0.4, 0.5]])) == 0    # above it
    assert count(np.array([[2.0, 2.0, -0.5], [2.1, 2.0, 0.5], [2.0, 2.1, 0.5]])) == 0   # beside it
    assert count(np.array([[0.2, 0.2, 0.0], [0.3, 0.2, 0.5], [0.2, 0.4, 0.5]])) == 0    # touches at a vertex
    assert count(np.array([[0.1, 0.1, 0.0], [0.6, 0.1, 0.0], [0.1, 0.6, 0.0]])) == 0    # coplanar overlap
    # a triangle whose edge passes through the face with its vertices outside it
    assert count(np.array([[-0.5, 0.25, -0.3], [1.5, 0.25, 0.3], [0.5, 2.0, 0.0]])) == 1
    # pairs sharing a vertex are skipped even when they cross
    assert ref.crossings(np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0.3, 0.3, 0.5], [0.3, 0.3, -0.5]]).T,
                         V([0, 1, 2, 0, 3, 4], dtype=np.uint32)) == 0
    rng = np.random.default_rng(7)
    hits = 0
    for _ in range(3000):
        P, Q = rng.random((3, 3)), rng.random((3, 3))
        want = _cross_np(P, Q)
        got = ref.crossings(np.concatenate([P, Q]).T, tri)
        assert got == int(want)
        hits += got
    assert 300 < hits < 2700


def test_crossings_ignore_coplanar_rounding_noise():
    """Two triangles of one flat sheet that share no vertex, 2 mm apart, in
    randomly rotated planes 1 m from the origin: the volumes are rounding
    noise with random signs, and no crossing may be reported."""
    h = 0.2 / 64
    A = np.array([[0.0, 0.0, 0], [0.0, h, 0], [h, h, 0]])
    B = np.array([[-h, h, 0], [-h, 2 * h, 0], [0.0, 2 * h, 0]])
    tri = V([0, 1, 2, 3, 4, 5], dtype=np.uint32)
    rng = np.random.default_rng(3)
    for _ in range(2000):
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        off = rng.normal(size=3)
        P = np.concatenate([A, B]) @ q.T + off
        assert ref.crossings(P.T, tri) == 0


def test_crossings_box_pairs_equal_all_pairs():
    """Two crumpled sheets through each other (one mesh): the box-pair path
    counts what testing every pair counts; one sheet alone has none."""
    for c, amp, seed in ((8, 0.02, 0), (10, 0.05, 1)):
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
        a, b = ref.crossings(pos, idx2), ref.crossings(pos, idx2, brute=False)
        assert a == b and a > 10
        assert ref.crossings(one, idx) == 0


# ---- ipc_response.rs / pipeline (ipc_response_tests.rs) ----------------------

def _two_triangles(z):
    # a free triangle (0,1,2) and a vertex-on-top configuration: vertex 3 of
    # triangle (3,4,5) sits above the face of (0,1,2)
    px = V([0.0, 1.0, 0.0, 0.25, 1.25, 0.25])
    py = V([0.0, 0.0, 1.0, 0.25, 0.25, 1.25])
    pz = V([0.0, 0.0, 0.0, z, z + 0.5, z + 0.5])
    return np.stack([px, py, pz]), V([0, 1, 2, 3, 4, 5], dtype=np.uint32)


def test_vertex_triangle_barrier_gradient():
    """ipc_detect_close_vertex_triangle / ipc_gradient_nonzero (ipc_response_tests.rs:14-63):
    a vertex 0.05 above a face at d_hat = 0.1 is one contact; its barrier
    gradient on the vertex is kappa b'(d^2) * 2 (p - closest)."""
    pos, tri = _two_triangles(0.05)
    th = 0.03  # broad-phase margin and self-contact thickness: the boxes overlap in z
    q = ref.ipc_params(self_collision=True, thickness=th, exclusion_depth=1)
    dh, kappa = 0.1, 1e4
    r = ref.detect(pos, tri, q, dh, kappa)
    assert r["self_contacts"] >= 1 and r["max_violation"] == 0.0 and r["pairs"] == 1
    g = r["grad"][:, 3]
    d2 = 0.05 ** 2
    want = kappa * ref.barrier_gradient_thick(d2, dh, th) * 2 * V([0, 0, 0.05])
    # vertex 3 also takes part in edge-edge contacts with nonzero gradient
    # only through its edges; the z component is dominated by the VT term
    assert g[2] == pytest.approx(want[2], rel=0.5)
    # the per-vertex normal is the normalised energy gradient (collision_pipeline.rs:241-251),
    # which points into the triangle (opposite to the push)
    assert r["in_contact"][3] == 1 and r["normal"][2, 3] < -0.9
    far = ref.detect(_two_triangles(10.0)[0], tri, q, 0.01, kappa)    # ipc_detect_far_vertex_triangle
    assert far["self_contacts"] == 0 and not far["in_contact"].any()


def test_ground_and_sphere_closed_forms():
    n = 4
    pos = np.zeros((3, n))
    pos[1] = [0.01, -0.02, 0.5, 0.0]  # above (in range), below, far, on
    q = ref.ipc_params(ground=0.0)
    dh, kappa = 1e-3, 1e3
    r = ref.detect(pos, V([0, 1, 2], dtype=np.uint32), q, dh, kappa)
    g = r["grad"][1]
    assert g[0] == pytest.approx(ref.scaled_barrier_gradient(1e-4, dh, kappa) * 2 * 0.01, rel=1e-14)
    assert g[1] == pytest.approx(ref.scaled_barrier_gradient(4e-4, dh, kappa) * 2 * 0.02, rel=1e-14)
    assert g[2] == 0.0 and r["in_contact"].tolist() == [1, 1, 0, 1]
    assert r["max_violation"] == pytest.approx(0.02) and r["active"] == 3
    # hessian_diag (collision_pipeline.rs:272-276): kappa |b''(d^2)| 4 d^2 with
    # d clamped to >= 1e-6; b'' from finite differences of b'
    def bpp(d):
        e = 1e-7 * d
        return (ref.barrier_gradient(d + e, dh) - ref.barrier_gradient(d - e, dh)) / (2 * e)
    hd = r["hessian_diag"]
    assert hd[0] == pytest.approx(kappa * abs(bpp(1e-4)) * 4e-4, rel=1e-6)
    assert hd[1] == pytest.approx(kappa * abs(ref.barrier_hessian(1e-12, dh)) * 4e-12, rel=1e-12)
    assert hd[2] == 0.0 and hd[3] == hd[1]
    # sphere: vertex at radius 1.01 from the centre of a unit sphere
    pos = np.zeros((3, 2))
    pos[:, 0] = [1.01, 0, 0]
    pos[:, 1] = [0, 3.0, 0]
    q = ref.ipc_params(sphere=((0.0, 0.0, 0.0), 1.0))
    r = ref.detect(pos, V([0, 1, 0], dtype=np.uint32), q, dh, kappa)
    assert r["grad"][0, 0] == pytest.approx(ref.scaled_barrier_gradient(1e-4, dh, kappa) * 2 * 0.01, rel=1e-12)
    assert r["grad"][:, 1].tolist() == [0.0, 0.0, 0.0]
    assert r["normal"][0, 0] == pytest.approx(1.0)


def test_ccd_step_closed_forms():
    tri = V([0, 1, 2], dtype=np.uint32)
    p0 = np.zeros((3, 3)); p1 = np.zeros((3, 3))
    p0[1] = [1.0, 2.0, 3.0]
    p1[1] = [-1.0, 2.0, 3.0]  # vertex 0 crosses y = 0 at t = 0.5
    a = ref.ccd_step(p0, p1, tri, ref.ipc_params(ground=0.0))
    assert a == pytest.approx(0.45, rel=1e-14)
    p0 = np.zeros((3, 3)); p1 = np.zeros((3, 3))
    p0[0] = [3.0, 5.0, 5.0]; p1[0] = [-3.0, 5.0, 5.0]  # hits the unit sphere at x = 1: t = 1/3
    a = ref.ccd_step(p0, p1, tri, ref.ipc_params(sphere=((0.0, 0.0, 0.0), 1.0)))
    assert a == pytest.approx(0.9 / 3, rel=1e-12)
    assert ref.ccd_step(p0, p0, tri, ref.ipc_params(ground=-10.0)) == 1.0


# ---- the whole step ----------------------------------------------------------

def _sheet(c):
    px, py, pz, idx = mesh.quad_grid(c, c, 1.0, 1.0)
    hg = mesh.hinges4(idx, px.size)
    pin = mesh.pinned_top_row(c, c)
    return px, py, pz, idx, hg, pin


def test_step_ipc_without_colliders_equals_step():
    """With no collider the AL loop converges at once and, while speeds stay
    under the CFL clamp (0.5 sqrt(d_hat) / dt), step_with_ipc is step()."""
    px, py, pz, idx, hg, pin = _sheet(8)
    a = ref.Oracle(ref.F64, px, py, pz, idx, hg, pin, ref.params())
    b = ref.Oracle(ref.F64, px, py, pz, idx, hg, pin, ref.params())
    q = ref.ipc_params()
    for _ in range(3):
        a.step()
        r, st = b.step_ipc(q)
        assert st.al_iterations == 1 and st.al_converged == 1 and r.iterations == 10
    assert np.abs(a.positions() - b.positions()).max() < 1e-13
    assert np.abs(a.get(1) - b.get(1)).max() < 1e-12


def test_step_ipc_sheet_on_ground_stays_above():
    """A flat sheet dropped 5 cm onto a ground plane: the barrier keeps every
    vertex above the plane, velocities go to rest, lambda is zero off contact."""
    c = 8
    px, py, pz, idx = mesh.quad_grid(c, c, 0.5, 0.5)
    # rotate flat into the XZ plane at height 0.05
    pz2, py2 = py.copy(), np.full_like(py, 0.05)
    hg = mesh.hinges4(idx, px.size)
    pin = np.zeros(px.size, dtype=np.uint8)
    o = ref.Oracle(ref.F64, px, py2, pz2, idx, hg, pin, ref.params(rayleigh=0.01))
    kappa = ref.estimate_kappa(0.0, 0.002, 9.81, 1e-3)
    q = ref.ipc_params(ground=0.0, kappa=kappa)
    miny = []
    for _ in range(40):
        r, st = o.step_ipc(q)
        miny.append(o.positions()[1].min())
    assert min(miny) > 0.0
    assert abs(o.get(1)).max() < 0.05
    assert np.all(np.isfinite(o.positions()))


def test_product_estimate_initial_kappa():
    """vistio_amd.estimate_initial_kappa (the product's barrier.rs:97-115):
    the reference's KATs (barrier_tests.rs:73-87), the closed form at the
    midpoint of the barrier zone and both clamps. With d = d_hat / 2 the
    barrier gradient is |b'(d)| = (d_hat / 2)(1 + 2 ln 2), so
    kappa = sqrt(2) m g / (d_hat^1.5 (1 + 2 ln 2))."""
    import vistio_amd as va
    k1 = va.estimate_initial_kappa(0.0, 0.001, 9.81, 1e-3)
    k2 = va.estimate_initial_kappa(0.0, 0.01, 9.81, 1e-3)
    assert 0 < k1 < k2 < 1e8
    assert va.estimate_initial_kappa(5000.0, 0.001, 9.81, 1e-3) == 5000.0
    for m, g, d in ((0.002, 9.81, 1e-3), (0.01, 9.81, 1e-3), (0.002, 9.81, 4e-4), (0.05, 3.0, 2e-3)):
        want = np.sqrt(2.0) * m * g / (d ** 1.5 * (1.0 + 2.0 * np.log(2.0)))
        assert 1.0 < want < 1e6
        assert va.estimate_initial_kappa(0.0, m, g, d) == pytest.approx(want, rel=1e-13)
        assert va.estimate_initial_kappa(0.0, m, g, d) == pytest.approx(ref.estimate_kappa(0.0, m, g, d), rel=1e-13)
    assert va.estimate_initial_kappa(0.0, 1e-9, 9.81, 1e-3) == 1.0     # clamped below
    assert va.estimate_initial_kappa(0.0, 1e3, 9.81, 1e-3) == 1e6      # clamped above
