"""The oracle's VertexTriangleTest (oracle/contact_ref.hpp, vt_detect and
vt_proximity; crates/vistio-contact/src/vertex_triangle.rs:25-152), pinned by
closed forms and an independent float64 restatement.

The reference's own tests only check the strategy's name
(contact_tests.rs:366-369), so every value here comes from geometry: a point
over a right triangle has its height as distance, (0, 0, +-1) as normal and
known barycentric coordinates; the -0.01 tolerance, the thickness bound
(inclusive) and the degenerate-triangle exit are each crossed on both sides;
and the candidate loop's rules (only triangles that contain vj, none that
contain vi, ascending triangle order) are checked on a fan of triangles.
"""
import numpy as np
import pytest

from oracle import ref

S = 0.5  # leg of the right triangle (0,0,0), (S,0,0), (0,S,0)


def _one(p, thickness=0.1, tri=((0.0, 0.0, 0.0), (S, 0.0, 0.0), (0.0, S, 0.0))):
    """Contacts of vertex 3 (at p) against the triangle (0, 1, 2), candidate (3, 0)."""
    pts = np.array(list(tri) + [p], dtype=np.float32)
    return ref.narrow_vt(pts[:, 0], pts[:, 1], pts[:, 2], np.array([0, 1, 2], np.uint32), [[3, 0]], thickness)


def test_point_above_centroid():
    c = _one((S / 3, S / 3, 0.03))
    assert len(c) == 1
    r = c[0]
    assert list(r["indices"]) == [3, 0, 1, 2]
    assert r["distance"] == np.float32(-0.03)
    assert list(r["normal"]) == [0.0, 0.0, 1.0]
    assert np.allclose(r["barycentric"], 1.0 / 3.0, atol=1e-6)


def test_point_below_flips_the_normal():
    r = _one((S / 3, S / 3, -0.03))[0]
    assert r["distance"] == np.float32(-0.03)
    assert list(r["normal"]) == [0.0, 0.0, -1.0]


def test_barycentric_order():
    # the projection at (0.6 S, 0.1 S): v along (0 -> 1), w along (0 -> 2), u = 1 - v - w
    r = _one((0.6 * S, 0.1 * S, 0.01))[0]
    assert np.allclose(r["barycentric"], [0.3, 0.6, 0.1], atol=1e-6)


def test_thickness_bound_is_inclusive():
    assert len(_one((S / 3, S / 3, 0.25), thickness=0.25)) == 1
    assert len(_one((S / 3, S / 3, 0.2500001), thickness=0.25)) == 0
    assert len(_one((S / 3, S / 3, -0.26), thickness=0.25)) == 0


@pytest.mark.parametrize("x,hit", [(-0.005, True), (-0.009, True), (-0.011, False), (-0.02, False)])
def test_barycentric_tolerance(x, hit):
    # v = x / S: inside down to -0.01
    assert (len(_one((x * S, 0.5 * S, 0.01))) == 1) == hit


def test_degenerate_triangle_gives_nothing():
    tri = ((0.0, 0.0, 0.0), (S, 0.0, 0.0), (2 * S, 0.0, 0.0))
    assert len(_one((S, 0.0, 0.01), tri=tri)) == 0


def test_candidate_loop_rules():
    # a fan of four triangles around vertex 0, plus a lone vertex 5 above it
    pts = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [-1, 0, 0], [0, -1, 0], [0.1, 0.1, 0.02]], np.float32) * S
    tri = np.array([0, 1, 2, 0, 2, 3, 0, 3, 4, 0, 4, 1], np.uint32)
    x, y, z = pts[:, 0], pts[:, 1], pts[:, 2]
    # vj = 0 is in every triangle; only the first contains the point's projection
    c = ref.narrow_vt(x, y, z, tri, [[5, 0]], 0.1)
    assert [list(r["indices"]) for r in c] == [[5, 0, 1, 2]]
    # vj = 2 lies in triangles 0 and 1; a vertex near both gets their contacts in triangle order
    pts[5] = [-0.001 * S, 0.3 * S, 0.01 * S]
    c = ref.narrow_vt(pts[:, 0], pts[:, 1], pts[:, 2], tri, [[5, 2]], 0.1)
    assert [list(r["indices"]) for r in c] == [[5, 0, 1, 2], [5, 0, 2, 3]]
    # a triangle that contains vi is skipped; vi == vj gives nothing
    c = ref.narrow_vt(x, y, z, tri, [[1, 0], [0, 0], [3, 3]], 10.0)
    assert all(1 not in list(r["indices"][1:]) for r in c)
    assert all(r["indices"][0] == 1 for r in c)
    # candidates are kept in order and each repeats its contacts
    c = ref.narrow_vt(pts[:, 0], pts[:, 1], pts[:, 2], tri, [[5, 2], [5, 2]], 0.1)
    assert len(c) == 4


def _f64_detect(pos, tri, cand, thickness):
    """The same test restated independently in float64 with numpy; returns
    {(k, t): (distance, normal, bary, margin)} where margin is how far the
    case is from a threshold (the f32 decision may differ within it)."""
    T = tri.reshape(-1, 3)
    out = {}
    for k, (vi, vj) in enumerate(cand):
        for t, (a, b, c) in enumerate(T):
            if vi in (a, b, c) or vj not in (a, b, c):
                continue
            P, A, B, C = (pos[:, i].astype(np.float64) for i in (vi, a, b, c))
            n = np.cross(B - A, C - A)
            area2 = np.linalg.norm(n)
            if area2 < 1e-6:
                continue
            n = n / area2
            sd = float(np.dot(P - A, n))
            e1, e2 = B - A, C - A
            G = np.array([[e1 @ e1, e1 @ e2], [e1 @ e2, e2 @ e2]])
            v, w = np.linalg.solve(G, np.array([(P - A) @ e1, (P - A) @ e2]))
            u = 1.0 - v - w
            margin = min(abs(abs(sd) - thickness), abs(u + 0.01), abs(v + 0.01), abs(w + 0.01))
            hit = abs(sd) <= thickness and min(u, v, w) >= -0.01
            out[(k, t)] = (hit, -abs(sd), n if sd >= 0 else -n, (u, v, w), margin)
    return out


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_matches_float64_restatement(seed):
    from oracle import mesh
    px, py, pz, idx = mesh.quad_grid(8, 8, 1.0, 1.0)
    rng = np.random.default_rng(seed)
    pz = (rng.normal(size=px.size) * 0.01).astype(np.float32)
    right = px > 0.0  # the right half folded back over the left, 2 cm above it
    px = px.copy()
    px[right] = -px[right]
    pz[right] += 0.02
    pos = np.stack([px, py, pz])
    d2 = ((pos[:, :, None] - pos[:, None, :]) ** 2).sum(0)
    a, b = np.nonzero((d2 < 0.2 ** 2) & (d2 > 0))
    cand = np.stack([a, b], 1).astype(np.uint32)
    thick = 0.03
    got = ref.narrow_vt(px, py, pz, idx, cand, thick)
    want = _f64_detect(pos, idx, cand, thick)
    T = idx.reshape(-1, 3)
    # the contacts of each vertex (a vertex's candidates share its contacts' first index)
    per_cand = {}
    for r in got:
        per_cand.setdefault(int(r["indices"][0]), []).append(r)
    checked = hits = 0
    for (k, t), (hit, dist, nrm, bary, margin) in want.items():
        if margin < 1e-4:
            continue
        vi = int(cand[k][0])
        rows = [r for r in per_cand.get(vi, []) if tuple(int(x) for x in r["indices"][1:]) == tuple(T[t].tolist())]
        assert bool(rows) == hit, (k, t, hit, margin)
        if hit:
            hits += 1
            r = rows[0]
            assert abs(r["distance"] - dist) < 1e-6
            assert np.allclose(r["normal"], nrm, atol=1e-5)
            assert np.allclose(r["barycentric"], bary, atol=1e-4)
        checked += 1
    assert checked > 100 and hits > 10
