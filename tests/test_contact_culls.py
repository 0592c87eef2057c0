"""The pair culls of the GPU contact kernels are exact: a culled pair can
never contribute. The bounds are restated here as the kernels compute them
(vh_contact.hpp: k_ccd_tri_bounds + k_ccd_cull, k_tri_spheres +
k_self_cull) and checked against the oracle's own vertex-triangle CCD and
distance functions on random triangle pairs at many scales. A bound made too
loose on purpose (here, on the CPU) is caught by the same checks."""
import numpy as np

from oracle import ref

F32 = np.float32


def _round_up(v):
    return float(F32(v * (1.0 + 1e-6))) + float(F32(1e-30))


def _tri_record(P, D=None):
    """(centre as float32, radius rounded up, mean displacement as float32,
    largest distance of a corner's displacement from it rounded up)."""
    c = np.array([F32(v) for v in P.mean(0)], dtype=np.float64)
    r = _round_up(np.sqrt(max(((p - c) ** 2).sum() for p in P)))
    if D is None:
        return c, r, np.zeros(3), 0.0
    g = np.array([F32(v) for v in D.mean(0)], dtype=np.float64)
    s = _round_up(np.sqrt(max(((d - g) ** 2).sum() for d in D)))
    return c, r, g, s


def _ccd_culled(A0, DA, B0, DB, half_eps=0.5e-8, ee=False):
    ca, ra, ga, sa = _tri_record(A0, DA)
    cb, rb, gb, sb = _tri_record(B0, DB)
    D = np.sqrt(((ca - cb) ** 2).sum()) * (1.0 + 1e-12) + ra + rb
    W = np.sqrt(((ga - gb) ** 2).sum()) * (1.0 + 1e-12) + sa + sb
    if ee:
        W = max(W, 2.0 * max(sa, sb))
    return W ** 3 < half_eps and 3 * D * W * W < half_eps and 3 * D * D * W < half_eps


def _ccd_hits(A0, A1, B0, B1, ee=False):
    hits = []
    for P0, P1, T0, T1 in ((A0, A1, B0, B1), (B0, B1, A0, A1)):
        for q in range(3):
            t = ref.vt_ccd(P0[q], P1[q], T0[0], T1[0], T0[1], T1[1], T0[2], T1[2])
            if t is not None:
                hits.append(t)
    if ee:
        for i in range(3):
            for j in range(3):
                i1, j1 = (i + 1) % 3, (j + 1) % 3
                t = ref.ee_ccd(A0[i], A1[i], A0[i1], A1[i1], B0[j], B1[j], B0[j1], B1[j1])
                if t is not None:
                    hits.append(t)
    return hits


def _random_pair(rng, scale, motion, common):
    """Two triangles near each other; each corner moves by a shared
    translation of size common plus its own motion of size motion."""
    A0 = rng.normal(size=(3, 3)) * scale
    B0 = A0.mean(0) + rng.normal(size=(3, 3)) * scale * rng.uniform(0.5, 3.0)
    T = rng.normal(size=3) * common
    DA = T + rng.normal(size=(3, 3)) * motion
    DB = T + rng.normal(size=(3, 3)) * motion
    return A0, DA, B0, DB


def _ccd_trials(half_eps, ee=False):
    rng = np.random.default_rng(11)
    culled = culled_with_hit = 0
    hits_seen = 0
    for scale in (3e-4, 1e-3, 3e-3, 1e-2, 3e-2):
        for motion in (1e-7, 1e-6, 1e-5, 1e-4, 1e-3):
            for common in (0.0, 1e-2):
                for _ in range(80):
                    A0, DA, B0, DB = _random_pair(rng, scale, motion, common)
                    hits = _ccd_hits(A0, A0 + DA, B0, B0 + DB, ee)
                    hits_seen += bool(hits)
                    if _ccd_culled(A0, DA, B0, DB, half_eps, ee):
                        culled += 1
                        culled_with_hit += bool(hits)
    return culled, culled_with_hit, hits_seen


def test_ccd_pair_cull_never_drops_a_hit():
    culled, bad, hits = _ccd_trials(0.5e-8)
    print(f"ccd cull: {culled} of 4000 pairs culled, {hits} pairs with a hit", flush=True)
    assert culled > 1000 and hits > 50
    assert bad == 0


def test_ccd_pair_cull_with_edge_edge_never_drops_a_hit():
    culled, bad, hits = _ccd_trials(0.5e-8, ee=True)
    print(f"ccd cull (vt + ee): {culled} of 4000 pairs culled, {hits} pairs with a hit", flush=True)
    assert culled > 1000 and hits > 50
    assert bad == 0


def test_ccd_cull_uses_the_relative_motion():
    """A pair that translates as a piece by 1 cm, with 1e-7 m of motion of
    its own, is culled: its cubic's leading coefficients vanish, whatever the
    shared translation."""
    rng = np.random.default_rng(5)
    A0, DA, B0, DB = _random_pair(rng, 1e-3, 1e-7, 1e-2)
    assert _ccd_culled(A0, DA, B0, DB)
    assert not _ccd_hits(A0, A0 + DA, B0, B0 + DB, ee=True)
    # the same pair with 1 cm of relative motion is kept
    assert not _ccd_culled(A0, DA, B0, DB + np.array([1e-2, 0.0, 0.0]))


def test_ccd_cull_check_catches_a_loose_bound():
    """At 50x the kernel's threshold the bound admits pairs whose cubic has
    roots the solver reports, and the check above flags them."""
    culled, bad, _ = _ccd_trials(25e-8)
    assert bad > 0, culled


def _sphere_culled(A, B, dh, rel=1e-6):
    ca, ra, _, _ = _tri_record(A)
    cb, rb, _, _ = _tri_record(B)
    gap = np.sqrt(((ca - cb) ** 2).sum()) * (1.0 - 1e-12) - ra - rb
    return gap > 0 and gap * gap > dh * (1.0 + rel)


def _min_dist2(A, B):
    d = [ref.pt_dist2(A[q], B[0], B[1], B[2])[0] for q in range(3)]
    d += [ref.pt_dist2(B[q], A[0], A[1], A[2])[0] for q in range(3)]
    d += [ref.ee_dist2(A[i], A[(i + 1) % 3], B[j], B[(j + 1) % 3])[0] for i in range(3) for j in range(3)]
    return min(d)


def _sphere_trials(shrink):
    rng = np.random.default_rng(12)
    culled = bad = near = 0
    for scale in (1e-3, 3e-3, 1e-2):
        for _ in range(700):
            A = rng.normal(size=(3, 3)) * scale
            u = rng.normal(size=3)
            B = A.mean(0) + u / np.linalg.norm(u) * scale * rng.uniform(0.5, 6.0) + rng.normal(size=(3, 3)) * scale
            dh = (scale * rng.uniform(0.05, 3.0)) ** 2
            d2 = _min_dist2(A, B)
            near += d2 < dh
            # shrink < 1 makes the spheres too small: a planted mistake
            ca, ra, _, _ = _tri_record(A)
            cb, rb, _, _ = _tri_record(B)
            gap = np.sqrt(((ca - cb) ** 2).sum()) * (1.0 - 1e-12) - shrink * (ra + rb)
            if gap > 0 and gap * gap > dh * (1.0 + 1e-6):
                culled += 1
                bad += d2 < dh
    return culled, bad, near


def test_sphere_cull_never_drops_a_contact():
    culled, bad, near = _sphere_trials(1.0)
    print(f"sphere cull: {culled} of 2100 pairs culled, {near} pairs within reach", flush=True)
    assert culled > 300 and near > 300
    assert bad == 0
    # and the helper agrees with the test's inline form
    rng = np.random.default_rng(3)
    A, B = rng.normal(size=(3, 3)), rng.normal(size=(3, 3)) + 5.0
    assert _sphere_culled(A, B, 1e-2)


def test_sphere_cull_check_catches_a_loose_bound():
    culled, bad, _ = _sphere_trials(0.5)
    assert bad > 0, culled


def _box_gap2(P, Q):
    """vh_contact.hpp box_gap2: the squared gap between the axis-aligned boxes
    of two point sets, 0 where they overlap."""
    g = np.maximum(0.0, np.maximum(P.min(0) - Q.max(0), Q.min(0) - P.max(0)))
    return float((g * g).sum())


def _box_trials(shrink):
    """Per-test pre-test of self_pair_contacts: for random triangle pairs at
    several scales, every vertex-triangle and edge-edge test the box gap
    skips (gap^2 > dh (1 + 1e-6)) must have d2 >= dh by the oracle's own
    distance functions. shrink < 1 scales the gap up: a planted mistake."""
    rng = np.random.default_rng(13)
    skipped = bad = near = 0
    for scale in (1e-3, 3e-3, 1e-2):
        for _ in range(400):
            A = rng.normal(size=(3, 3)) * scale
            u = rng.normal(size=3)
            B = A.mean(0) + u / np.linalg.norm(u) * scale * rng.uniform(0.2, 4.0) + rng.normal(size=(3, 3)) * scale
            dh = (scale * rng.uniform(0.05, 2.0)) ** 2
            tests = [(A[q:q + 1], B, ref.pt_dist2(A[q], B[0], B[1], B[2])[0]) for q in range(3)]
            tests += [(B[q:q + 1], A, ref.pt_dist2(B[q], A[0], A[1], A[2])[0]) for q in range(3)]
            tests += [(A[[i, (i + 1) % 3]], B[[j, (j + 1) % 3]],
                       ref.ee_dist2(A[i], A[(i + 1) % 3], B[j], B[(j + 1) % 3])[0]) for i in range(3) for j in range(3)]
            for P, Q, d2 in tests:
                near += d2 < dh
                if _box_gap2(P, Q) / shrink ** 2 > dh * (1.0 + 1e-6):
                    skipped += 1
                    bad += d2 < dh
    return skipped, bad, near


def test_box_pretest_never_skips_a_contact():
    skipped, bad, near = _box_trials(1.0)
    print(f"box pre-test: {skipped} of 18000 tests skipped, {near} tests within reach", flush=True)
    assert skipped > 2000 and near > 1000
    assert bad == 0


def test_box_pretest_check_catches_a_loose_bound():
    skipped, bad, _ = _box_trials(0.5)
    assert bad > 0, skipped
