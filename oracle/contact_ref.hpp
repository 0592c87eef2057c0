// CPU restatement of Vistio's IPC contact path, for tests only (the oracle).
//
// Restates, in floating type R (the oracle instantiates double):
//   barrier.rs            log-barrier, its derivatives, thickness variant, kappa
//   distance_primitives.rs point-triangle / edge-edge squared distances and
//                          their gradients
//   ccd.rs                cubic/quadratic roots, vertex-triangle and edge-edge
//                          continuous collision tests, max step size
//   ground_plane.rs, sphere.rs   per-vertex IPC barrier and CCD
//   exclusion.rs          n-ring exclusion sets
//   bvh.rs                AABB predicate (closed intervals, f32 with margin);
//                          candidate pairs are every overlapping triangle pair
//                          (ta < tb), found here by sort-and-sweep on min_x
//   self_collision.rs, ipc_response.rs, collision_pipeline.rs
//                          narrow phase and barrier-force assembly
//   vertex_triangle.rs    the pre-IPC narrow phase (VertexTriangleTest), in f32
// Each function cites the reference lines it follows.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <utility>
#include <vector>

namespace oc {

template <class R> struct P3 {
  R x = 0, y = 0, z = 0;
  P3() = default;
  P3(R a, R b, R c) : x(a), y(b), z(c) {}
  P3 operator+(const P3& o) const { return {x + o.x, y + o.y, z + o.z}; }
  P3 operator-(const P3& o) const { return {x - o.x, y - o.y, z - o.z}; }
  P3 operator*(R s) const { return {x * s, y * s, z * s}; }
  P3 operator-() const { return {-x, -y, -z}; }
  R dot(const P3& o) const { return x * o.x + y * o.y + z * o.z; }
  P3 cross(const P3& o) const { return {y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x}; }
  R len2() const { return dot(*this); }
};

// ---- barrier.rs -----------------------------------------------------------

// barrier.rs:25-34: b(d) = -(d - dh)^2 ln(d / dh) for 0 < d < dh (d is a
// squared distance).
template <class R> R barrier_energy(R d, R dh) {
  if (d >= dh || d <= 0) return 0;
  const R diff = d - dh;
  return -(diff * diff) * std::log(d / dh);
}

// barrier.rs:40-53: db/dd with d clamped to [1e-3 dh, dh].
template <class R> R barrier_gradient(R d, R dh) {
  const R dc = std::min(std::max(d, dh * R(1e-3)), dh);
  if (dc >= dh) return 0;
  const R diff = dc - dh;
  return -diff * (R(2) * std::log(dc / dh) + diff / dc);
}

// barrier.rs:59-70
template <class R> R barrier_hessian(R d, R dh) {
  if (d >= dh || d <= 0) return 0;
  const R diff = d - dh, inv = R(1) / d;
  return -R(2) * std::log(d / dh) - R(4) * diff * inv + diff * diff * inv * inv;
}

// barrier.rs:82-93: kappa * db/dd, +-1e6 if not finite.
template <class R> R scaled_barrier_gradient(R d, R dh, R kappa) {
  const R r = kappa * barrier_gradient(d, dh);
  return std::isfinite(r) ? r : std::copysign(R(1e6), r);
}

// barrier.rs:140-148: d_eff = d - thickness^2; non-positive d_eff takes the
// gradient at d = 1e-12.
template <class R> R barrier_gradient_thick(R d, R dh, R th) {
  const R de = d - th * th;
  if (de <= 0) return barrier_gradient(R(1e-12), dh);
  return barrier_gradient(de, dh);
}

// barrier.rs:151-163
template <class R> R scaled_barrier_gradient_thick(R d, R dh, R kappa, R th) {
  const R r = kappa * barrier_gradient_thick(d, dh, th);
  return std::isfinite(r) ? r : std::copysign(R(1e6), r);
}

// barrier.rs:129-137: energy with thickness offset, 1e6 inside it.
template <class R> R barrier_energy_thick(R d, R dh, R th) {
  const R de = d - th * th;
  if (de <= 0) return R(1e6);
  return barrier_energy(de, dh);
}

// barrier.rs:102-121: kappa = m g / (2 sqrt(dh/2) |b'(dh/2)|), clamped to
// [1, 1e6]; a configured kappa > 0 wins.
template <class R> R estimate_kappa(R configured, R avg_mass, R gravity, R dh) {
  if (configured > 0) return configured;
  const R dm = dh * R(0.5);
  const R g = std::fabs(barrier_gradient(dm, dh));
  if (g < R(1e-12)) return R(1e4);
  const R k = avg_mass * gravity / (R(2) * std::sqrt(dm) * g);
  return std::min(std::max(k, R(1)), R(1e6));
}

// ---- distance_primitives.rs ---------------------------------------------

enum DistType { kFace = 0, kEdge0 = 1, kEdge1 = 2, kEdge2 = 3, kVert0 = 4, kVert1 = 5, kVert2 = 6, kEEInterior = 7,
                kDegenerate = 8 };

// distance_primitives.rs:300-319
template <class R> R point_segment_dist2(const P3<R>& p, const P3<R>& a, const P3<R>& b) {
  const P3<R> ab = b - a, ap = p - a;
  const R l2 = ab.len2();
  if (l2 < R(1e-20)) return ap.len2();
  const R t = std::min(std::max(ap.dot(ab) / l2, R(0)), R(1));
  return (p - (a + ab * t)).len2();
}

// distance_primitives.rs:36-116: Voronoi-region closest point (Ericson,
// Real-Time Collision Detection, 5.1.5); degenerate triangles take the
// nearest of the three edges.
template <class R> R pt_dist2(const P3<R>& p, const P3<R>& a, const P3<R>& b, const P3<R>& c, int* type) {
  const P3<R> ab = b - a, ac = c - a, ap = p - a, bp = p - b, cp = p - c;
  if (ab.cross(ac).len2() < R(1e-20)) {
    const R e0 = point_segment_dist2(p, a, b), e1 = point_segment_dist2(p, b, c), e2 = point_segment_dist2(p, c, a);
    if (e0 <= e1 && e0 <= e2) { *type = kEdge0; return e0; }
    if (e1 <= e2) { *type = kEdge1; return e1; }
    *type = kEdge2; return e2;
  }
  const R d1 = ab.dot(ap), d2 = ac.dot(ap);
  if (d1 <= 0 && d2 <= 0) { *type = kVert0; return ap.len2(); }
  const R d3 = ab.dot(bp), d4 = ac.dot(bp);
  if (d3 >= 0 && d4 <= d3) { *type = kVert1; return bp.len2(); }
  const R d5 = ab.dot(cp), d6 = ac.dot(cp);
  if (d6 >= 0 && d5 <= d6) { *type = kVert2; return cp.len2(); }
  const R vc = d1 * d4 - d3 * d2;
  if (vc <= 0 && d1 >= 0 && d3 <= 0) {
    const R v = d1 / (d1 - d3);
    *type = kEdge0; return (p - (a + ab * v)).len2();
  }
  const R va = d3 * d6 - d5 * d4;
  if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
    const R w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    *type = kEdge1; return (p - (b + (c - b) * w)).len2();
  }
  const R vb = d5 * d2 - d1 * d6;
  if (vb <= 0 && d2 >= 0 && d6 <= 0) {
    const R w = d2 / (d2 - d6);
    *type = kEdge2; return (p - (a + ac * w)).len2();
  }
  const R den = R(1) / (va + vb + vc);
  const R v = vb * den, w = vc * den;
  *type = kFace;
  return (p - (a + ab * v + ac * w)).len2();
}

// distance_primitives.rs:124-177: gradient through the plane projection's
// barycentric coordinates (exact in the face region only, as the reference).
template <class R> void pt_grad(const P3<R>& p, const P3<R>& a, const P3<R>& b, const P3<R>& c, P3<R> g[4]) {
  for (int k = 0; k < 4; ++k) g[k] = P3<R>();
  const P3<R> ab = b - a, ac = c - a;
  const P3<R> nrm = ab.cross(ac);
  const R a2 = nrm.len2();
  if (a2 < R(1e-20)) return;
  const P3<R> n = nrm * (R(1) / std::sqrt(a2));
  const R sd = (p - a).dot(n);
  const P3<R> proj = p - n * sd;
  const P3<R> v2 = proj - a;
  const R d00 = ab.dot(ab), d01 = ab.dot(ac), d11 = ac.dot(ac), d20 = v2.dot(ab), d21 = v2.dot(ac);
  const R den = d00 * d11 - d01 * d01;
  if (std::fabs(den) < R(1e-20)) return;
  const R inv = R(1) / den;
  const R v = (d11 * d20 - d01 * d21) * inv, w = (d00 * d21 - d01 * d20) * inv, u = R(1) - v - w;
  const P3<R> diff = p - (a * u + b * v + c * w);
  g[0] = diff * R(2);
  g[1] = -diff * (R(2) * u);
  g[2] = -diff * (R(2) * v);
  g[3] = -diff * (R(2) * w);
}

// distance_primitives.rs:185-246: segment-segment closest points (Ericson
// 5.1.9) with the reference's eps = 1e-10 branches.
template <class R> R ee_dist2(const P3<R>& a0, const P3<R>& a1, const P3<R>& b0, const P3<R>& b1, int* type) {
  const P3<R> d1 = a1 - a0, d2 = b1 - b0, r = a0 - b0;
  const R a = d1.dot(d1), e = d2.dot(d2), f = d2.dot(r);
  const R eps = R(1e-10);
  if (a < eps && e < eps) { *type = kDegenerate; return r.len2(); }
  R s, t;
  if (a < eps) {
    s = 0; t = std::min(std::max(f / e, R(0)), R(1));
  } else {
    const R c = d1.dot(r);
    if (e < eps) {
      t = 0; s = std::min(std::max(-c / a, R(0)), R(1));
    } else {
      const R b = d1.dot(d2);
      const R den = a * e - b * b;
      s = std::fabs(den) > eps ? std::min(std::max((b * f - c * e) / den, R(0)), R(1)) : R(0);
      t = (b * s + f) / e;
      if (t < 0) { t = 0; s = std::min(std::max(-c / a, R(0)), R(1)); }
      else if (t > 1) { t = 1; s = std::min(std::max((b - c) / a, R(0)), R(1)); }
    }
  }
  *type = kEEInterior;
  return ((a0 + d1 * s) - (b0 + d2 * t)).len2();
}

// distance_primitives.rs:251-303: gradient at the closest-point parameters
// (both degenerate cases collapse to s = t = 0, as the reference).
template <class R> void ee_grad(const P3<R>& a0, const P3<R>& a1, const P3<R>& b0, const P3<R>& b1, P3<R> g[4]) {
  const P3<R> d1 = a1 - a0, d2 = b1 - b0, r = a0 - b0;
  const R a = d1.dot(d1), e = d2.dot(d2), f = d2.dot(r), c = d1.dot(r), b = d1.dot(d2);
  const R eps = R(1e-10);
  R s = 0, t = 0;
  if (!(a < eps || e < eps)) {
    const R den = a * e - b * b;
    s = std::fabs(den) > eps ? std::min(std::max((b * f - c * e) / den, R(0)), R(1)) : R(0);
    t = (b * s + f) / e;
    if (t < 0) { t = 0; s = std::min(std::max(-c / a, R(0)), R(1)); }
    else if (t > 1) { t = 1; s = std::min(std::max((b - c) / a, R(0)), R(1)); }
  }
  const P3<R> diff = (a0 + d1 * s) - (b0 + d2 * t);
  g[0] = diff * (R(2) * (R(1) - s));
  g[1] = diff * (R(2) * s);
  g[2] = -diff * (R(2) * (R(1) - t));
  g[3] = -diff * (R(2) * t);
}

// ---- ccd.rs ---------------------------------------------------------------

constexpr double kCcdSafety = 0.9;  // ccd.rs:13
constexpr double kCcdMinToi = 1e-6;  // ccd.rs:16

// ccd.rs:352-377: real roots of a t^2 + b t + c, eps = 1e-8 decisions.
template <class R> int quadratic_roots(R a, R b, R c, R out[2]) {
  const R eps = R(1e-8);
  if (std::fabs(a) < eps) {
    if (std::fabs(b) < eps) return 0;
    out[0] = -c / b;
    return 1;
  }
  const R disc = b * b - R(4) * a * c;
  if (disc < -eps) return 0;
  if (std::fabs(disc) <= eps) { out[0] = -b / (R(2) * a); return 1; }
  const R sq = std::sqrt(disc);
  out[0] = (-b + sq) / (R(2) * a);
  out[1] = (-b - sq) / (R(2) * a);
  return 2;
}

// ccd.rs:299-349: Cardano on the depressed cubic; three real roots by the
// trigonometric form.
template <class R> int cubic_roots(R a, R b, R c, R d, R out[3]) {
  const R eps = R(1e-8);
  if (std::fabs(a) < eps) return quadratic_roots(b, c, d, out);
  const R p = b / a, q = c / a, r = d / a;
  const R ad = q - p * p / R(3);
  const R bd = r - p * q / R(3) + R(2) * p * p * p / R(27);
  const R disc = bd * bd / R(4) + ad * ad * ad / R(27);
  const R shift = -p / R(3);
  if (disc > eps) {
    const R sq = std::sqrt(disc);
    out[0] = std::cbrt(-bd / R(2) + sq) + std::cbrt(-bd / R(2) - sq) + shift;
    return 1;
  }
  if (std::fabs(disc) <= eps) {
    if (std::fabs(bd) < eps) { out[0] = shift; return 1; }
    const R u = std::cbrt(bd / R(2));
    out[0] = R(-2) * u + shift;
    out[1] = u + shift;
    return 2;
  }
  const R rv = std::sqrt(-ad * ad * ad / R(27));
  if (std::fabs(rv) < R(1e-12)) { out[0] = shift; return 1; }
  const R cv = std::min(std::max(-bd / (R(2) * rv), R(-1)), R(1));
  const R th = std::acos(cv);
  const R m = R(2) * std::cbrt(rv);
  const R pi = R(3.14159265358979323846);
  out[0] = m * std::cos(th / R(3)) + shift;
  out[1] = m * std::cos((th + R(2) * pi) / R(3)) + shift;
  out[2] = m * std::cos((th + R(4) * pi) / R(3)) + shift;
  return 3;
}

// ccd.rs:32-129: coplanarity cubic of the vertex relative to the moving
// triangle, then a barycentric inside test (tolerance 0.01) at each root in
// [-1e-6, 1 + 1e-6]; returns the smallest 0.9 t (at least 1e-6).
template <class R>
bool vt_ccd(const P3<R>& p0, const P3<R>& p1, const P3<R>& a0, const P3<R>& a1, const P3<R>& b0, const P3<R>& b1,
            const P3<R>& c0, const P3<R>& c1, R* toi) {
  const P3<R> dp = p1 - p0;
  const P3<R> ar = a0 - p0, br = b0 - p0, cr = c0 - p0;
  const P3<R> av = (a1 - a0) - dp, bv = (b1 - b0) - dp, cv = (c1 - c0) - dp;
  const P3<R> vv = bv.cross(cv), rv = br.cross(cv), vr = bv.cross(cr), rr = br.cross(cr);
  const R ca = av.dot(vv);
  const R cb = ar.dot(vv) + av.dot(rv + vr);
  const R cc = ar.dot(rv + vr) + av.dot(rr);
  const R cd = ar.dot(rr);
  R roots[3];
  const int nr = cubic_roots(ca, cb, cc, cd, roots);
  bool hit = false;
  R best = 0;
  for (int k = 0; k < nr; ++k) {
    R t = roots[k];
    if (t < -R(kCcdMinToi) || t > R(1) + R(kCcdMinToi)) continue;
    t = std::min(std::max(t, R(0)), R(1));
    const P3<R> pa = ar + av * t, pb = br + bv * t, pc = cr + cv * t;
    if ((pb - pa).cross(pc - pa).len2() < R(1e-20)) continue;
    const P3<R> v0 = pb - pa, v1 = pc - pa, v2 = -pa;
    const R d00 = v0.dot(v0), d01 = v0.dot(v1), d11 = v1.dot(v1), d20 = v2.dot(v0), d21 = v2.dot(v1);
    const R den = d00 * d11 - d01 * d01;
    if (std::fabs(den) < R(1e-20)) continue;
    const R inv = R(1) / den;
    const R u = (d11 * d20 - d01 * d21) * inv, w = (d00 * d21 - d01 * d20) * inv, v = R(1) - u - w;
    const R tol = R(0.01);
    if (v >= -tol && u >= -tol && w >= -tol) {
      const R safe = std::max(t * R(kCcdSafety), R(kCcdMinToi));
      best = hit ? std::min(best, safe) : safe;
      hit = true;
    }
  }
  if (hit) *toi = best;
  return hit;
}

// ccd.rs:138-229: edge-edge coplanarity cubic and segment test (tolerance
// 0.05). The reference has no caller of it on the step path.
template <class R>
bool ee_ccd(const P3<R>& a0, const P3<R>& a1, const P3<R>& b0, const P3<R>& b1, const P3<R>& c0, const P3<R>& c1,
            const P3<R>& d0, const P3<R>& d1, R* toi) {
  const P3<R> da = a1 - a0, db = b1 - b0, dc = c1 - c0, dd = d1 - d0;
  const P3<R> r0 = c0 - a0, rv = dc - da, e10 = b0 - a0, e1v = db - da, e20 = d0 - c0, e2v = dd - dc;
  const P3<R> vv = e1v.cross(e2v), rvx = e10.cross(e2v), vrx = e1v.cross(e20), rr = e10.cross(e20);
  const R ca = rv.dot(vv);
  const R cb = r0.dot(vv) + rv.dot(rvx + vrx);
  const R cc = r0.dot(rvx + vrx) + rv.dot(rr);
  const R cd = r0.dot(rr);
  R roots[3];
  const int nr = cubic_roots(ca, cb, cc, cd, roots);
  bool hit = false;
  R best = 0;
  for (int k = 0; k < nr; ++k) {
    R t = roots[k];
    if (t < -R(kCcdMinToi) || t > R(1) + R(kCcdMinToi)) continue;
    t = std::min(std::max(t, R(0)), R(1));
    const P3<R> at = a0 + da * t, bt = b0 + db * t, ct = c0 + dc * t, dt = d0 + dd * t;
    const P3<R> e1 = bt - at, e2 = dt - ct, r = ct - at;
    const R l1 = e1.dot(e1), l2 = e2.dot(e2);
    if (l1 < R(1e-20) || l2 < R(1e-20)) continue;
    const R bb = e1.dot(e2), ff = e2.dot(r), cc2 = e1.dot(r);
    const R den = l1 * l2 - bb * bb;
    if (std::fabs(den) < R(1e-10)) continue;
    const R s = std::min(std::max((bb * ff - cc2 * l2) / den, R(0)), R(1));
    const R u = std::min(std::max((bb * s + ff) / l2, R(0)), R(1));
    const R tol = R(0.05);
    if (s >= -tol && s <= R(1) + tol && u >= -tol && u <= R(1) + tol) {
      const R safe = std::max(t * R(kCcdSafety), R(kCcdMinToi));
      best = hit ? std::min(best, safe) : safe;
      hit = true;
    }
  }
  if (hit) *toi = best;
  return hit;
}

// ---- broad phase: bvh.rs Aabb (f32) ---------------------------------------

struct Box32 { float lo[3], hi[3]; };

// bvh.rs:28-42: per-triangle AABB of f32 positions, inflated by margin.
inline Box32 tri_box(const float* px, const float* py, const float* pz, const int v[3], float margin) {
  Box32 b;
  const float* c[3] = {px, py, pz};
  for (int a = 0; a < 3; ++a) {
    const float x0 = c[a][v[0]], x1 = c[a][v[1]], x2 = c[a][v[2]];
    b.lo[a] = std::min(std::min(x0, x1), x2) - margin;
    b.hi[a] = std::max(std::max(x0, x1), x2) + margin;
  }
  return b;
}

// bvh.rs:57-62: closed-interval overlap on all three axes.
inline bool overlaps(const Box32& a, const Box32& b) {
  return a.lo[0] <= b.hi[0] && a.hi[0] >= b.lo[0] && a.lo[1] <= b.hi[1] && a.hi[1] >= b.lo[1] &&
         a.lo[2] <= b.hi[2] && a.hi[2] >= b.lo[2];
}

// Every overlapping triangle pair (ta < tb), sorted. Sort-and-sweep on
// lo[0] gives the same set as testing all pairs (the predicate is the same).
inline std::vector<std::pair<int, int>> tri_pairs(const std::vector<Box32>& box) {
  const int t = int(box.size());
  std::vector<int> ord(t);
  for (int i = 0; i < t; ++i) ord[i] = i;
  std::sort(ord.begin(), ord.end(), [&](int a, int b) { return box[a].lo[0] < box[b].lo[0] || (box[a].lo[0] == box[b].lo[0] && a < b); });
  std::vector<std::pair<int, int>> out;
  for (int ii = 0; ii < t; ++ii) {
    const Box32& A = box[ord[ii]];
    for (int jj = ii + 1; jj < t; ++jj) {
      const Box32& B = box[ord[jj]];
      if (B.lo[0] > A.hi[0]) break;
      if (overlaps(A, B)) out.push_back({std::min(ord[ii], ord[jj]), std::max(ord[ii], ord[jj])});
    }
  }
  std::sort(out.begin(), out.end());
  return out;
}

// ---- intersection-freeness (no reference function) -------------------------

// The check behind the contact scenes' promise of no interpenetration: two
// triangles with no shared vertex cross when an edge of one properly passes
// through the other (signs of tetrahedron volumes; touching and coplanar
// overlap are not counted).
// Volumes below 1e-10 of |b-a| |c-a| |d-a| are 0 (coplanar rounding noise).
inline int orient_sign(const P3<double>& a, const P3<double>& b, const P3<double>& c, const P3<double>& d) {
  const P3<double> u = b - a, w = c - a, z = d - a;
  const double v = u.cross(w).dot(z);
  if (v * v <= 1e-20 * (u.len2() * w.len2() * z.len2())) return 0;
  return (v > 0.0) - (v < 0.0);
}
inline bool seg_crosses_tri(const P3<double>& p, const P3<double>& q, const P3<double>& a, const P3<double>& b,
                            const P3<double>& c) {
  const int s1 = orient_sign(a, b, c, p), s2 = orient_sign(a, b, c, q);
  if (s1 * s2 >= 0) return false;
  const int t1 = orient_sign(p, q, a, b), t2 = orient_sign(p, q, b, c), t3 = orient_sign(p, q, c, a);
  return t1 != 0 && t1 == t2 && t2 == t3;
}
inline bool tris_cross(const P3<double> P[3], const P3<double> Q[3]) {
  for (int e = 0; e < 3; ++e) {
    if (seg_crosses_tri(P[e], P[(e + 1) % 3], Q[0], Q[1], Q[2])) return true;
    if (seg_crosses_tri(Q[e], Q[(e + 1) % 3], P[0], P[1], P[2])) return true;
  }
  return false;
}

// ---- exclusion.rs ---------------------------------------------------------

// exclusion.rs:22-41 with topology.rs:126-141: the set of v plus every vertex
// reachable through `depth` one-ring expansions (one ring = the other corners
// of v's triangles). Returned as sorted CSR.
inline void exclusion_sets(int n, int t, const uint32_t* tri, int depth, std::vector<int>& ptr,
                           std::vector<int>& idx) {
  std::vector<std::vector<int>> ring(n);
  for (int e = 0; e < t; ++e)
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        if (a != b) ring[tri[3 * e + a]].push_back(int(tri[3 * e + b]));
  for (auto& r : ring) { std::sort(r.begin(), r.end()); r.erase(std::unique(r.begin(), r.end()), r.end()); }
  ptr.assign(n + 1, 0);
  idx.clear();
  std::vector<int> cur, next;
  for (int v = 0; v < n; ++v) {
    cur.assign(1, v);
    for (int d = 0; d < depth; ++d) {
      next = cur;
      for (int u : cur) next.insert(next.end(), ring[u].begin(), ring[u].end());
      std::sort(next.begin(), next.end());
      next.erase(std::unique(next.begin(), next.end()), next.end());
      cur.swap(next);
    }
    idx.insert(idx.end(), cur.begin(), cur.end());
    ptr[v + 1] = int(idx.size());
  }
}

inline bool excluded(const std::vector<int>& ptr, const std::vector<int>& idx, int i, int j) {
  return std::binary_search(idx.begin() + ptr[i], idx.begin() + ptr[i + 1], j);
}

// ---- collision_pipeline.rs::detect_ipc_contacts ---------------------------

struct IpcScene {
  bool has_ground = false;
  double ground = 0;
  bool has_sphere = false;
  double center[3] = {0, 0, 0};
  double radius = 0;
  bool self_collision = false;
  double thickness = 0;  // broad-phase margin and self-contact thickness
  bool ccd_edge_edge = false;  // the CCD step also runs edge_edge_ccd on each candidate pair
};

template <class R> struct Forces {
  std::vector<R> gx, gy, gz, nx, ny, nz;
  std::vector<R> hd;  // hessian_diag: ground and sphere terms (collision_pipeline.rs:272-276, 301-305)
  std::vector<uint8_t> in_contact;
  R max_violation = 0;
  long long active = 0, self_contacts = 0;
  void reset(int n) {
    gx.assign(n, 0); gy.assign(n, 0); gz.assign(n, 0);
    nx.assign(n, 0); ny.assign(n, 0); nz.assign(n, 0);
    hd.assign(n, 0);
    in_contact.assign(n, 0);
    max_violation = 0; active = 0; self_contacts = 0;
  }
};

// collision_pipeline.rs:207-351 with self_collision.rs:199-276,
// ipc_response.rs:58-219, ground_plane.rs:68-110, sphere.rs:102-176.
// Self-contact pairs come from `pairs` (triangle indices into tri).
template <class R>
void detect(int n, const R* px, const R* py, const R* pz, int /*t*/, const uint32_t* tri,
            const std::vector<std::pair<int, int>>& pairs, const std::vector<int>& xptr,
            const std::vector<int>& xidx, const IpcScene& sc, R dh, R kappa, Forces<R>& F) {
  F.reset(n);
  auto P = [&](int i) { return P3<R>(px[i], py[i], pz[i]); };
  if (sc.self_collision) {
    std::vector<R> sx(n, 0), sy(n, 0), sz(n, 0);
    const R th = R(sc.thickness);
    auto add = [&](const int idx[4], const P3<R> g[4], R bg) {
      for (int k = 0; k < 4; ++k) { sx[idx[k]] += bg * g[k].x; sy[idx[k]] += bg * g[k].y; sz[idx[k]] += bg * g[k].z; }
    };
    auto bgrad = [&](R d2) { return th > 0 ? scaled_barrier_gradient_thick(d2, dh, kappa, th) : scaled_barrier_gradient(d2, dh, kappa); };
    for (const auto& pr : pairs) {
      const int A[3] = {int(tri[3 * pr.first]), int(tri[3 * pr.first + 1]), int(tri[3 * pr.first + 2])};
      const int B[3] = {int(tri[3 * pr.second]), int(tri[3 * pr.second + 1]), int(tri[3 * pr.second + 2])};
      for (int side = 0; side < 2; ++side) {  // vertices of A vs triangle B, then B vs A
        const int* V = side == 0 ? A : B;
        const int* T = side == 0 ? B : A;
        for (int k = 0; k < 3; ++k) {
          const int v = V[k];
          if (v == T[0] || v == T[1] || v == T[2]) continue;
          if (excluded(xptr, xidx, v, T[0]) || excluded(xptr, xidx, v, T[1]) || excluded(xptr, xidx, v, T[2])) continue;
          int ty;
          const R d2 = pt_dist2(P(v), P(T[0]), P(T[1]), P(T[2]), &ty);
          if (d2 < dh && d2 > 0) {
            const int idx[4] = {v, T[0], T[1], T[2]};
            P3<R> g[4];
            pt_grad(P(v), P(T[0]), P(T[1]), P(T[2]), g);
            add(idx, g, bgrad(d2));
            F.self_contacts++;
            if (d2 <= R(1e-12)) F.max_violation = std::max(F.max_violation, R(1e-3));
          }
        }
      }
      const int EA[3][2] = {{A[0], A[1]}, {A[1], A[2]}, {A[2], A[0]}};
      const int EB[3][2] = {{B[0], B[1]}, {B[1], B[2]}, {B[2], B[0]}};
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          const int a0 = EA[i][0], a1 = EA[i][1], b0 = EB[j][0], b1 = EB[j][1];
          if (a0 == b0 || a0 == b1 || a1 == b0 || a1 == b1) continue;
          if (excluded(xptr, xidx, a0, b0) && excluded(xptr, xidx, a0, b1) && excluded(xptr, xidx, a1, b0) &&
              excluded(xptr, xidx, a1, b1))
            continue;
          int ty;
          const R d2 = ee_dist2(P(a0), P(a1), P(b0), P(b1), &ty);
          if (d2 < dh && d2 > 0) {
            const int idx[4] = {a0, a1, b0, b1};
            P3<R> g[4];
            ee_grad(P(a0), P(a1), P(b0), P(b1), g);
            add(idx, g, bgrad(d2));
            F.self_contacts++;
            if (d2 <= R(1e-12)) F.max_violation = std::max(F.max_violation, R(1e-3));
          }
        }
    }
    for (int i = 0; i < n; ++i) {
      F.gx[i] += sx[i]; F.gy[i] += sy[i]; F.gz[i] += sz[i];
      const R g2 = sx[i] * sx[i] + sy[i] * sy[i] + sz[i] * sz[i];
      if (g2 > R(1e-12)) {
        const R gl = std::sqrt(g2);
        F.nx[i] += sx[i] / gl; F.ny[i] += sy[i] / gl; F.nz[i] += sz[i] / gl;
        F.in_contact[i] = 1;
      }
    }
    F.active += F.self_contacts;
  }
  if (sc.has_ground) {
    const R h = R(sc.ground);
    for (int i = 0; i < n; ++i) {
      const R ds = py[i] - h;
      if (ds <= 0) {
        F.max_violation = std::max(F.max_violation, -ds);
        const R dcl = std::max(-ds, R(1e-6));
        F.active++;
        F.gy[i] += scaled_barrier_gradient(dcl * dcl, dh, kappa) * R(2) * dcl;
      } else if (ds * ds < dh) {
        F.active++;
        F.gy[i] += scaled_barrier_gradient(ds * ds, dh, kappa) * R(2) * ds;
      }
      if (ds * ds < dh || ds <= 0) {
        F.ny[i] += R(1); F.in_contact[i] = 1;
        const R de = std::max(ds, R(1e-6)), d2 = de * de;
        F.hd[i] += std::fabs(kappa * barrier_hessian(d2, dh) * R(4) * d2);
      }
    }
  }
  if (sc.has_sphere) {
    const R cx = R(sc.center[0]), cy = R(sc.center[1]), cz = R(sc.center[2]), rad = R(sc.radius);
    for (int i = 0; i < n; ++i) {
      const R dx = px[i] - cx, dy = py[i] - cy, dz = pz[i] - cz;
      const R r = std::sqrt(dx * dx + dy * dy + dz * dz);
      const R ds = r - rad;
      if (ds <= 0) {
        F.max_violation = std::max(F.max_violation, -ds);
        const R dcl = std::max(-ds, R(1e-6));
        F.active++;
        const R bg = scaled_barrier_gradient(dcl * dcl, dh, kappa);
        if (r > R(1e-6)) {
          const R f = bg * R(2) * dcl;
          F.gx[i] += f * (dx / r); F.gy[i] += f * (dy / r); F.gz[i] += f * (dz / r);
        } else {
          F.gy[i] += bg * R(2) * dcl;
        }
      } else if (ds * ds < dh) {
        F.active++;
        const R f = scaled_barrier_gradient(ds * ds, dh, kappa) * R(2) * ds;
        F.gx[i] += f * (dx / r); F.gy[i] += f * (dy / r); F.gz[i] += f * (dz / r);
      }
      if (ds * ds < dh || ds <= 0) {
        if (r > R(1e-6)) { F.nx[i] += dx / r; F.ny[i] += dy / r; F.nz[i] += dz / r; }
        F.in_contact[i] = 1;
        const R de = std::max(ds, R(1e-6)), d2 = de * de;
        F.hd[i] += std::fabs(kappa * barrier_hessian(d2, dh) * R(4) * d2);
      }
    }
  }
  for (int i = 0; i < n; ++i)
    if (F.in_contact[i]) {
      const R l2 = F.nx[i] * F.nx[i] + F.ny[i] * F.ny[i] + F.nz[i] * F.nz[i];
      if (l2 > R(1e-12)) {
        const R inv = R(1) / std::sqrt(l2);
        F.nx[i] *= inv; F.ny[i] *= inv; F.nz[i] *= inv;
      }
    }
}

// collision_pipeline.rs:355-400 with ccd.rs:238-292, ground_plane.rs:112-134,
// sphere.rs:178-224: the largest safe step fraction from (p0) to (p1).
template <class R>
R ccd_step(int n, const R* x0, const R* y0, const R* z0, const R* x1, const R* y1, const R* z1, const uint32_t* tri,
           const std::vector<std::pair<int, int>>& pairs, const IpcScene& sc) {
  R toi = 1;
  if (sc.self_collision) {
    R self = 1;
    auto P0 = [&](int i) { return P3<R>(x0[i], y0[i], z0[i]); };
    auto P1 = [&](int i) { return P3<R>(x1[i], y1[i], z1[i]); };
    for (const auto& pr : pairs) {
      const int A[3] = {int(tri[3 * pr.first]), int(tri[3 * pr.first + 1]), int(tri[3 * pr.first + 2])};
      const int B[3] = {int(tri[3 * pr.second]), int(tri[3 * pr.second + 1]), int(tri[3 * pr.second + 2])};
      for (int side = 0; side < 2; ++side) {
        const int* V = side == 0 ? A : B;
        const int* T = side == 0 ? B : A;
        for (int k = 0; k < 3; ++k) {
          const int v = V[k];
          if (v == T[0] || v == T[1] || v == T[2]) continue;
          R tt;
          if (vt_ccd(P0(v), P1(v), P0(T[0]), P1(T[0]), P0(T[1]), P1(T[1]), P0(T[2]), P1(T[2]), &tt)) self = std::min(self, tt);
        }
      }
      if (sc.ccd_edge_edge)  // edges (0,1), (1,2), (2,0) of A against those of B, no shared vertex
        for (int ea = 0; ea < 3; ++ea)
          for (int eb = 0; eb < 3; ++eb) {
            const int a0 = A[ea], a1 = A[(ea + 1) % 3], b0 = B[eb], b1 = B[(eb + 1) % 3];
            if (a0 == b0 || a0 == b1 || a1 == b0 || a1 == b1) continue;
            R tt;
            if (ee_ccd(P0(a0), P1(a0), P0(a1), P1(a1), P0(b0), P1(b0), P0(b1), P1(b1), &tt)) self = std::min(self, tt);
          }
    }
    toi = std::min(toi, std::max(self, R(kCcdMinToi)));
  }
  if (sc.has_ground) {
    R g = 1;
    const R h = R(sc.ground);
    for (int i = 0; i < n; ++i)
      if (y0[i] > h && y1[i] < h) {
        const R t = (h - y0[i]) / (y1[i] - y0[i]);
        if (t >= 0 && t <= 1) g = std::min(g, t * R(0.9));
      }
    toi = std::min(toi, std::max(g, R(1e-6)));
  }
  if (sc.has_sphere) {
    R s = 1;
    const R r2 = R(sc.radius * sc.radius);
    for (int i = 0; i < n; ++i) {
      const R ax = x0[i] - R(sc.center[0]), ay = y0[i] - R(sc.center[1]), az = z0[i] - R(sc.center[2]);
      if (ax * ax + ay * ay + az * az < r2) continue;
      const R bx = x1[i] - R(sc.center[0]), by = y1[i] - R(sc.center[1]), bz = z1[i] - R(sc.center[2]);
      const R vx = bx - ax, vy = by - ay, vz = bz - az;
      const R a = vx * vx + vy * vy + vz * vz;
      const R b = R(2) * (ax * vx + ay * vy + az * vz);
      const R c = ax * ax + ay * ay + az * az - r2;
      if (a > R(1e-8)) {
        const R disc = b * b - R(4) * a * c;
        if (disc >= 0) {
          const R t = (-b - std::sqrt(disc)) / (R(2) * a);
          if (t >= 0 && t <= 1) s = std::min(s, t * R(0.9));
        }
      }
    }
    toi = std::min(toi, std::max(s, R(1e-6)));
  }
  return std::max(toi, R(1e-6));
}

// ---- vertex_triangle.rs: the pre-IPC narrow phase ---------------------------

// One contact of VertexTriangleTest (contact.rs:22-44 ContactPair; the type is
// always VertexTriangle and is_self always true there).
struct VtContact {
  uint32_t idx[4];  // vertex, triangle's three vertices
  float distance;   // -|plane distance|
  float normal[3];  // from the triangle to the vertex
  float bary[3];    // (u, v, w)
};

// point_triangle_proximity, vertex_triangle.rs:88-152, in f32 with the
// reference's operation order (glam Vec3: component-wise ops, cross as
// y*z' - y'*z, dot left to right, length = sqrt(dot), division by the
// scalar per component); the oracle is compiled without FMA contraction.
inline bool vt_proximity(const float p[3], const float a[3], const float b[3], const float c[3], float thickness,
                         VtContact& o) {
  const float ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  const float ac[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  const float ap[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
  float nrm[3] = {ab[1] * ac[2] - ac[1] * ab[2], ab[2] * ac[0] - ac[2] * ab[0], ab[0] * ac[1] - ac[0] * ab[1]};
  const float area2 = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
  if (area2 < 1e-10f) return false;  // :108-111 degenerate triangle
  nrm[0] = nrm[0] / area2; nrm[1] = nrm[1] / area2; nrm[2] = nrm[2] / area2;
  const float sd = ap[0] * nrm[0] + ap[1] * nrm[1] + ap[2] * nrm[2];  // :114 signed plane distance
  const float ad = std::fabs(sd);
  if (ad > thickness) return false;  // :117-119
  const float proj[3] = {p[0] - nrm[0] * sd, p[1] - nrm[1] * sd, p[2] - nrm[2] * sd};
  const float app[3] = {proj[0] - a[0], proj[1] - a[1], proj[2] - a[2]};
  const float d00 = ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2];
  const float d01 = ab[0] * ac[0] + ab[1] * ac[1] + ab[2] * ac[2];
  const float d11 = ac[0] * ac[0] + ac[1] * ac[1] + ac[2] * ac[2];
  const float d20 = app[0] * ab[0] + app[1] * ab[1] + app[2] * ab[2];
  const float d21 = app[0] * ac[0] + app[1] * ac[1] + app[2] * ac[2];
  const float den = d00 * d11 - d01 * d01;
  if (std::fabs(den) < 1e-10f) return false;  // :133-136
  const float inv = 1.0f / den;
  const float v = (d11 * d20 - d01 * d21) * inv;
  const float w = (d00 * d21 - d01 * d20) * inv;
  const float u = 1.0f - v - w;
  const float tol = -0.01f;
  if (u < tol || v < tol || w < tol) return false;  // :142-146
  const float sgn = sd >= 0.0f ? 1.0f : -1.0f;          // :148-149 normal toward the vertex
  o.distance = -ad;
  o.normal[0] = sgn > 0 ? nrm[0] : -nrm[0];
  o.normal[1] = sgn > 0 ? nrm[1] : -nrm[1];
  o.normal[2] = sgn > 0 ? nrm[2] : -nrm[2];
  o.bary[0] = u; o.bary[1] = v; o.bary[2] = w;
  return true;
}

// VertexTriangleTest::detect, vertex_triangle.rs:25-81: for each candidate
// (a = vertex vi, b = vertex vj) in order, every triangle t, ascending, that
// contains vj and not vi.
inline std::vector<VtContact> vt_detect(const float* x, const float* y, const float* z, int t, const uint32_t* tri,
                                        long long nc, const uint32_t* ca, const uint32_t* cb, float thickness) {
  std::vector<VtContact> out;
  for (long long k = 0; k < nc; ++k) {
    const uint32_t vi = ca[k], vj = cb[k];
    const float p[3] = {x[vi], y[vi], z[vi]};
    for (int e = 0; e < t; ++e) {
      const uint32_t A = tri[3 * e], B = tri[3 * e + 1], Cc = tri[3 * e + 2];
      if (vi == A || vi == B || vi == Cc) continue;
      if (vj != A && vj != B && vj != Cc) continue;
      const float pa[3] = {x[A], y[A], z[A]}, pb[3] = {x[B], y[B], z[B]}, pc[3] = {x[Cc], y[Cc], z[Cc]};
      VtContact o;
      if (vt_proximity(p, pa, pb, pc, thickness, o)) {
        o.idx[0] = vi; o.idx[1] = A; o.idx[2] = B; o.idx[3] = Cc;
        out.push_back(o);
      }
    }
  }
  return out;
}

}  // namespace oc
