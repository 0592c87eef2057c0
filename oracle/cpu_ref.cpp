// cpu_ref — CPU restatement of Vistio's Projective Dynamics step.
//
// TEST INFRASTRUCTURE ONLY. This file is the oracle: it is compiled into
// oracle/_build/libvh_oracle.so and is loaded by tests/, by the smoke check
// and by the CPU-baseline leg of bench.py, always as the reference the GPU
// results are compared with, or as the CPU baseline, never as the product path. The product (vistio_amd/) never
// links, loads or calls it.
//
// It follows the reference algorithm line by line (citations are
// file:line in the reference tree) and is templated on the element/state
// precision R:
//   * mode ORACLE_F64      R = double everywhere, params taken as double.
//                          This is "the reference algorithm re-evaluated in
//                          f64", the FP64 oracle the GPU paths are held to.
//   * mode ORACLE_F32_REF  R = float for state, elements and projections;
//                          the global solve runs in f64 with f32 in/out,
//                          like FaerSolver (crates/vistio-math/src/faer_solver.rs:95-129).
//                          This reproduces the reference's own precision split.
//
// Deliberate, documented differences from the reference (none changes the
// mathematics):
//   * interior edges are sorted by (v0, v1) instead of HashMap order
//     (crates/vistio-mesh/src/topology.rs:73-112 iterates a randomly seeded
//     HashMap, so the reference's own order is not reproducible);
//   * the sparse LL^T is this file's own simplicial (up-looking) Cholesky
//     with a geometric nested-dissection ordering, not faer 0.24 (absent).
//     Only the rounding of the f64 solve can differ from faer's.
//
// Build: see oracle/Makefile (g++ -O3 -ffp-contract=off, no FMA contraction,
// matching Rust's default of never fusing a*b+c).

#ifdef _OPENMP
#include <omp.h>
#endif
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "contact_ref.hpp"

namespace {

// ---------------------------------------------------------------------------
// Small vector types with glam 0.29 evaluation order (scalar Vec3/Vec2:
// dot = x*x' + y*y' + z*z' left to right; normalize = v * (1/|v|)).
// ---------------------------------------------------------------------------
template <class R> struct V3 {
  R x, y, z;
  V3() : x(0), y(0), z(0) {}
  V3(R a, R b, R c) : x(a), y(b), z(c) {}
  V3 operator+(const V3& o) const { return V3(x + o.x, y + o.y, z + o.z); }
  V3 operator-(const V3& o) const { return V3(x - o.x, y - o.y, z - o.z); }
  V3 operator*(R s) const { return V3(x * s, y * s, z * s); }
  V3 operator/(R s) const { return V3(x / s, y / s, z / s); }
  R dot(const V3& o) const { return x * o.x + y * o.y + z * o.z; }
  V3 cross(const V3& o) const {
    return V3(y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x);
  }
  R length() const { return std::sqrt(dot(*this)); }
  V3 normalize() const { R r = R(1) / length(); return *this * r; }
};

template <class R> struct V2 {
  R x, y;
  V2() : x(0), y(0) {}
  V2(R a, R b) : x(a), y(b) {}
  V2 normalize() const {
    R len = std::sqrt(x * x + y * y);
    R r = R(1) / len;
    return V2(x * r, y * r);
  }
};

// 3x2 column-major matrix (crates/vistio-math/src/mat3x2.rs:16-68).
template <class R> struct M32 {
  V3<R> c0, c1;
  // result = self * [[m0, m2], [m1, m3]]  (mat3x2.rs:63-68)
  M32 mul2(const R m[4]) const {
    M32 o;
    o.c0 = c0 * m[0] + c1 * m[1];
    o.c1 = c0 * m[2] + c1 * m[3];
    return o;
  }
};

template <class R> struct Polar {
  M32<R> rot;
  R stretch[4];
  R s0, s1;
  V2<R> v0, v1;
};

// polar_decomposition_3x2 — crates/vistio-math/src/decomposition.rs:76-138.
template <class R> Polar<R> polar_decomposition(const M32<R>& f) {
  const R a = f.c0.dot(f.c0);
  const R b = f.c0.dot(f.c1);
  const R d = f.c1.dot(f.c1);
  const R half_trace = R(0.5) * (a + d);
  const R half_diff = R(0.5) * (a - d);
  const R disc = std::sqrt(half_diff * half_diff + b * b);
  R lambda0 = half_trace + disc; if (!(lambda0 >= R(0))) lambda0 = R(0);  // .max(0.0)
  R lambda1 = half_trace - disc; if (!(lambda1 >= R(0))) lambda1 = R(0);
  const R s0 = std::sqrt(lambda0);
  const R s1 = std::sqrt(lambda1);
  const R eps = R(1.0e-10);  // DEGENERATE_AREA_THRESHOLD, crates/vistio-types/src/constants.rs
  Polar<R> p;
  if (s0 < eps) {
    p.rot.c0 = V3<R>(1, 0, 0); p.rot.c1 = V3<R>(0, 1, 0);
    p.stretch[0] = p.stretch[1] = p.stretch[2] = p.stretch[3] = 0;
    p.s0 = p.s1 = 0; p.v0 = V2<R>(1, 0); p.v1 = V2<R>(0, 1);
    return p;
  }
  V2<R> v0, v1;
  if (std::fabs(b) > eps) {
    v0 = V2<R>(lambda0 - d, b).normalize();
    v1 = V2<R>(-v0.y, v0.x);
  } else if (a >= d) {
    v0 = V2<R>(1, 0); v1 = V2<R>(0, 1);
  } else {
    v0 = V2<R>(0, 1); v1 = V2<R>(1, 0);
  }
  const R s00 = v0.x * v0.x * s0 + v1.x * v1.x * s1;
  const R s01 = v0.x * v0.y * s0 + v1.x * v1.y * s1;
  const R s11 = v0.y * v0.y * s0 + v1.y * v1.y * s1;
  const R inv_s0 = R(1) / s0;
  const R inv_s1 = s1 > eps ? R(1) / s1 : R(0);
  const R si00 = v0.x * v0.x * inv_s0 + v1.x * v1.x * inv_s1;
  const R si01 = v0.x * v0.y * inv_s0 + v1.x * v1.y * inv_s1;
  const R si11 = v0.y * v0.y * inv_s0 + v1.y * v1.y * inv_s1;
  const R sinv[4] = {si00, si01, si01, si11};
  p.rot = f.mul2(sinv);
  p.stretch[0] = s00; p.stretch[1] = s01; p.stretch[2] = s01; p.stretch[3] = s11;
  p.s0 = s0; p.s1 = s1; p.v0 = v0; p.v1 = v1;
  return p;
}

// clamp_stretch — decomposition.rs:50-66 (sigma -> min(sigma, 1)).
template <class R> void clamp_stretch(const Polar<R>& p, R out[4]) {
  const R cs0 = std::min(p.s0, R(1));
  const R cs1 = std::min(p.s1, R(1));
  const R s00 = p.v0.x * p.v0.x * cs0 + p.v1.x * p.v1.x * cs1;
  const R s01 = p.v0.x * p.v0.y * cs0 + p.v1.x * p.v1.y * cs1;
  const R s11 = p.v0.y * p.v0.y * cs0 + p.v1.y * p.v1.y * cs1;
  out[0] = s00; out[1] = s01; out[2] = s01; out[3] = s11;
}

// compute_dm_inv — crates/vistio-solver/src/element.rs:228-263.
template <class R> void compute_dm_inv(const V3<R>& e1, const V3<R>& e2, R out[4]) {
  const R len_e1 = e1.length();
  const R tiny = R(1e-10);
  out[0] = 1; out[1] = 0; out[2] = 0; out[3] = 1;
  if (len_e1 < tiny) return;
  const V3<R> u = e1 / len_e1;
  const V3<R> n = e1.cross(e2);
  const R n_len = n.length();
  if (n_len < tiny) return;
  const V3<R> v = n.cross(u).normalize();
  const R dm00 = e1.dot(u), dm10 = e1.dot(v), dm01 = e2.dot(u), dm11 = e2.dot(v);
  const R det = dm00 * dm11 - dm01 * dm10;
  if (std::fabs(det) < tiny) return;
  const R inv_det = R(1) / det;
  out[0] = dm11 * inv_det;
  out[1] = -dm10 * inv_det;
  out[2] = -dm01 * inv_det;
  out[3] = dm00 * inv_det;
}

// compute_dihedral_angle — crates/vistio-solver/src/bending.rs:56-91.
template <class R> R dihedral_angle(const V3<R>& p0, const V3<R>& p1, const V3<R>& pa, const V3<R>& pb) {
  const R PI = R(M_PI);  // std::f32::consts::PI or std::f64::consts::PI, rounded to R
  const V3<R> edge = p1 - p0;
  const R edge_len = edge.length();
  if (edge_len < R(1e-10)) return PI;
  const V3<R> e = edge / edge_len;
  const V3<R> to_a = pa - p0, to_b = pb - p0;
  V3<R> perp_a = to_a - e * to_a.dot(e);
  V3<R> perp_b = to_b - e * to_b.dot(e);
  const R la = perp_a.length(), lb = perp_b.length();
  if (la < R(1e-10) || lb < R(1e-10)) return PI;
  perp_a = perp_a / la;
  perp_b = perp_b / lb;
  R c = perp_a.dot(perp_b);
  if (c < R(-1)) c = R(-1);
  if (c > R(1)) c = R(1);
  return std::acos(c);
}

// rotate_around_axis — bending.rs:231-239 (Rodrigues).
template <class R> V3<R> rotate_around_axis(const V3<R>& p, const V3<R>& center, const V3<R>& axis, R angle) {
  const V3<R> v = p - center;
  const R ca = std::cos(angle);
  const R sa = std::sin(angle);
  const R d = axis.dot(v);
  const V3<R> rotated = v * ca + axis.cross(v) * sa + axis * d * (R(1) - ca);
  return center + rotated;
}

// BendingData::project — bending.rs:182-227.
template <class R> void bending_project(const V3<R>& p0, const V3<R>& p1, const V3<R>& pa, const V3<R>& pb,
                                        R rest_angle, V3<R> out[4]) {
  out[0] = p0; out[1] = p1; out[2] = pa; out[3] = pb;
  const R cur = dihedral_angle(p0, p1, pa, pb);
  const R diff = cur - rest_angle;
  if (std::fabs(diff) < R(1e-8)) return;
  const V3<R> edge = p1 - p0;
  const R edge_len = edge.length();
  if (edge_len < R(1e-10)) return;
  const V3<R> axis = edge / edge_len;
  const V3<R> to_a = pa - p0, to_b = pb - p0;
  const V3<R> q_a = to_a - axis * to_a.dot(axis);
  const V3<R> q_b = to_b - axis * to_b.dot(axis);
  const V3<R> nc = q_a.cross(q_b);
  const R sign = nc.dot(axis) >= R(0) ? R(1) : R(-1);
  const R half = diff * R(0.5) * sign;
  const V3<R> mid = (p0 + p1) * R(0.5);
  out[2] = rotate_around_axis(pa, mid, axis, half);
  out[3] = rotate_around_axis(pb, mid, axis, -half);
}

enum Model { MODEL_ARAP = 0, MODEL_COROT = 1, MODEL_ISO = 2, MODEL_ORTHO = 3, MODEL_ANISO = 4 };

// Target deformation gradient of each membrane model (ConstitutiveModel::project):
//   ARAP        R                                     element.rs:127-144
//   COROT       R S(min(sigma, 1))                    corotational.rs:51-72
//   ISO         I (Mat3x2::IDENTITY)                  isotropic.rs:46-58
//   ORTHO       R V diag(s + (1 - s) k) V^T, k = (warp, weft) / max(warp, weft, 1e-6)
//                                                     orthotropic.rs:58-88
//   ANISO       R V diag(s > 1 ? 1 + (s - 1)(1 - k) : s) V^T with k0 = kw c^2 + kf s^2,
//               k1 = kw s^2 + kf c^2 (c, s from v0; kw, kf clamped to [0, 1])
//                                                     anisotropic.rs:78-143
template <class R> M32<R> target_of(const M32<R>& f, int model, R warp, R weft) {
  if (model == MODEL_ISO) {
    M32<R> I; I.c0 = V3<R>(R(1), R(0), R(0)); I.c1 = V3<R>(R(0), R(1), R(0));
    return I;
  }
  const Polar<R> pd = polar_decomposition(f);
  if (model == MODEL_COROT) {
    R cs[4]; clamp_stretch(pd, cs);
    return pd.rot.mul2(cs);
  }
  if (model == MODEL_ORTHO || model == MODEL_ANISO) {
    R t0, t1;
    if (model == MODEL_ORTHO) {
      const R mk = std::max(std::max(warp, weft), R(1e-6));
      const R kw = warp / mk, kf = weft / mk;
      t0 = pd.s0 + (R(1) - pd.s0) * kw;
      t1 = pd.s1 + (R(1) - pd.s1) * kf;
    } else {
      const R kw = std::min(std::max(warp, R(0)), R(1)), kf = std::min(std::max(weft, R(0)), R(1));
      const R c2 = pd.v0.x * pd.v0.x, s2 = pd.v0.y * pd.v0.y;
      const R k0 = kw * c2 + kf * s2, k1 = kw * s2 + kf * c2;
      t0 = pd.s0 > R(1) ? R(1) + (pd.s0 - R(1)) * (R(1) - k0) : pd.s0;
      t1 = pd.s1 > R(1) ? R(1) + (pd.s1 - R(1)) * (R(1) - k1) : pd.s1;
    }
    const R t00 = pd.v0.x * pd.v0.x * t0 + pd.v1.x * pd.v1.x * t1;
    const R t01 = pd.v0.x * pd.v0.y * t0 + pd.v1.x * pd.v1.y * t1;
    const R t11 = pd.v0.y * pd.v0.y * t0 + pd.v1.y * pd.v1.y * t1;
    const R ts[4] = {t00, t01, t01, t11};
    return pd.rot.mul2(ts);
  }
  return pd.rot;
}

// ElementData::project / project_with_model + reconstruct_from_target_f —
// element.rs:127-144, 153-174, 181-217, with the model's target above.
template <class R> void membrane_project(const V3<R>& p0, const V3<R>& p1, const V3<R>& p2,
                                         const R dm_inv[4], int model, V3<R> out[3], R warp = R(0),
                                         R weft = R(0)) {
  M32<R> ds; ds.c0 = p1 - p0; ds.c1 = p2 - p0;
  const M32<R> f = ds.mul2(dm_inv);  // deformation_gradient, decomposition.rs:147-155
  const M32<R> target = target_of(f, model, warp, weft);
  const V3<R> centroid = (p0 + p1 + p2) / R(3);
  const R det = dm_inv[0] * dm_inv[3] - dm_inv[1] * dm_inv[2];
  if (!(std::fabs(det) > R(1e-12))) { out[0] = p0; out[1] = p1; out[2] = p2; return; }
  const R inv_det = R(1) / det;
  const R dm00 = dm_inv[3] * inv_det, dm01 = -dm_inv[2] * inv_det;
  const R dm10 = -dm_inv[1] * inv_det, dm11 = dm_inv[0] * inv_det;
  const V3<R> re1 = target.c0 * dm00 + target.c1 * dm10;
  const V3<R> re2 = target.c0 * dm01 + target.c1 * dm11;
  const V3<R> rest_c = (re1 + re2) / R(3);
  const V3<R> q0 = centroid - rest_c;
  out[0] = q0; out[1] = q0 + re1; out[2] = q0 + re2;
}

// ---------------------------------------------------------------------------
// Sparse LL^T (double), this file's own: geometric nested dissection ordering
// + up-looking simplicial Cholesky (elimination tree + row reach).
// ---------------------------------------------------------------------------
struct SparseChol {
  int n = 0;
  std::vector<int> perm, iperm;  // perm[new] = old
  std::vector<int> Lp, Li;       // column-compressed L, diagonal first
  std::vector<double> Lx;
  bool ok = false;

  // Nested dissection by recursive coordinate bisection; the vertex separator
  // is the smaller of the two one-sided boundaries of the cut.
  static void nd(const std::vector<int>& Ap, const std::vector<int>& Ai, const double* X, const double* Y,
                 const double* Z, std::vector<int>& verts, std::vector<int>& order, std::vector<int>& mark,
                 int& stamp) {
    if (verts.size() <= 32) { for (int v : verts) order.push_back(v); return; }
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int v : verts) {
      const double c[3] = {X[v], Y[v], Z[v]};
      for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], c[k]); hi[k] = std::max(hi[k], c[k]); }
    }
    int ax = 0;
    for (int k = 1; k < 3; ++k) if (hi[k] - lo[k] > hi[ax] - lo[ax]) ax = k;
    const double* C = ax == 0 ? X : (ax == 1 ? Y : Z);
    std::vector<int> s(verts);
    std::sort(s.begin(), s.end(), [&](int a, int b) { return C[a] < C[b] || (C[a] == C[b] && a < b); });
    const size_t half = s.size() / 2;
    // fresh stamps per call, so marks never need clearing
    const int L = ++stamp, Rm = ++stamp, Sep = ++stamp;
    for (size_t i = 0; i < s.size(); ++i) mark[s[i]] = i < half ? L : Rm;
    std::vector<int> bl, br;
    for (size_t i = 0; i < s.size(); ++i) {
      const int v = s[i];
      const int other = mark[v] == L ? Rm : L;
      bool touch = false;
      for (int p = Ap[v]; p < Ap[v + 1]; ++p) if (mark[Ai[p]] == other) { touch = true; break; }
      if (touch) (mark[v] == L ? bl : br).push_back(v);
    }
    std::vector<int>& sep = bl.size() <= br.size() ? bl : br;
    std::vector<int> left, right;
    for (int v : sep) mark[v] = Sep;
    for (int v : s) {
      if (mark[v] == L) left.push_back(v);
      else if (mark[v] == Rm) right.push_back(v);
    }
    if (left.empty() || right.empty()) { for (int v : verts) order.push_back(v); return; }
    nd(Ap, Ai, X, Y, Z, left, order, mark, stamp);
    nd(Ap, Ai, X, Y, Z, right, order, mark, stamp);
    for (int v : sep) order.push_back(v);
  }

  // Factor the symmetric matrix given by its full CSR pattern (Ap, Ai, Ax).
  bool factor(int n_, const std::vector<int>& Ap, const std::vector<int>& Ai, const std::vector<double>& Ax,
              const double* X, const double* Y, const double* Z) {
    n = n_;
    ok = false;
    {
      std::vector<int> verts(n), order, mark(n, 0);
      std::iota(verts.begin(), verts.end(), 0);
      order.reserve(n);
      int stamp = 0;
      nd(Ap, Ai, X, Y, Z, verts, order, mark, stamp);
      perm = order;
    }
    iperm.assign(n, 0);
    for (int k = 0; k < n; ++k) iperm[perm[k]] = k;
    // C = P A P^T, upper part stored by columns: column j holds rows i <= j.
    std::vector<int> Cp(n + 1, 0), Ci;
    std::vector<double> Cx;
    for (int r = 0; r < n; ++r)
      for (int p = Ap[r]; p < Ap[r + 1]; ++p) {
        const int i = iperm[r], j = iperm[Ai[p]];
        if (i <= j) Cp[j + 1]++;
      }
    for (int j = 0; j < n; ++j) Cp[j + 1] += Cp[j];
    Ci.resize(Cp[n]); Cx.resize(Cp[n]);
    {
      std::vector<int> w(Cp.begin(), Cp.end() - 1);
      for (int r = 0; r < n; ++r)
        for (int p = Ap[r]; p < Ap[r + 1]; ++p) {
          const int i = iperm[r], j = iperm[Ai[p]];
          if (i <= j) { Ci[w[j]] = i; Cx[w[j]] = Ax[p]; w[j]++; }
        }
    }
    // elimination tree (Liu) on the upper pattern
    std::vector<int> parent(n, -1), ancestor(n, -1);
    for (int k = 0; k < n; ++k)
      for (int p = Cp[k]; p < Cp[k + 1]; ++p) {
        int i = Ci[p];
        while (i != -1 && i < k) {
          const int next = ancestor[i];
          ancestor[i] = k;
          if (next == -1) { parent[i] = k; break; }
          i = next;
        }
      }
    // row reach of L(k,:) : nodes i < k reachable from A(0:k-1, k) via the etree
    std::vector<int> flag(n, -1), stack(n);
    auto ereach = [&](int k, int& top) {
      top = n;
      flag[k] = k;
      for (int p = Cp[k]; p < Cp[k + 1]; ++p) {
        int i = Ci[p];
        if (i > k) continue;
        int len = 0;
        for (; flag[i] != k; i = parent[i]) { stack[len++] = i; flag[i] = k; }
        while (len > 0) stack[--top] = stack[--len];
      }
    };
    // column counts
    std::vector<int> cnt(n, 1);
    for (int k = 0; k < n; ++k) {
      int top; ereach(k, top);
      for (int q = top; q < n; ++q) cnt[stack[q]]++;
    }
    Lp.assign(n + 1, 0);
    for (int k = 0; k < n; ++k) Lp[k + 1] = Lp[k] + cnt[k];
    Li.assign(Lp[n], 0); Lx.assign(Lp[n], 0.0);
    std::vector<int> c(Lp.begin(), Lp.end() - 1);
    std::vector<double> x(n, 0.0);
    std::fill(flag.begin(), flag.end(), -1);
    for (int k = 0; k < n; ++k) {
      int top; ereach(k, top);
      x[k] = 0;
      for (int p = Cp[k]; p < Cp[k + 1]; ++p) if (Ci[p] <= k) x[Ci[p]] += Cx[p];
      double d = x[k];
      x[k] = 0;
      for (; top < n; ++top) {
        const int i = stack[top];
        const double lki = x[i] / Lx[Lp[i]];
        x[i] = 0;
        for (int p = Lp[i] + 1; p < c[i]; ++p) x[Li[p]] -= Lx[p] * lki;
        d -= lki * lki;
        const int p = c[i]++;
        Li[p] = k; Lx[p] = lki;
      }
      if (!(d > 0)) return false;
      const int p = c[k]++;
      Li[p] = k; Lx[p] = std::sqrt(d);
    }
    ok = true;
    return true;
  }

  // x = A^{-1} b (b, x in original order); work has length n.
  void solve(const double* b, double* x, std::vector<double>& w) const {
    for (int k = 0; k < n; ++k) w[k] = b[perm[k]];
    for (int j = 0; j < n; ++j) {  // L y = b
      const double yj = w[j] / Lx[Lp[j]];
      w[j] = yj;
      for (int p = Lp[j] + 1; p < Lp[j + 1]; ++p) w[Li[p]] -= Lx[p] * yj;
    }
    for (int j = n - 1; j >= 0; --j) {  // L^T x = y
      double s = w[j];
      for (int p = Lp[j] + 1; p < Lp[j + 1]; ++p) s -= Lx[p] * w[Li[p]];
      w[j] = s / Lx[Lp[j]];
    }
    for (int k = 0; k < n; ++k) x[perm[k]] = w[k];
  }
};

}  // namespace

// ---------------------------------------------------------------------------
// Public C API (ctypes). Everything below is test infrastructure.
// ---------------------------------------------------------------------------
extern "C" {

struct or_params {
  int model;              // 0 = ARAP (init()), 1 = co-rotational (init_with_material())
  double k_membrane;      // per-area membrane stiffness: weight = k * area (element.rs:77)
  double k_bending;       // bending stiffness: weight = k * edge_len (bending.rs:131-132)
  double density;         // kg/m^2 (pd_solver.rs:1221 divides g/m^2 by 1000)
  double dt_matrix;       // dt used in A (pd_solver.rs:130 hard-codes 1/60)
  double gravity[3];
  double damping;         // v *= 1 - damping (state.rs:233-240)
  double rayleigh;        // v *= 1/(1 + a*dt) for free vertices (pd_solver.rs:456-466)
  int max_iterations;
  double tolerance;
  int has_ground;
  double ground_height;
  int bending_kind;       // 1 = dihedral (bending.rs), 2 = Discrete Shells (discrete_shells.rs)
  double warp, weft;      // membrane model stiffness (orthotropic / anisotropic models)
  double bend_warp, bend_weft;  // Discrete Shells warp/weft modulation (0, 0 = isotropic)
};

struct or_step_result { int iterations; double final_residual; int converged; double wall_time; };

// Contact settings of ProjectiveDynamicsSolver::step_with_ipc
// (crates/vistio-solver/src/config.rs:47-120) and of the collision pipeline
// the runner builds (crates/vistio-bench/src/runner.rs:60-95).
struct or_ipc_params {
  int has_ground;          // ground plane y = ground_height (also the state's ground)
  double ground_height;
  int has_sphere;          // analytic sphere collider
  double sphere_center[3];
  double sphere_radius;
  int self_collision;      // triangle-pair VT/EE barrier + VT CCD
  int exclusion_depth;     // n-ring exclusion (runner: 1)
  double d_hat;            // barrier_d_hat (squared distance)
  double kappa;            // barrier stiffness (runner: estimate_initial_kappa)
  double thickness;        // broad-phase margin and self-contact thickness (runner: 1.5 d_hat)
  int al_max_iterations;
  double al_mu_initial, al_mu_growth, al_tolerance;
  int compliant_contact;
  double compliant_d_hat_scale;
  double friction_coefficient, contact_damping;
  int adaptive_contact_damping;
  double contact_damping_max, contact_velocity_threshold;
  // 1: the handler's set_d_hat is IpcCollisionHandler's default no-op
  // (pd_solver.rs:488), as in the reference's own drape test
  // (crates/vistio-bench/tests/bench_tests.rs:181-199): detection keeps
  // d_hat while the CFL clamp uses the compliant-scaled one.
  int keep_d_hat;
  // 1: the CCD step also runs edge_edge_ccd (ccd.rs:138-229) on the 9 edge
  // pairs of every candidate triangle pair; 0: vertex-triangle only, as
  // compute_max_step_size (ccd.rs:238-292).
  int ccd_edge_edge;
};

// Diagnostics of one IPC step (no counterpart in the reference's StepResult).
struct or_ipc_stats {
  int al_iterations;        // outer iterations run
  int al_converged;
  double mu;                // penalty after the step
  double max_violation;     // of the last detection
  long long active;         // active contacts of the last detection
  long long self_contacts;
  long long pairs;          // candidate triangle pairs of the last detection
  double last_alpha;        // step fraction of the last PD iteration
};

}  // extern "C"

namespace {

// ---- pieces of step_with_ipc, each exported for closed-form tests ----------
// Test hook: a nonzero value plants one known mistake in the pieces below, so
// that a test can show its closed-form cases reject that mistake. The oracle
// sets it only through or_set_mutation; 0 (the default) is the reference.
//   1 friction acts on the normal component and the tangential one is dropped
//   2 lambda -= mu grad          3 the normal component is kept
//   4 backtracking halves when the barrier pushes along the step
//   5 global contact damping multiplies by 0.05 instead of 0.95
//   6 mu grows before the multiplier update
//   7 adaptive damping blends the wrong way (fast vertices damped most)
//   8 the 0.8 CCD safety factor also scales a full step
int g_mutation = 0;

// Step fraction of one PD iteration from the CCD bound (pd_solver.rs:895-906)
// and the backtracking on the frozen barrier gradient (pd_solver.rs:909-932):
// grad_dot_d = sum over free vertices of grad . (sol - pos), diff_sq = sum of
// |sol - pos|^2.
template <class R> R ipc_step_fraction(R ccd_alpha, double grad_dot_d, R mu, double diff_sq) {
  R max_alpha = ccd_alpha;
  if (max_alpha < R(1) || g_mutation == 8) max_alpha *= R(0.8);
  max_alpha = std::min(max_alpha, R(1));
  R alpha = max_alpha;
  for (int bt = 0; bt < 4; ++bt) {
    if (alpha < R(0.1)) break;
    const double g = grad_dot_d * double(mu);
    const bool against = g_mutation == 4 ? -g > 1e-4 * diff_sq * 100.0 : g > 1e-4 * diff_sq * 100.0;
    if (against) alpha *= R(0.5);
    else break;
  }
  return alpha;
}

// Augmented-Lagrangian multiplier update and penalty growth,
// pd_solver.rs:1006-1021: lambda += mu grad where in contact, then mu *= growth.
template <class R>
void ipc_al_update(int n, const uint8_t* in_contact, const R* gx, const R* gy, const R* gz, R* lx, R* ly, R* lz,
                   R& mu, R growth) {
  const R m = g_mutation == 6 ? mu * growth : mu;
  for (int i = 0; i < n; ++i)
    if (in_contact[i]) {
      if (g_mutation == 2) { lx[i] -= m * gx[i]; ly[i] -= m * gy[i]; lz[i] -= m * gz[i]; }
      else { lx[i] += m * gx[i]; ly[i] += m * gy[i]; lz[i] += m * gz[i]; }
    }
  mu *= growth;
}

// The inelastic velocity filter, pd_solver.rs:1036-1111: for free vertices
// that were in contact during the step and have a unit contact normal,
// remove the normal velocity, apply Coulomb friction with |v.n| as the
// normal-impulse proxy (stick when mu_f |v_n| >= |v_t|, else scale v_t by
// 1 - mu_f |v_n| / |v_t|), then damp by the contact damping (adaptive: blended
// towards contact_damping_max as the speed falls below the threshold).
template <class R> struct VelFilter {
  R mu_f, cd, cd_max, v_thresh;
  bool adaptive;
};
template <class R>
void ipc_velocity_filter(int n, R* vx, R* vy, R* vz, const R* inv_mass, const uint8_t* ever, const R* fnx,
                         const R* fny, const R* fnz, const VelFilter<R>& f) {
  for (int i = 0; i < n; ++i) {
    if (inv_mass[i] == R(0) || !ever[i]) continue;
    const R nx = fnx[i], ny = fny[i], nz = fnz[i];
    if (nx * nx + ny * ny + nz * nz < R(0.5)) continue;
    const R vn = vx[i] * nx + vy[i] * ny + vz[i] * nz;
    if (g_mutation == 1) {  // the components swapped: friction on the normal part
      const R tx = vx[i] - vn * nx, ty = vy[i] - vn * ny, tz = vz[i] - vn * nz;
      const R vt = std::sqrt(tx * tx + ty * ty + tz * tz), vnm = std::fabs(vn);
      const R s2 = vnm > R(1e-8) && f.mu_f * vt < vnm ? R(1) - f.mu_f * vt / vnm : R(0);
      vx[i] = vn * nx * s2; vy[i] = vn * ny * s2; vz[i] = vn * nz * s2;
    } else {
      if (g_mutation != 3) { vx[i] -= vn * nx; vy[i] -= vn * ny; vz[i] -= vn * nz; }
      const R vnm = std::fabs(vn);
      const R tx = vx[i], ty = vy[i], tz = vz[i];
      const R vt = std::sqrt(tx * tx + ty * ty + tz * tz);
      if (vt > R(1e-8)) {
        const R fi = f.mu_f * vnm;
        if (fi >= vt) { vx[i] = 0; vy[i] = 0; vz[i] = 0; }
        else { const R s2 = R(1) - fi / vt; vx[i] = tx * s2; vy[i] = ty * s2; vz[i] = tz * s2; }
      }
    }
    R damp_i = f.cd;
    if (f.adaptive) {
      const R v = std::sqrt(vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]);
      const R q = v / std::max(f.v_thresh, R(1e-6));
      const R tt = std::min(std::max(g_mutation == 7 ? q : R(1) - q, R(0)), R(1));
      damp_i = f.cd + tt * (f.cd_max - f.cd);
    }
    vx[i] *= R(1) - damp_i; vy[i] *= R(1) - damp_i; vz[i] *= R(1) - damp_i;
  }
}

// Global contact damping, pd_solver.rs:1115-1117: state.damp_velocities(0.05)
// (state.rs:233-240, every vertex) when the last detection had any contact.
template <class R> void ipc_contact_damping(int n, long long active, R* vx, R* vy, R* vz) {
  if (active <= 0) return;
  const R f = g_mutation == 5 ? R(0.05) : R(1) - R(0.05);
  for (int i = 0; i < n; ++i) { vx[i] *= f; vy[i] *= f; vz[i] *= f; }
}

struct OracleBase {
  virtual ~OracleBase() {}
  int n = 0, t = 0, h = 0;
  SparseChol chol;
  std::vector<int> Ap, Ai;  // summed matrix pattern, full symmetric CSR
  std::vector<double> Ax;
  std::string err;
  double solve_time = 0.0;
  virtual int step(double dt, or_step_result* res) = 0;
  virtual void get(int which, double* out) const = 0;  // which: 0 pos,1 vel,2 prev,3 pred (3n, SoA)
  virtual void set(int which, const double* in) = 0;
  virtual void masses(double* out) const = 0;
  virtual void local_rhs(double* rhs_out, double* tri_targets, double* hinge_targets) = 0;
  virtual void elements(double* dm, double* area, double* weight) const = 0;
  virtual void hinges(double* rest, double* weight) const = 0;
  virtual void hinge_stencils(double* st) const = 0;
  virtual int step_ipc(double dt, const or_ipc_params& ip, or_step_result* res, or_ipc_stats* st) = 0;
  virtual void get_lambda(double* out) const = 0;
  virtual void set_lambda(const double* in) = 0;
};

template <class R> struct Oracle : OracleBase {
  struct Tri { int i[3]; R dm_inv[4]; R area; R weight; };
  struct Hinge { int v[4]; R rest; R weight; R st[4]; };
  std::vector<Tri> tris;
  std::vector<Hinge> hinges_;
  std::vector<R> mass;  // solver lumped masses (pinned -> 1e8)
  std::vector<R> px, py, pz, vx, vy, vz, qx, qy, qz, rx, ry, rz;  // pos, vel, prev, pred
  std::vector<R> inv_mass;
  or_params P;
  R k_mem, k_bend, dens, dtm, grav[3], damp, rayl;
  bool ground = false;
  R ground_h = 0;

  bool init(int n_, const float* x, const float* y, const float* z, int t_, const uint32_t* tri, int h_,
            const uint32_t* hg, const uint8_t* pinned, const or_params* p) {
    n = n_; t = t_; h = h_; P = *p;
    k_mem = R(p->k_membrane); k_bend = R(p->k_bending); dens = R(p->density); dtm = R(p->dt_matrix);
    for (int k = 0; k < 3; ++k) grav[k] = R(p->gravity[k]);
    damp = R(p->damping); rayl = R(p->rayleigh);
    ground = p->has_ground != 0; ground_h = R(p->ground_height);
    px.assign(x, x + n); py.assign(y, y + n); pz.assign(z, z + n);
    // ElementData::from_mesh — element.rs:54-91
    tris.resize(t);
    for (int e = 0; e < t; ++e) {
      Tri& T = tris[e];
      for (int k = 0; k < 3; ++k) T.i[k] = int(tri[3 * e + k]);
      const V3<R> p0(px[T.i[0]], py[T.i[0]], pz[T.i[0]]);
      const V3<R> p1(px[T.i[1]], py[T.i[1]], pz[T.i[1]]);
      const V3<R> p2(px[T.i[2]], py[T.i[2]], pz[T.i[2]]);
      const V3<R> e1 = p1 - p0, e2 = p2 - p0;
      T.area = R(0.5) * e1.cross(e2).length();
      compute_dm_inv(e1, e2, T.dm_inv);
      T.weight = k_mem * T.area;
    }
    // compute_lumped_masses — pd_solver.rs:1219-1242
    mass.assign(n, R(0));
    for (const Tri& T : tris) {
      const R third = (T.area * dens) / R(3);
      for (int k = 0; k < 3; ++k) mass[T.i[k]] += third;
    }
    for (int i = 0; i < n; ++i) {
      if (pinned[i]) mass[i] = R(1e8);
      else if (mass[i] < R(1e-8)) mass[i] = R(1e-8);
    }
    // BendingData::from_topology — bending.rs:98-148
    hinges_.resize(h);
    const R PI = R(M_PI);
    for (int e = 0; e < h; ++e) {
      Hinge& H = hinges_[e];
      for (int k = 0; k < 4; ++k) H.v[k] = int(hg[4 * e + k]);
      V3<R> q[4];
      for (int k = 0; k < 4; ++k) q[k] = V3<R>(px[H.v[k]], py[H.v[k]], pz[H.v[k]]);
      R rest = dihedral_angle(q[0], q[1], q[2], q[3]);
      if (std::fabs(rest - PI) < R(1e-3)) rest = PI;
      else if (std::fabs(rest) < R(1e-3)) rest = R(0);
      H.rest = rest;
      H.weight = k_bend * (q[1] - q[0]).length();
      H.st[0] = R(-1); H.st[1] = R(-1); H.st[2] = R(1); H.st[3] = R(1);  // assembly.rs:97-101
      if (p->bending_kind == 2) discrete_shells(H, q, p);
    }
    // state — state.rs:85-140
    vx.assign(n, 0); vy.assign(n, 0); vz.assign(n, 0);
    qx = px; qy = py; qz = pz;
    rx.assign(n, 0); ry.assign(n, 0); rz.assign(n, 0);
    inv_mass.assign(n, 0);
    for (int i = 0; i < n; ++i) inv_mass[i] = pinned[i] ? R(0) : (mass[i] > R(1e-12) ? R(1) / mass[i] : R(0));
    // assemble_system_matrix — crates/vistio-solver/src/assembly.rs:44-134.
    // Triplet values are computed in R, then summed in f64 (faer sums the
    // duplicate triplets of CsrMatrix::from_triplets, sparse.rs:49-107).
    struct Trip { int r, c; double v; };
    std::vector<Trip> trip;
    trip.reserve(size_t(n) + 9 * size_t(t) + 16 * size_t(h));
    const R inv_dt2 = R(1) / (dtm * dtm);
    for (int i = 0; i < n; ++i) trip.push_back({i, i, double(mass[i] * inv_dt2)});
    for (const Tri& T : tris) {
      const R a = T.dm_inv[0], b = T.dm_inv[1], c = T.dm_inv[2], d = T.dm_inv[3];
      const R g[3][2] = {{-(a + b), -(c + d)}, {a, c}, {b, d}};
      for (int ia = 0; ia < 3; ++ia)
        for (int ib = 0; ib < 3; ++ib) {
          const R dot = g[ia][0] * g[ib][0] + g[ia][1] * g[ib][1];
          if (std::fabs(dot) > R(1e-12)) trip.push_back({T.i[ia], T.i[ib], double(T.weight * dot)});
        }
    }
    for (const Hinge& H : hinges_)
      for (int ia = 0; ia < 4; ++ia)
        for (int ib = 0; ib < 4; ++ib) {
          const R val = H.weight * H.st[ia] * H.st[ib];  // assembly.rs:96-128
          if (std::fabs(val) > R(1e-12)) trip.push_back({H.v[ia], H.v[ib], double(val)});
        }
    std::sort(trip.begin(), trip.end(), [](const Trip& A, const Trip& B) {
      return A.r < B.r || (A.r == B.r && A.c < B.c);
    });
    Ap.assign(n + 1, 0); Ai.clear(); Ax.clear();
    for (size_t k = 0; k < trip.size();) {
      size_t j = k;
      double s = 0;
      while (j < trip.size() && trip[j].r == trip[k].r && trip[j].c == trip[k].c) { s += trip[j].v; ++j; }
      Ai.push_back(trip[k].c); Ax.push_back(s); Ap[trip[k].r + 1]++;
      k = j;
    }
    for (int i = 0; i < n; ++i) Ap[i + 1] += Ap[i];
    std::vector<double> X(n), Y(n), Z(n);
    for (int i = 0; i < n; ++i) { X[i] = px[i]; Y[i] = py[i]; Z[i] = pz[i]; }
    if (!chol.factor(n, Ap, Ai, Ax, X.data(), Y.data(), Z.data())) { err = "Cholesky factorization failed"; return false; }
    return true;
  }

  // Local step (membrane + bending) and RHS assembly for one axis each,
  // exactly as pd_solver.rs:296-389 and assembly.rs:157-240.
  // DiscreteShellsBendingData::from_topology (discrete_shells.rs:106-195) and
  // the warp/weft modulation of from_topology_with_anisotropic_material
  // (:206-268): normalised cotangent stencil; weight = k * edge length, times
  // k_edge / k_avg for the edge's alignment. A zero-length edge has no element.
  void discrete_shells(Hinge& H, const V3<R> q[4], const or_params* p) {
    const V3<R> &p0 = q[0], &p1 = q[1], &pa = q[2], &pb = q[3];
    const R area_a = R(0.5) * (p1 - p0).cross(pa - p0).length();
    const R area_b = R(0.5) * (p1 - p0).cross(pb - p0).length();
    const V3<R> edge = p1 - p0;
    const R el = edge.length();
    if (el < R(1e-10)) { H.weight = R(0); return; }
    const R inv_e2 = R(1) / (el * el);
    const R c_wa = el / (R(2) * std::max(area_a, R(1e-8)));
    const R c_wb = el / (R(2) * std::max(area_b, R(1e-8)));
    const R dot0_a = (p1 - pa).dot(edge) * inv_e2, dot0_b = (p1 - pb).dot(edge) * inv_e2;
    const R c_0 = -dot0_a * c_wa - dot0_b * c_wb;
    const R dot1_a = (pa - p0).dot(edge) * inv_e2, dot1_b = (pb - p0).dot(edge) * inv_e2;
    const R c_1 = -dot1_a * c_wa - dot1_b * c_wb;
    const R raw[4] = {c_0, c_1, c_wa, c_wb};
    const R nrm = std::sqrt(raw[0] * raw[0] + raw[1] * raw[1] + raw[2] * raw[2] + raw[3] * raw[3]);
    if (nrm > R(1e-10)) {
      for (int k = 0; k < 4; ++k) H.st[k] = raw[k] / nrm;
    } else {
      H.st[0] = R(-0.5); H.st[1] = R(-0.5); H.st[2] = R(0.5); H.st[3] = R(0.5);
    }
    const R kw = R(p->bend_warp), kf = R(p->bend_weft);
    const R avg_raw = (kw + kf) / R(2);
    if (avg_raw > R(1e-8)) {
      const V3<R> dir = edge / el;
      const V3<R> hz(dir.x, R(0), dir.z);
      const R hl = hz.length();
      if (hl >= R(1e-10)) {
        const R c = std::fabs(hz.x / hl);
        const R c2 = c * c, s2 = R(1) - c2;
        H.weight *= (kf * c2 + kw * s2) / avg_raw;
      }
    }
  }

  void local_and_rhs(const R dt, std::vector<R> rhs[3], double* tri_targets, double* hinge_targets) {
    std::vector<V3<R>> tt(size_t(3) * t);
    std::vector<V3<R>> ht(size_t(4) * h);
    const int model = P.model;
    // Parallel where it cannot change a result: per-element projections,
    // then each RHS axis summed serially in the reference's order.
#pragma omp parallel for schedule(static)
    for (int e = 0; e < t; ++e) {
      const Tri& T = tris[e];
      V3<R> p[3];
      for (int k = 0; k < 3; ++k) p[k] = V3<R>(px[T.i[k]], py[T.i[k]], pz[T.i[k]]);
      membrane_project(p[0], p[1], p[2], T.dm_inv, model, &tt[size_t(3) * e], R(P.warp), R(P.weft));
    }
#pragma omp parallel for schedule(static)
    for (int e = 0; e < h; ++e) {
      const Hinge& H = hinges_[e];
      V3<R> q[4];
      for (int k = 0; k < 4; ++k) q[k] = V3<R>(px[H.v[k]], py[H.v[k]], pz[H.v[k]]);
      bending_project(q[0], q[1], q[2], q[3], H.rest, &ht[size_t(4) * e]);
    }
    if (tri_targets)
      for (size_t k = 0; k < tt.size(); ++k) {
        tri_targets[3 * k] = tt[k].x; tri_targets[3 * k + 1] = tt[k].y; tri_targets[3 * k + 2] = tt[k].z;
      }
    if (hinge_targets)
      for (size_t k = 0; k < ht.size(); ++k) {
        hinge_targets[3 * k] = ht[k].x; hinge_targets[3 * k + 1] = ht[k].y; hinge_targets[3 * k + 2] = ht[k].z;
      }
    const R inv_dt2 = R(1) / (dt * dt);
    const std::vector<R>* pred[3] = {&rx, &ry, &rz};
#pragma omp parallel for schedule(static, 1)
    for (int ax = 0; ax < 3; ++ax) {
      std::vector<R>& b = rhs[ax];
      b.assign(n, R(0));
      for (int i = 0; i < n; ++i) b[i] += mass[i] * inv_dt2 * (*pred[ax])[i];
      for (int e = 0; e < t; ++e) {
        const Tri& T = tris[e];
        const R a = T.dm_inv[0], bb = T.dm_inv[1], c = T.dm_inv[2], d = T.dm_inv[3];
        const R g[3][2] = {{-(a + bb), -(c + d)}, {a, c}, {bb, d}};
        R tg[3];
        for (int k = 0; k < 3; ++k) {
          const V3<R>& v = tt[size_t(3) * e + k];
          tg[k] = ax == 0 ? v.x : (ax == 1 ? v.y : v.z);
        }
        const R te1 = tg[1] - tg[0], te2 = tg[2] - tg[0];
        const R f0 = te1 * a + te2 * bb;
        const R f1 = te1 * c + te2 * d;
        for (int k = 0; k < 3; ++k) {
          const R contrib = g[k][0] * f0 + g[k][1] * f1;
          b[T.i[k]] += T.weight * contrib;
        }
      }
      for (int e = 0; e < h; ++e) {
        const Hinge& H = hinges_[e];
        const R* st = H.st;  // [-1, -1, 1, 1] (assembly.rs:210) or the Discrete Shells stencil (:222-236)
        R tg[4];
        for (int k = 0; k < 4; ++k) {
          const V3<R>& v = ht[size_t(4) * e + k];
          tg[k] = ax == 0 ? v.x : (ax == 1 ? v.y : v.z);
        }
        // (0..4).map(|k| s[k]*t[k]).sum(): Rust's float sum starts from -0.0
        R l = R(-0.0);
        for (int k = 0; k < 4; ++k) l = l + st[k] * tg[k];
        for (int k = 0; k < 4; ++k) b[H.v[k]] += H.weight * st[k] * l;
      }
    }
  }

  void local_rhs(double* rhs_out, double* tri_targets, double* hinge_targets) override {
    std::vector<R> rhs[3];
    local_and_rhs(R(P.dt_matrix), rhs, tri_targets, hinge_targets);
    for (int ax = 0; ax < 3; ++ax)
      for (int i = 0; i < n; ++i) rhs_out[size_t(ax) * n + i] = double(rhs[ax][i]);
  }

  // ProjectiveDynamicsSolver::step — pd_solver.rs:259-476.
  int step(double dt_d, or_step_result* res) override {
    const auto t0 = std::chrono::steady_clock::now();
    const R dt = R(dt_d);
    qx = px; qy = py; qz = pz;  // save_previous, state.rs:226-230
    const R dt2 = dt * dt;      // predict, state.rs:145-169
    for (int i = 0; i < n; ++i) {
      if (inv_mass[i] == R(0)) { rx[i] = px[i]; ry[i] = py[i]; rz[i] = pz[i]; continue; }
      rx[i] = px[i] + dt * vx[i] + dt2 * grav[0];
      ry[i] = py[i] + dt * vy[i] + dt2 * grav[1];
      rz[i] = pz[i] + dt * vz[i] + dt2 * grav[2];
      if (ground && ry[i] < ground_h) ry[i] = ground_h;
    }
    px = rx; py = ry; pz = rz;
    std::vector<R> rhs[3];
    std::vector<double> bd3[3] = {std::vector<double>(n), std::vector<double>(n), std::vector<double>(n)};
    std::vector<double> xd3[3] = {std::vector<double>(n), std::vector<double>(n), std::vector<double>(n)};
    std::vector<double> work3[3] = {std::vector<double>(n), std::vector<double>(n), std::vector<double>(n)};
    std::vector<R> sol[3] = {std::vector<R>(n), std::vector<R>(n), std::vector<R>(n)};
    int iterations = 0;
    double final_residual = std::numeric_limits<double>::max();
    for (int iter = 0; iter < P.max_iterations; ++iter) {
      local_and_rhs(dt, rhs, nullptr, nullptr);
      const auto s0 = std::chrono::steady_clock::now();
#pragma omp parallel for schedule(static, 1)
      for (int ax = 0; ax < 3; ++ax) {  // FaerSolver::solve x3, faer_solver.rs:95-129 (independent)
        std::vector<double>& bd = bd3[ax];
        std::vector<double>& xd = xd3[ax];
        for (int i = 0; i < n; ++i) bd[i] = double(rhs[ax][i]);
        chol.solve(bd.data(), xd.data(), work3[ax]);
        for (int i = 0; i < n; ++i) sol[ax][i] = R(xd[i]);
      }
      solve_time += std::chrono::duration<double>(std::chrono::steady_clock::now() - s0).count();
      double diff = 0, norm = 0;  // pd_solver.rs:402-421
      for (int i = 0; i < n; ++i) {
        const double dx = double(sol[0][i] - px[i]), dy = double(sol[1][i] - py[i]), dz = double(sol[2][i] - pz[i]);
        diff += dx * dx + dy * dy + dz * dz;
        const double ox = double(px[i]), oy = double(py[i]), oz = double(pz[i]);
        norm += ox * ox + oy * oy + oz * oz;
      }
      final_residual = norm > 1e-12 ? std::sqrt(diff / norm) : std::sqrt(diff);
      px = sol[0]; py = sol[1]; pz = sol[2];
      for (int i = 0; i < n; ++i)
        if (inv_mass[i] == R(0)) { px[i] = qx[i]; py[i] = qy[i]; pz[i] = qz[i]; }
      if (ground)  // enforce_ground, state.rs:180-189
        for (int i = 0; i < n; ++i) if (inv_mass[i] > R(0) && py[i] < ground_h) py[i] = ground_h;
      iterations = iter + 1;
      if (final_residual < P.tolerance) break;
    }
    const R inv_dt = R(1) / dt;  // update_velocities, state.rs:216-223
    for (int i = 0; i < n; ++i) {
      vx[i] = (px[i] - qx[i]) * inv_dt; vy[i] = (py[i] - qy[i]) * inv_dt; vz[i] = (pz[i] - qz[i]) * inv_dt;
    }
    if (ground)  // enforce_ground_velocities, state.rs:195-213
      for (int i = 0; i < n; ++i)
        if (inv_mass[i] > R(0) && py[i] <= ground_h) {
          if (vy[i] < R(0)) vy[i] = R(0);
          qy[i] = ground_h;
          vx[i] *= R(0.5); vz[i] *= R(0.5);
        }
    const R fdamp = R(1) - damp;  // damp_velocities, state.rs:233-240
    for (int i = 0; i < n; ++i) { vx[i] *= fdamp; vy[i] *= fdamp; vz[i] *= fdamp; }
    if (rayl > R(0)) {  // pd_solver.rs:456-466
      const R f = R(1) / (R(1) + rayl * dt);
      for (int i = 0; i < n; ++i)
        if (inv_mass[i] > R(0)) { vx[i] *= f; vy[i] *= f; vz[i] *= f; }
    }
    res->iterations = iterations;
    res->final_residual = final_residual;
    res->converged = final_residual < P.tolerance;
    res->wall_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return 0;
  }

  // ---- IPC contact step --------------------------------------------------
  std::vector<R> lx, ly, lz;          // SimulationState.al_lambda (state.rs:55-63)
  std::vector<uint32_t> tri_flat;
  std::vector<int> xptr, xidx;        // exclusion sets
  int xdepth = -1;

  void ensure_contact_data(int depth) {
    if (lx.size() != size_t(n)) { lx.assign(n, 0); ly.assign(n, 0); lz.assign(n, 0); }
    if (tri_flat.empty()) {
      tri_flat.resize(3 * size_t(t));
      for (int e = 0; e < t; ++e)
        for (int k = 0; k < 3; ++k) tri_flat[3 * e + k] = uint32_t(tris[e].i[k]);
    }
    if (xdepth != depth) { oc::exclusion_sets(n, t, tri_flat.data(), depth, xptr, xidx); xdepth = depth; }
  }

  // Broad phase of SelfCollisionSystem::detect_ipc_contacts: every pair of
  // f32 triangle boxes (margin = thickness) that overlap.
  std::vector<std::pair<int, int>> pairs_at(const std::vector<R>& X, const std::vector<R>& Y,
                                            const std::vector<R>& Z, double margin) const {
    std::vector<float> fx(X.begin(), X.end()), fy(Y.begin(), Y.end()), fz(Z.begin(), Z.end());
    std::vector<oc::Box32> box(t);
    for (int e = 0; e < t; ++e) box[e] = oc::tri_box(fx.data(), fy.data(), fz.data(), tris[e].i, float(margin));
    return oc::tri_pairs(box);
  }

  // ProjectiveDynamicsSolver::step_with_ipc — pd_solver.rs:549-1157, with
  // CollisionPipeline as the handler (runner.rs:17-34).
  int step_ipc(double dt_d, const or_ipc_params& ip, or_step_result* res, or_ipc_stats* st) override {
    const auto t0 = std::chrono::steady_clock::now();
    ensure_contact_data(ip.exclusion_depth);
    const R dt = R(dt_d);
    oc::IpcScene sc;
    sc.has_ground = ip.has_ground != 0; sc.ground = ip.ground_height;
    sc.has_sphere = ip.has_sphere != 0;
    for (int k = 0; k < 3; ++k) sc.center[k] = ip.sphere_center[k];
    sc.radius = ip.sphere_radius;
    sc.self_collision = ip.self_collision != 0;
    sc.thickness = ip.thickness;
    const bool gnd = sc.has_ground;  // runner.rs:118-122 sets state.ground_height
    const R gh = R(ip.ground_height);
    qx = px; qy = py; qz = pz;  // save_previous
    // d_hat_effective (pd_solver.rs:612-618) drives the CFL clamp; the
    // handler detects with it unless its set_d_hat is the no-op default.
    const R dh_eff = R(ip.compliant_contact ? ip.d_hat * ip.compliant_d_hat_scale : ip.d_hat);
    const R dh = ip.keep_d_hat ? R(ip.d_hat) : dh_eff;
    sc.ccd_edge_edge = ip.ccd_edge_edge != 0;
    const R kappa = R(ip.kappa);
    {  // CFL-like velocity clamp, pd_solver.rs:572-595
      const R max_vel = R(0.5) * std::sqrt(dh_eff) / dt;
      for (int i = 0; i < n; ++i) {
        if (inv_mass[i] == R(0)) continue;
        const R v2 = vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i];
        if (v2 > max_vel * max_vel) {
          const R sc2 = max_vel / std::sqrt(v2);
          vx[i] *= sc2; vy[i] *= sc2; vz[i] *= sc2;
        }
      }
    }
    const R dt2 = dt * dt;  // predict, state.rs:145-169
    for (int i = 0; i < n; ++i) {
      if (inv_mass[i] == R(0)) { rx[i] = px[i]; ry[i] = py[i]; rz[i] = pz[i]; continue; }
      rx[i] = px[i] + dt * vx[i] + dt2 * grav[0];
      ry[i] = py[i] + dt * vy[i] + dt2 * grav[1];
      rz[i] = pz[i] + dt * vz[i] + dt2 * grav[2];
      if (gnd && ry[i] < gh) ry[i] = gh;
    }
    px = rx; py = ry; pz = rz;
    R mu = R(ip.al_mu_initial);
    int total_iterations = 0;
    double final_residual = std::numeric_limits<double>::max();
    bool al_converged = false;
    std::vector<uint8_t> ever(n, 0);
    std::vector<R> fnx(n, 0), fny(n, 0), fnz(n, 0);
    oc::Forces<R> bf, uf;
    uf.reset(n);
    std::vector<std::pair<int, int>> pairs;
    std::vector<R> rhs[3];
    std::vector<double> bd(n), xd(n), work(n);
    std::vector<R> sol[3] = {std::vector<R>(n), std::vector<R>(n), std::vector<R>(n)};
    int al_done = 0;
    R alpha = 1;
    auto detect_at = [&](oc::Forces<R>& F) {
      if (sc.self_collision) pairs = pairs_at(px, py, pz, ip.thickness);
      else pairs.clear();
      oc::detect<R>(n, px.data(), py.data(), pz.data(), t, tri_flat.data(), pairs, xptr, xidx, sc, dh, kappa, F);
      for (int i = 0; i < n; ++i)
        if (F.in_contact[i]) { ever[i] = 1; fnx[i] = F.nx[i]; fny[i] = F.ny[i]; fnz[i] = F.nz[i]; }
    };
    for (int al = 0; al < ip.al_max_iterations; ++al) {
      al_done = al + 1;
      detect_at(bf);
      for (int i = 0; i < n; ++i)
        if (!bf.in_contact[i]) { lx[i] = 0; ly[i] = 0; lz[i] = 0; }
      for (int iter = 0; iter < P.max_iterations; ++iter) {
        local_and_rhs(dt, rhs, nullptr, nullptr);
        const std::vector<R>* G[3] = {&bf.gx, &bf.gy, &bf.gz};
        const std::vector<R>* L[3] = {&lx, &ly, &lz};
        for (int ax = 0; ax < 3; ++ax)  // assemble_barrier_rhs, assembly.rs:272-282
          for (int i = 0; i < n; ++i) rhs[ax][i] -= mu * (*G[ax])[i] + (*L[ax])[i];
        // sol = A^{-1} rhs, formed as pos + A^{-1} (rhs - A pos): the same
        // vector, without the cancellation of sol - pos in the CCD and the
        // step below (their inputs are these small differences).
        const std::vector<R>* Pp[3] = {&px, &py, &pz};
        for (int ax = 0; ax < 3; ++ax) {
          const std::vector<R>& X = *Pp[ax];
          for (int i = 0; i < n; ++i) {
            double ax_i = 0;
            for (int q = Ap[i]; q < Ap[i + 1]; ++q) ax_i += Ax[q] * double(X[Ai[q]]);
            bd[i] = double(rhs[ax][i]) - ax_i;
          }
          chol.solve(bd.data(), xd.data(), work);
          for (int i = 0; i < n; ++i) sol[ax][i] = X[i] + R(xd[i]);
        }
        double diff = 0, norm = 0;  // pd_solver.rs:873-891
        for (int i = 0; i < n; ++i) {
          const double dx = double(sol[0][i] - px[i]), dy = double(sol[1][i] - py[i]), dz = double(sol[2][i] - pz[i]);
          diff += dx * dx + dy * dy + dz * dz;
          norm += double(px[i]) * double(px[i]) + double(py[i]) * double(py[i]) + double(pz[i]) * double(pz[i]);
        }
        final_residual = norm > 1e-12 ? std::sqrt(diff / norm) : std::sqrt(diff);
        // CCD-limited step with the pairs of the last detection, pd_solver.rs:897-906
        const R ccd_alpha = oc::ccd_step<R>(n, px.data(), py.data(), pz.data(), sol[0].data(), sol[1].data(),
                                            sol[2].data(), tri_flat.data(), pairs, sc);
        double gdd = 0;
        for (int i = 0; i < n; ++i)
          if (inv_mass[i] > R(0))
            gdd += double(bf.gx[i] * (sol[0][i] - px[i]) + bf.gy[i] * (sol[1][i] - py[i]) + bf.gz[i] * (sol[2][i] - pz[i]));
        alpha = ipc_step_fraction<R>(ccd_alpha, gdd, mu, diff);
        for (int i = 0; i < n; ++i)  // pd_solver.rs:934-950
          if (inv_mass[i] > R(0)) {
            px[i] += alpha * (sol[0][i] - px[i]);
            py[i] += alpha * (sol[1][i] - py[i]);
            pz[i] += alpha * (sol[2][i] - pz[i]);
          }
        total_iterations++;
        if (final_residual < P.tolerance) break;
      }
      detect_at(uf);  // pd_solver.rs:966-978
      if (uf.max_violation < R(ip.al_tolerance)) { al_converged = true; break; }
      ipc_al_update<R>(n, uf.in_contact.data(), uf.gx.data(), uf.gy.data(), uf.gz.data(), lx.data(), ly.data(),
                       lz.data(), mu, R(ip.al_mu_growth));
    }
    const R inv_dt = R(1) / dt;  // update_velocities
    for (int i = 0; i < n; ++i) {
      vx[i] = (px[i] - qx[i]) * inv_dt; vy[i] = (py[i] - qy[i]) * inv_dt; vz[i] = (pz[i] - qz[i]) * inv_dt;
    }
    const VelFilter<R> vf{R(ip.friction_coefficient), R(ip.contact_damping), R(ip.contact_damping_max),
                          R(ip.contact_velocity_threshold), ip.adaptive_contact_damping != 0};
    ipc_velocity_filter<R>(n, vx.data(), vy.data(), vz.data(), inv_mass.data(), ever.data(), fnx.data(), fny.data(),
                           fnz.data(), vf);
    ipc_contact_damping<R>(n, uf.active, vx.data(), vy.data(), vz.data());
    if (gnd)  // enforce_ground_velocities, state.rs:195-213
      for (int i = 0; i < n; ++i)
        if (inv_mass[i] > R(0) && py[i] <= gh) {
          if (vy[i] < R(0)) vy[i] = R(0);
          qy[i] = gh;
          vx[i] *= R(0.5); vz[i] *= R(0.5);
        }
    const R fdamp = R(1) - damp;
    for (int i = 0; i < n; ++i) { vx[i] *= fdamp; vy[i] *= fdamp; vz[i] *= fdamp; }
    if (rayl > R(0)) {
      const R f = R(1) / (R(1) + rayl * dt);
      for (int i = 0; i < n; ++i)
        if (inv_mass[i] > R(0)) { vx[i] *= f; vy[i] *= f; vz[i] *= f; }
    }
    res->iterations = total_iterations;
    res->final_residual = final_residual;
    res->converged = (al_converged || final_residual < P.tolerance) ? 1 : 0;
    res->wall_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (st) {
      st->al_iterations = al_done; st->al_converged = al_converged ? 1 : 0; st->mu = double(mu);
      st->max_violation = double(uf.max_violation); st->active = uf.active; st->self_contacts = uf.self_contacts;
      st->pairs = (long long)pairs.size(); st->last_alpha = double(alpha);
    }
    return 0;
  }
  void get_lambda(double* out) const override {
    for (int i = 0; i < n; ++i) {
      out[i] = lx.empty() ? 0.0 : double(lx[i]);
      out[n + i] = ly.empty() ? 0.0 : double(ly[i]);
      out[2 * size_t(n) + i] = lz.empty() ? 0.0 : double(lz[i]);
    }
  }
  void set_lambda(const double* in) override {
    lx.assign(n, 0); ly.assign(n, 0); lz.assign(n, 0);
    for (int i = 0; i < n; ++i) { lx[i] = R(in[i]); ly[i] = R(in[n + i]); lz[i] = R(in[2 * size_t(n) + i]); }
  }

  void get(int which, double* out) const override {
    const std::vector<R>* s[4][3] = {{&px, &py, &pz}, {&vx, &vy, &vz}, {&qx, &qy, &qz}, {&rx, &ry, &rz}};
    for (int a = 0; a < 3; ++a)
      for (int i = 0; i < n; ++i) out[size_t(a) * n + i] = double((*s[which][a])[i]);
  }
  void set(int which, const double* in) override {
    std::vector<R>* s[4][3] = {{&px, &py, &pz}, {&vx, &vy, &vz}, {&qx, &qy, &qz}, {&rx, &ry, &rz}};
    for (int a = 0; a < 3; ++a)
      for (int i = 0; i < n; ++i) (*s[which][a])[i] = R(in[size_t(a) * n + i]);
  }
  void masses(double* out) const override { for (int i = 0; i < n; ++i) out[i] = double(mass[i]); }
  void elements(double* dm, double* area, double* weight) const override {
    for (int e = 0; e < t; ++e) {
      for (int k = 0; k < 4; ++k) dm[4 * e + k] = double(tris[e].dm_inv[k]);
      area[e] = double(tris[e].area);
      weight[e] = double(tris[e].weight);
    }
  }
  void hinge_stencils(double* st) const override {
    for (int e = 0; e < h; ++e)
      for (int k = 0; k < 4; ++k) st[4 * e + k] = double(hinges_[e].st[k]);
  }
  void hinges(double* rest, double* weight) const override {
    for (int e = 0; e < h; ++e) { rest[e] = double(hinges_[e].rest); weight[e] = double(hinges_[e].weight); }
  }
};

}  // namespace

extern "C" {

// mode: 0 = f64 oracle, 1 = f32 reference emulation. Returns nullptr on error.
void* or_create(int mode, int n, const float* x, const float* y, const float* z, int t, const uint32_t* tri, int h,
                const uint32_t* hinge4, const uint8_t* pinned, const or_params* p) {
  if (n <= 0 || !p) return nullptr;
  if (mode == 0) {
    auto* o = new Oracle<double>();
    if (!o->init(n, x, y, z, t, tri, h, hinge4, pinned, p)) { delete o; return nullptr; }
    return static_cast<OracleBase*>(o);
  }
  auto* o = new Oracle<float>();
  if (!o->init(n, x, y, z, t, tri, h, hinge4, pinned, p)) { delete o; return nullptr; }
  return static_cast<OracleBase*>(o);
}
void or_destroy(void* o) { delete static_cast<OracleBase*>(o); }
int or_step(void* o, double dt, int nsteps, or_step_result* res) {
  auto* b = static_cast<OracleBase*>(o);
  for (int s = 0; s < nsteps; ++s) b->step(dt, res);
  return 0;
}
void or_get(void* o, int which, double* out) { static_cast<OracleBase*>(o)->get(which, out); }
void or_set(void* o, int which, const double* in) { static_cast<OracleBase*>(o)->set(which, in); }
void or_masses(void* o, double* out) { static_cast<OracleBase*>(o)->masses(out); }
double or_solve_time(void* o) { return static_cast<OracleBase*>(o)->solve_time; }
int or_matrix_nnz(void* o) { return int(static_cast<OracleBase*>(o)->Ai.size()); }
void or_matrix(void* o, int* Ap, int* Ai, double* Ax) {
  auto* b = static_cast<OracleBase*>(o);
  std::copy(b->Ap.begin(), b->Ap.end(), Ap);
  std::copy(b->Ai.begin(), b->Ai.end(), Ai);
  std::copy(b->Ax.begin(), b->Ax.end(), Ax);
}
long long or_factor_nnz(void* o) { return (long long)static_cast<OracleBase*>(o)->chol.Li.size(); }
void or_solve(void* o, const double* b, double* x) {
  auto* s = static_cast<OracleBase*>(o);
  std::vector<double> w(s->n);
  s->chol.solve(b, x, w);
}
// Local step at the current positions with pred as set: RHS (3n, SoA) plus the
// per-element targets (3t x xyz and 4h x xyz, AoS) of ElementData::project*
// and BendingData::project. Uses dt_matrix as dt.
void or_elements(void* o, double* dm, double* area, double* weight) {
  static_cast<OracleBase*>(o)->elements(dm, area, weight);
}
void or_hinges(void* o, double* rest, double* weight) { static_cast<OracleBase*>(o)->hinges(rest, weight); }
void or_local_rhs(void* o, double* rhs, double* tri_targets, double* hinge_targets) {
  static_cast<OracleBase*>(o)->local_rhs(rhs, tri_targets, hinge_targets);
}

// ---- pure functions for known-answer tests (mode 0 = double, 1 = float) ----
// F is 6 values (col0 xyz, col1 xyz). Out: rot (6), stretch (4), s0, s1, v0 (2), v1 (2) = 16 values.
void or_polar(int mode, const double* F, double* out) {
  auto run = [&](auto zero) {
    using R = decltype(zero);
    M32<R> f; f.c0 = V3<R>(R(F[0]), R(F[1]), R(F[2])); f.c1 = V3<R>(R(F[3]), R(F[4]), R(F[5]));
    Polar<R> p = polar_decomposition(f);
    const R v[16] = {p.rot.c0.x, p.rot.c0.y, p.rot.c0.z, p.rot.c1.x, p.rot.c1.y, p.rot.c1.z,
                     p.stretch[0], p.stretch[1], p.stretch[2], p.stretch[3], p.s0, p.s1,
                     p.v0.x, p.v0.y, p.v1.x, p.v1.y};
    for (int k = 0; k < 16; ++k) out[k] = double(v[k]);
  };
  if (mode == 0) run(0.0); else run(0.0f);
}
// Co-rotational (model=1) or ARAP (model=0) target gradient R*S_clamped / R. out: 6 values.
void or_target_f_ex(int mode, int model, const double* F, double warp, double weft, double* out);
void or_target_f(int mode, int model, const double* F, double* out) { or_target_f_ex(mode, model, F, 0.0, 0.0, out); }
void or_target_f_ex(int mode, int model, const double* F, double warp, double weft, double* out) {
  auto run = [&](auto zero) {
    using R = decltype(zero);
    M32<R> f; f.c0 = V3<R>(R(F[0]), R(F[1]), R(F[2])); f.c1 = V3<R>(R(F[3]), R(F[4]), R(F[5]));
    const M32<R> tg = target_of(f, model, R(warp), R(weft));
    const R v[6] = {tg.c0.x, tg.c0.y, tg.c0.z, tg.c1.x, tg.c1.y, tg.c1.z};
    for (int k = 0; k < 6; ++k) out[k] = double(v[k]);
  };
  if (mode == 0) run(0.0); else run(0.0f);
}
// e1, e2: 3 values each. out: dm_inv (4).
void or_dm_inv(int mode, const double* e1, const double* e2, double* out) {
  auto run = [&](auto zero) {
    using R = decltype(zero);
    R d[4];
    compute_dm_inv(V3<R>(R(e1[0]), R(e1[1]), R(e1[2])), V3<R>(R(e2[0]), R(e2[1]), R(e2[2])), d);
    for (int k = 0; k < 4; ++k) out[k] = double(d[k]);
  };
  if (mode == 0) run(0.0); else run(0.0f);
}
// p: 12 values (v0, v1, wa, wb). Returns the dihedral angle.
double or_dihedral(int mode, const double* p) {
  auto run = [&](auto zero) -> double {
    using R = decltype(zero);
    V3<R> q[4];
    for (int k = 0; k < 4; ++k) q[k] = V3<R>(R(p[3 * k]), R(p[3 * k + 1]), R(p[3 * k + 2]));
    return double(dihedral_angle(q[0], q[1], q[2], q[3]));
  };
  return mode == 0 ? run(0.0) : run(0.0f);
}
// p: 12 values; out: 12 values (projected v0, v1, wa, wb).
void or_bend_project(int mode, const double* p, double rest, double* out) {
  auto run = [&](auto zero) {
    using R = decltype(zero);
    V3<R> q[4], o[4];
    for (int k = 0; k < 4; ++k) q[k] = V3<R>(R(p[3 * k]), R(p[3 * k + 1]), R(p[3 * k + 2]));
    bending_project(q[0], q[1], q[2], q[3], R(rest), o);
    for (int k = 0; k < 4; ++k) { out[3 * k] = double(o[k].x); out[3 * k + 1] = double(o[k].y); out[3 * k + 2] = double(o[k].z); }
  };
  if (mode == 0) run(0.0); else run(0.0f);
}
// p: 9 values (p0, p1, p2), dm_inv: 4. out: 9 values (targets).
void or_hinge_stencils(void* o, double* st) { static_cast<OracleBase*>(o)->hinge_stencils(st); }
void or_membrane_project(int mode, int model, const double* p, const double* dm, double* out) {
  auto run = [&](auto zero) {
    using R = decltype(zero);
    V3<R> q[3], o[3];
    for (int k = 0; k < 3; ++k) q[k] = V3<R>(R(p[3 * k]), R(p[3 * k + 1]), R(p[3 * k + 2]));
    const R d[4] = {R(dm[0]), R(dm[1]), R(dm[2]), R(dm[3])};
    membrane_project(q[0], q[1], q[2], d, model, o);
    for (int k = 0; k < 3; ++k) { out[3 * k] = double(o[k].x); out[3 * k + 1] = double(o[k].y); out[3 * k + 2] = double(o[k].z); }
  };
  if (mode == 0) run(0.0); else run(0.0f);
}


// ---- IPC contact step and contact primitives --------------------------------
// Threads of the oracle's parallel loops (1 = the reference's single thread).
void or_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n > 0 ? n : 1);
#else
  (void)n;
#endif
}
int or_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
int or_step_ipc(void* o, double dt, int nsteps, const or_ipc_params* ip, or_step_result* res, or_ipc_stats* st) {
  for (int k = 0; k < nsteps; ++k)
    if (static_cast<OracleBase*>(o)->step_ipc(dt, *ip, res, st)) return 1;
  return 0;
}
void or_get_lambda(void* o, double* out) { static_cast<OracleBase*>(o)->get_lambda(out); }
void or_set_lambda(void* o, const double* in) { static_cast<OracleBase*>(o)->set_lambda(in); }

// The pieces of step_with_ipc that step_ipc calls, in f64, for closed-form
// tests. or_set_mutation plants one of the known mistakes listed at
// g_mutation (0 restores the reference).
void or_set_mutation(int m) { g_mutation = m; }
double or_ipc_step_fraction(double ccd_alpha, double grad_dot_d, double mu, double diff_sq) {
  return ipc_step_fraction<double>(ccd_alpha, grad_dot_d, mu, diff_sq);
}
// g, lambda: 3n SoA; mu in/out.
void or_ipc_al_update(int n, const uint8_t* in_contact, const double* g, double* lambda, double* mu, double growth) {
  ipc_al_update<double>(n, in_contact, g, g + n, g + 2 * size_t(n), lambda, lambda + n, lambda + 2 * size_t(n), *mu,
                        growth);
}
// v, nrm: 3n SoA; v in/out.
void or_ipc_velocity_filter(int n, double* v, const double* inv_mass, const uint8_t* ever, const double* nrm,
                            double mu_f, double cd, int adaptive, double cd_max, double v_thresh) {
  const VelFilter<double> f{mu_f, cd, cd_max, v_thresh, adaptive != 0};
  ipc_velocity_filter<double>(n, v, v + n, v + 2 * size_t(n), inv_mass, ever, nrm, nrm + n, nrm + 2 * size_t(n), f);
}
void or_ipc_contact_damping(int n, long long active, double* v) {
  ipc_contact_damping<double>(n, active, v, v + n, v + 2 * size_t(n));
}

double oc_barrier_energy(double d, double dh) { return oc::barrier_energy(d, dh); }
double oc_barrier_gradient(double d, double dh) { return oc::barrier_gradient(d, dh); }
double oc_barrier_hessian(double d, double dh) { return oc::barrier_hessian(d, dh); }
double oc_scaled_barrier_gradient(double d, double dh, double k) { return oc::scaled_barrier_gradient(d, dh, k); }
double oc_barrier_energy_thick(double d, double dh, double th) { return oc::barrier_energy_thick(d, dh, th); }
double oc_barrier_gradient_thick(double d, double dh, double th) { return oc::barrier_gradient_thick(d, dh, th); }
double oc_estimate_kappa(double c, double m, double g, double dh) { return oc::estimate_kappa(c, m, g, dh); }

static oc::P3<double> p3(const double* v) { return oc::P3<double>(v[0], v[1], v[2]); }
// p = [p, a, b, c] (12 doubles); returns d^2, type
double oc_pt_dist2(const double* p, int* type) { return oc::pt_dist2(p3(p), p3(p + 3), p3(p + 6), p3(p + 9), type); }
void oc_pt_grad(const double* p, double* g) {
  oc::P3<double> G[4];
  oc::pt_grad(p3(p), p3(p + 3), p3(p + 6), p3(p + 9), G);
  for (int k = 0; k < 4; ++k) { g[3 * k] = G[k].x; g[3 * k + 1] = G[k].y; g[3 * k + 2] = G[k].z; }
}
double oc_ee_dist2(const double* p, int* type) { return oc::ee_dist2(p3(p), p3(p + 3), p3(p + 6), p3(p + 9), type); }
void oc_ee_grad(const double* p, double* g) {
  oc::P3<double> G[4];
  oc::ee_grad(p3(p), p3(p + 3), p3(p + 6), p3(p + 9), G);
  for (int k = 0; k < 4; ++k) { g[3 * k] = G[k].x; g[3 * k + 1] = G[k].y; g[3 * k + 2] = G[k].z; }
}
int oc_cubic_roots(double a, double b, double c, double d, double* out) { return oc::cubic_roots(a, b, c, d, out); }
// p = [p0, p1, a0, a1, b0, b1, c0, c1] (24 doubles)
int oc_vt_ccd(const double* p, double* toi) {
  return oc::vt_ccd(p3(p), p3(p + 3), p3(p + 6), p3(p + 9), p3(p + 12), p3(p + 15), p3(p + 18), p3(p + 21), toi) ? 1 : 0;
}
int oc_ee_ccd(const double* p, double* toi) {
  return oc::ee_ccd(p3(p), p3(p + 3), p3(p + 6), p3(p + 9), p3(p + 12), p3(p + 15), p3(p + 18), p3(p + 21), toi) ? 1 : 0;
}
// Candidate triangle pairs (ta < tb, sorted) of f32 positions with margin;
// brute = 1 tests all pairs. Returns the count; writes up to cap pairs.
long long oc_tri_pairs(int n, const float* x, const float* y, const float* z, int t, const uint32_t* tri, float margin,
                       int brute, int* out, long long cap) {
  (void)n;
  std::vector<oc::Box32> box(t);
  for (int e = 0; e < t; ++e) {
    const int v[3] = {int(tri[3 * e]), int(tri[3 * e + 1]), int(tri[3 * e + 2])};
    box[e] = oc::tri_box(x, y, z, v, margin);
  }
  std::vector<std::pair<int, int>> pr;
  if (brute) {
    for (int a = 0; a < t; ++a)
      for (int b = a + 1; b < t; ++b)
        if (oc::overlaps(box[a], box[b])) pr.push_back({a, b});
  } else {
    pr = oc::tri_pairs(box);
  }
  for (long long k = 0; k < (long long)pr.size() && k < cap; ++k) { out[2 * k] = pr[k].first; out[2 * k + 1] = pr[k].second; }
  return (long long)pr.size();
}
// Triangle pairs with no shared vertex that properly cross, at f64
// positions: brute = 1 tests every pair, 0 only the margin-0 box pairs.
long long oc_crossings(int n, const double* pos, int t, const uint32_t* tri, int brute) {
  auto P = [&](uint32_t i) { return oc::P3<double>(pos[i], pos[n + i], pos[2 * (size_t)n + i]); };
  std::vector<std::pair<int, int>> pr;
  if (brute) {
    for (int a = 0; a < t; ++a)
      for (int b = a + 1; b < t; ++b) pr.push_back({a, b});
  } else {
    std::vector<float> x(n), y(n), z(n);
    for (int i = 0; i < n; ++i) { x[i] = (float)pos[i]; y[i] = (float)pos[n + i]; z[i] = (float)pos[2 * (size_t)n + i]; }
    std::vector<oc::Box32> box(t);
    for (int e = 0; e < t; ++e) {
      const int v[3] = {int(tri[3 * e]), int(tri[3 * e + 1]), int(tri[3 * e + 2])};
      box[e] = oc::tri_box(x.data(), y.data(), z.data(), v, 0.0f);
    }
    pr = oc::tri_pairs(box);
  }
  long long hits = 0;
  for (const auto& q : pr) {
    const uint32_t* A = tri + 3 * (size_t)q.first;
    const uint32_t* B = tri + 3 * (size_t)q.second;
    bool shared = false;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) shared |= A[i] == B[j];
    if (shared) continue;
    const oc::P3<double> PA[3] = {P(A[0]), P(A[1]), P(A[2])}, PB[3] = {P(B[0]), P(B[1]), P(B[2])};
    hits += oc::tris_cross(PA, PB) ? 1 : 0;
  }
  return hits;
}
static oc::IpcScene scene_of(const or_ipc_params* ip) {
  oc::IpcScene sc;
  sc.has_ground = ip->has_ground != 0; sc.ground = ip->ground_height;
  sc.has_sphere = ip->has_sphere != 0;
  for (int k = 0; k < 3; ++k) sc.center[k] = ip->sphere_center[k];
  sc.radius = ip->sphere_radius;
  sc.self_collision = ip->self_collision != 0;
  sc.thickness = ip->thickness;
  sc.ccd_edge_edge = ip->ccd_edge_edge != 0;
  return sc;
}
// CollisionPipeline::detect_ipc_contacts at positions (SoA, 3n doubles),
// d_hat and kappa as given. out_grad / out_normal: 3n SoA; in_contact: n.
// stats = [max_violation, active, self_contacts, pairs].
void oc_detect(int n, const double* pos, int t, const uint32_t* tri, const or_ipc_params* ip, double dh, double kappa,
               double* out_grad, double* out_normal, uint8_t* in_contact, double* out_hess, double* stats) {
  const double *x = pos, *y = pos + n, *z = pos + 2 * size_t(n);
  std::vector<int> xp, xi;
  oc::exclusion_sets(n, t, tri, ip->exclusion_depth, xp, xi);
  std::vector<std::pair<int, int>> pairs;
  oc::IpcScene sc = scene_of(ip);
  if (sc.self_collision) {
    std::vector<float> fx(x, x + n), fy(y, y + n), fz(z, z + n);
    std::vector<oc::Box32> box(t);
    for (int e = 0; e < t; ++e) {
      const int v[3] = {int(tri[3 * e]), int(tri[3 * e + 1]), int(tri[3 * e + 2])};
      box[e] = oc::tri_box(fx.data(), fy.data(), fz.data(), v, float(ip->thickness));
    }
    pairs = oc::tri_pairs(box);
  }
  oc::Forces<double> F;
  oc::detect<double>(n, x, y, z, t, tri, pairs, xp, xi, sc, dh, kappa, F);
  for (int i = 0; i < n; ++i) {
    out_grad[i] = F.gx[i]; out_grad[n + i] = F.gy[i]; out_grad[2 * size_t(n) + i] = F.gz[i];
    out_normal[i] = F.nx[i]; out_normal[n + i] = F.ny[i]; out_normal[2 * size_t(n) + i] = F.nz[i];
    in_contact[i] = F.in_contact[i];
    out_hess[i] = F.hd[i];
  }
  stats[0] = F.max_violation; stats[1] = double(F.active); stats[2] = double(F.self_contacts); stats[3] = double(pairs.size());
}
// CollisionPipeline::compute_ccd_step from p0 to p1 (SoA 3n), with the
// candidate pairs of p0 (margin = thickness).
double oc_ccd_step(int n, const double* p0, const double* p1, int t, const uint32_t* tri, const or_ipc_params* ip) {
  oc::IpcScene sc = scene_of(ip);
  std::vector<std::pair<int, int>> pairs;
  if (sc.self_collision) {
    std::vector<float> fx(p0, p0 + n), fy(p0 + n, p0 + 2 * size_t(n)), fz(p0 + 2 * size_t(n), p0 + 3 * size_t(n));
    std::vector<oc::Box32> box(t);
    for (int e = 0; e < t; ++e) {
      const int v[3] = {int(tri[3 * e]), int(tri[3 * e + 1]), int(tri[3 * e + 2])};
      box[e] = oc::tri_box(fx.data(), fy.data(), fz.data(), v, float(ip->thickness));
    }
    pairs = oc::tri_pairs(box);
  }
  return oc::ccd_step<double>(n, p0, p0 + n, p0 + 2 * size_t(n), p1, p1 + n, p1 + 2 * size_t(n), tri, pairs, sc);
}

// VertexTriangleTest::detect (vertex_triangle.rs:25-81). out: up to cap
// contacts of 14 words each (4 uint32 indices as their bits, distance,
// normal[3], barycentric[3], then 3 zero words); returns the total count.
long long oc_narrow_vt(int n, const float* x, const float* y, const float* z, int t, const uint32_t* tri,
                       long long nc, const uint32_t* ca, const uint32_t* cb, float thickness, float* out,
                       long long cap) {
  (void)n;
  const std::vector<oc::VtContact> c = oc::vt_detect(x, y, z, t, tri, nc, ca, cb, thickness);
  for (long long k = 0; k < (long long)c.size() && k < cap; ++k) {
    float* o = out + 14 * k;
    std::memcpy(o, c[k].idx, 16);
    o[4] = c[k].distance;
    for (int q = 0; q < 3; ++q) { o[5 + q] = c[k].normal[q]; o[8 + q] = c[k].bary[q]; }
    o[11] = o[12] = o[13] = 0.0f;
  }
  return (long long)c.size();
}

}  // extern "C"
