// HIP kernels for the gfx950 Projective Dynamics step.
//
// Data layout (all per-vertex arrays in the nested-dissection order chosen
// at set-up, so the solve reads contiguous runs):
//   x, v, prev, pred, r : double [n][3]   (state is FP64 in both precisions)
//   ctri                : T [t][3][3]     membrane contribution per triangle corner
//   chinge              : T [h][3]        w_b * s^T(target - x) per hinge
//   y, d                : T [n][3]        forward result / solution of A d = r
//   U                   : T [sum r][3]    forward update vectors per supernode
//   M                   : T               per-supernode blocks [L^{-1}; L_IP L^{-1}]
//
// The local step is written in residual form: each element emits
// w * G^T (F_target - F) (membrane) or w_b * s (s^T target - s^T x) (bending),
// so the gather produces r = b - A x directly (assembly.rs:157-240 minus A x).
// In exact arithmetic this equals the reference's rhs minus A x; x + A^{-1} r
// is the reference's solve result (pd_solver.rs:381-400).
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>
#include <type_traits>

namespace vhk {

struct SnDev {
  int col0, w, r, rows_off, u_off, map_off, child_off, nchild;
  long long m_off;
  int fm_off, pad;
};

struct StepDev {
  int done;         // set when the relative update fell below the tolerance
  int iterations;   // PD iterations run in the current step
  double residual;  // final_residual of the current step
  unsigned counter; // arrival counter of the commit kernel's last-block reduce
  int pad;
};

__device__ __forceinline__ double3 ld3(const double* p, int i) {
  return make_double3(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
}
__device__ __forceinline__ double3 operator+(double3 a, double3 b) { return make_double3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ double3 operator-(double3 a, double3 b) { return make_double3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ double3 operator*(double3 a, double s) { return make_double3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ double3 ddiv(double3 a, double s) { return make_double3(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ double ddot(double3 a, double3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ double3 dcross(double3 a, double3 b) {
  return make_double3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double dlen(double3 a) { return sqrt(ddot(a, a)); }

__device__ __forceinline__ float3 f3(double3 a) { return make_float3((float)a.x, (float)a.y, (float)a.z); }
__device__ __forceinline__ float3 operator+(float3 a, float3 b) { return make_float3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 operator-(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 operator*(float3 a, float s) { return make_float3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float fdot(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float3 fcross(float3 a, float3 b) {
  return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float flen(float3 a) { return sqrtf(fdot(a, a)); }

// Membrane model (ConstitutiveModel, keyed by name) and its stiffness
// factors, already normalised on the host: orthotropic_linear divides warp
// and weft by their maximum (orthotropic.rs:64-66), anisotropic_corotational
// clamps them to [0, 1] (anisotropic.rs:88-89).
enum MemKind { kArap = 0, kCorot = 1, kIsoLinear = 2, kOrtho = 3, kAniso = 4 };
struct MemModel {
  int kind;
  double kw, kf;
};

// Target singular values of the stretch-based models for sigma = (s0, s1)
// with principal direction v0 = (v0x, v0y) of s0:
//   co-rotational  min(s, 1)                                  corotational.rs:51-72
//   orthotropic    s + (1 - s) (kw, kf)                       orthotropic.rs:68-70
//   anisotropic    s > 1 ? 1 + (s - 1)(1 - k) : s, k0 = kw c^2 + kf s^2,
//                  k1 = kw s^2 + kf c^2 with c^2 = v0x^2       anisotropic.rs:96-126
__device__ __forceinline__ void stretch_targets(const MemModel& m, double s0, double s1, double v0x, double v0y,
                                                double& t0, double& t1) {
  if (m.kind == kOrtho) {
    t0 = s0 + (1.0 - s0) * m.kw;
    t1 = s1 + (1.0 - s1) * m.kf;
  } else if (m.kind == kAniso) {
    const double c2 = v0x * v0x, s2 = v0y * v0y;
    const double k0 = m.kw * c2 + m.kf * s2, k1 = m.kw * s2 + m.kf * c2;
    t0 = s0 > 1.0 ? 1.0 + (s0 - 1.0) * (1.0 - k0) : s0;
    t1 = s1 > 1.0 ? 1.0 + (s1 - 1.0) * (1.0 - k1) : s1;
  } else {
    t0 = fmin(s0, 1.0);
    t1 = fmin(s1, 1.0);
  }
}

// ---------------------------------------------------------------------------
// Membrane projection, FP64, following the reference formulas:
// deformation_gradient (decomposition.rs:147-155), polar_decomposition_3x2
// (decomposition.rs:76-138), clamp_stretch (decomposition.rs:50-66),
// CoRotationalModel::project (corotational.rs:51-72), ARAP target = R
// (element.rs:127-144), degenerate Dm^{-1} -> unchanged positions
// (element.rs:192-202), and the other models' targets (isotropic_linear: I,
// isotropic.rs:46-58; orthotropic / anisotropic: stretch_targets above).
// Returns D = F_target - F.
// ---------------------------------------------------------------------------
__device__ inline void membrane_delta_f64(double3 p0, double3 p1, double3 p2, const double* dm, const MemModel& mm,
                                          double3& D0, double3& D1) {
  const double3 s0v = p1 - p0, s1v = p2 - p0;
  const double3 F0 = s0v * dm[0] + s1v * dm[1];
  const double3 F1 = s0v * dm[2] + s1v * dm[3];
  D0 = make_double3(0, 0, 0);
  D1 = make_double3(0, 0, 0);
  const double det = dm[0] * dm[3] - dm[1] * dm[2];
  if (!(fabs(det) > 1e-12)) return;
  if (mm.kind == kIsoLinear) {
    D0 = make_double3(1.0 - F0.x, -F0.y, -F0.z);
    D1 = make_double3(-F1.x, 1.0 - F1.y, -F1.z);
    return;
  }
  const bool stretch_model = mm.kind != kArap;
  const double a = ddot(F0, F0), b = ddot(F0, F1), d = ddot(F1, F1);
  const double ht = 0.5 * (a + d), hd = 0.5 * (a - d);
  const double disc = sqrt(hd * hd + b * b);
  const double l0 = fmax(ht + disc, 0.0), l1 = fmax(ht - disc, 0.0);
  const double s0 = sqrt(l0), s1 = sqrt(l1);
  const double eps = 1e-10;
  double3 R0, R1;
  double v0x, v0y, v1x, v1y;
  if (s0 < eps) {  // polar_decomposition_3x2's degenerate result: R = I, sigma = 0, V = I
    R0 = make_double3(1, 0, 0); R1 = make_double3(0, 1, 0);
    v0x = 1; v0y = 0; v1x = 0; v1y = 1;
    if (stretch_model) {
      double t0, t1;
      stretch_targets(mm, 0.0, 0.0, 1.0, 0.0, t0, t1);
      D0 = R0 * t0 - F0;
      D1 = R1 * t1 - F1;
    } else {
      D0 = R0 - F0; D1 = R1 - F1;
    }
    return;
  }
  if (fabs(b) > eps) {
    const double ex = l0 - d, ey = b;
    const double rl = 1.0 / sqrt(ex * ex + ey * ey);
    v0x = ex * rl; v0y = ey * rl;
    v1x = -v0y; v1y = v0x;
  } else if (a >= d) {
    v0x = 1; v0y = 0; v1x = 0; v1y = 1;
  } else {
    v0x = 0; v0y = 1; v1x = 1; v1y = 0;
  }
  const double is0 = 1.0 / s0, is1 = s1 > eps ? 1.0 / s1 : 0.0;
  const double si00 = v0x * v0x * is0 + v1x * v1x * is1;
  const double si01 = v0x * v0y * is0 + v1x * v1y * is1;
  const double si11 = v0y * v0y * is0 + v1y * v1y * is1;
  R0 = F0 * si00 + F1 * si01;
  R1 = F0 * si01 + F1 * si11;
  double3 T0 = R0, T1 = R1;
  if (stretch_model) {
    double c0, c1;
    stretch_targets(mm, s0, s1, v0x, v0y, c0, c1);
    const double t00 = v0x * v0x * c0 + v1x * v1x * c1;
    const double t01 = v0x * v0y * c0 + v1x * v1y * c1;
    const double t11 = v0y * v0y * c0 + v1y * v1y * c1;
    T0 = R0 * t00 + R1 * t01;
    T1 = R0 * t01 + R1 * t11;
  }
  D0 = T0 - F0;
  D1 = T1 - F1;
}

// Membrane projection, FP32 residual form. F and E = F^T F - I are formed in
// FP64 from FP64 positions; the eigen-decomposition runs in FP32 on E, so
// sigma - 1 = mu / (1 + sigma) keeps the strain's relative accuracy, and
//   D = F_target - F = F V diag(d_i / sigma_i) V^T,
//   d_i = min(sigma_i, 1) - sigma_i (co-rotational) or 1 - sigma_i (ARAP),
// which equals R S_clamped - F (resp. R - F) in exact arithmetic. The
// eigenvector uses whichever of (mu0 - e_d, b) and (b, mu0 - e_a) avoids
// cancellation (both are the reference's (lambda0 - d, b) direction).
// Degenerate singular values fall back to the FP64 path.
__device__ inline void membrane_delta_f32(double3 p0, double3 p1, double3 p2, const double* dm, const MemModel& mm,
                                          float3& D0, float3& D1) {
  const double3 s0v = p1 - p0, s1v = p2 - p0;
  const double3 F0 = s0v * dm[0] + s1v * dm[1];
  const double3 F1 = s0v * dm[2] + s1v * dm[3];
  D0 = make_float3(0, 0, 0);
  D1 = make_float3(0, 0, 0);
  const double det = dm[0] * dm[3] - dm[1] * dm[2];
  if (!(fabs(det) > 1e-12)) return;
  if (mm.kind == kIsoLinear) {  // I - F, formed in FP64 and rounded once
    D0 = make_float3((float)(1.0 - F0.x), (float)-F0.y, (float)-F0.z);
    D1 = make_float3((float)-F1.x, (float)(1.0 - F1.y), (float)-F1.z);
    return;
  }
  const float ea = (float)(ddot(F0, F0) - 1.0);
  const float eb = (float)ddot(F0, F1);
  const float ed = (float)(ddot(F1, F1) - 1.0);
  const float hd = 0.5f * (ea - ed), ht = 0.5f * (ea + ed);
  const float disc = sqrtf(hd * hd + eb * eb);
  const float mu0 = ht + disc, mu1 = ht - disc;
  const float s0 = sqrtf(fmaxf(1.0f + mu0, 0.0f)), s1 = sqrtf(fmaxf(1.0f + mu1, 0.0f));
  // Small sigma_1 (strong compression): 1 + mu1 loses its relative accuracy
  // in FP32, so such elements take the FP64 path.
  if (s0 < 1e-10f || s1 < 0.25f) {
    double3 E0, E1;
    membrane_delta_f64(p0, p1, p2, dm, mm, E0, E1);
    D0 = f3(E0); D1 = f3(E1);
    return;
  }
  float v0x, v0y;
  if (fabsf(eb) > 1e-10f) {
    float ex, ey;
    if (hd >= 0.0f) { ex = hd + disc; ey = eb; } else { ex = eb; ey = disc - hd; }
    const float rl = 1.0f / sqrtf(ex * ex + ey * ey);
    v0x = ex * rl; v0y = ey * rl;
  } else if (ea >= ed) {
    v0x = 1.0f; v0y = 0.0f;
  } else {
    v0x = 0.0f; v0y = 1.0f;
  }
  const float v1x = -v0y, v1y = v0x;  // sign of v1 is irrelevant in V diag V^T
  const float m0 = mu0 / (1.0f + s0), m1 = mu1 / (1.0f + s1);  // sigma - 1
  // d_i = target_i - sigma_i in terms of m_i = sigma_i - 1 (stretch_targets)
  float d0, d1;
  if (mm.kind == kArap) {
    d0 = -m0; d1 = -m1;
  } else if (mm.kind == kOrtho) {
    d0 = -m0 * (float)mm.kw; d1 = -m1 * (float)mm.kf;
  } else if (mm.kind == kAniso) {
    const float c2 = v0x * v0x, s2 = v0y * v0y;
    const float kw = (float)mm.kw, kf = (float)mm.kf;
    d0 = -fmaxf(m0, 0.0f) * (kw * c2 + kf * s2);
    d1 = -fmaxf(m1, 0.0f) * (kw * s2 + kf * c2);
  } else {
    d0 = -fmaxf(m0, 0.0f); d1 = -fmaxf(m1, 0.0f);
  }
  const float k0 = d0 / s0, k1 = d1 / s1;
  const float K00 = v0x * v0x * k0 + v1x * v1x * k1;
  const float K01 = v0x * v0y * k0 + v1x * v1y * k1;
  const float K11 = v0y * v0y * k0 + v1y * v1y * k1;
  const float3 f0 = f3(F0), f1 = f3(F1);
  D0 = f0 * K00 + f1 * K01;
  D1 = f0 * K01 + f1 * K11;
}

// Dihedral bending, FP64, following compute_dihedral_angle
// (bending.rs:56-91), BendingData::project (bending.rs:182-227) and
// rotate_around_axis (bending.rs:231-239); the Discrete Shells projection
// is the same rotation (discrete_shells.rs:285-340). Returns s^T(target - x)
// = s_a (t_a - x_a) + s_b (t_b - x_b) (the edge vertices are not moved);
// the dihedral stencil has s_a = s_b = 1.
__device__ inline double3 bending_delta_f64(double3 p0, double3 p1, double3 pa, double3 pb, double rest, double sa_,
                                           double sb_) {
  const double3 zero = make_double3(0, 0, 0);
  const double3 edge = p1 - p0;
  const double el = dlen(edge);
  double cur = M_PI;
  double3 axis = zero;
  if (el >= 1e-10) {
    axis = ddiv(edge, el);
    const double3 ta = pa - p0, tb = pb - p0;
    double3 qa = ta - axis * ddot(ta, axis);
    double3 qb = tb - axis * ddot(tb, axis);
    const double la = dlen(qa), lb = dlen(qb);
    if (!(la < 1e-10 || lb < 1e-10)) {
      qa = ddiv(qa, la); qb = ddiv(qb, lb);
      const double c = fmin(1.0, fmax(-1.0, ddot(qa, qb)));
      cur = acos(c);
    }
  }
  const double diff = cur - rest;
  if (fabs(diff) < 1e-8) return zero;
  if (el < 1e-10) return zero;
  const double3 ta = pa - p0, tb = pb - p0;
  const double3 qa = ta - axis * ddot(ta, axis);
  const double3 qb = tb - axis * ddot(tb, axis);
  const double sign = ddot(dcross(qa, qb), axis) >= 0.0 ? 1.0 : -1.0;
  const double half = diff * 0.5 * sign;
  const double3 mid = (p0 + p1) * 0.5;
  double sa, ca;
  sincos(half, &sa, &ca);
  const double3 va = pa - mid, vb = pb - mid;
  const double3 ra = va * ca + dcross(axis, va) * sa + axis * ddot(axis, va) * (1.0 - ca);
  const double3 rb = vb * ca + dcross(axis, vb) * (-sa) + axis * ddot(axis, vb) * (1.0 - ca);
  return ((mid + ra) - pa) * sa_ + ((mid + rb) - pb) * sb_;
}

// Dihedral bending, FP32 residual form: vectors relative to v0 (formed in
// FP64, rounded once), angle by atan2(|qa x qb|, qa . qb) (same value as the
// reference's acos of the normalised dot, without its loss near pi), and the
// displacement of each wing as sin(h) (axis x v) - 2 sin^2(h/2) v_perp.
__device__ inline float3 bending_delta_f32(double3 p0, double3 p1, double3 pa, double3 pb, float rest, float sa_,
                                          float sb_) {
  const float3 zero = make_float3(0, 0, 0);
  const float3 e = f3(p1 - p0), a = f3(pa - p0), b = f3(pb - p0);
  const float el = flen(e);
  if (!(el >= 1e-10f)) return zero;
  const float3 axis = e * (1.0f / el);
  const float3 qa = a - axis * fdot(a, axis);
  const float3 qb = b - axis * fdot(b, axis);
  const float3 cr = fcross(qa, qb);
  float cur = (float)M_PI;
  if (!(flen(qa) < 1e-10f || flen(qb) < 1e-10f)) cur = atan2f(flen(cr), fdot(qa, qb));
  const float diff = cur - rest;
  if (fabsf(diff) < 1e-8f) return zero;
  const float sign = fdot(cr, axis) >= 0.0f ? 1.0f : -1.0f;
  const float half = diff * 0.5f * sign;
  const float sh = sinf(half);
  const float q = sinf(0.5f * half);
  const float omc = 2.0f * q * q;  // 1 - cos(half)
  const float3 m = e * 0.5f;
  const float3 va = a - m, vb = b - m;
  const float3 da = fcross(axis, va) * sh - (va - axis * fdot(axis, va)) * omc;
  const float3 db = fcross(axis, vb) * (-sh) - (vb - axis * fdot(axis, vb)) * omc;
  return da * sa_ + db * sb_;
}

template <class T> struct Prec;
template <> struct Prec<double> { static constexpr bool f32 = false; };
template <> struct Prec<float> { static constexpr bool f32 = true; };

// Element data of the local step (k_local: one thread per element, corners
// written to their slots).
template <class T>
struct ElemArgs {
  MemModel mm;
  const double* x;
  const int* tri_idx;
  const double* tri_dm;
  const T* tri_w;
  const int* hinge_idx;
  const T* hinge_rest;
  const T* hinge_w;
  const T* hinge_st;  // null: dihedral springs
};

// Triangle e: w and its two D columns (the membrane projection's F_target - F).
template <class T>
__device__ __forceinline__ void tri_delta(const ElemArgs<T>& ea, int e, T& w, T (&g)[3][2], T (&D)[2][3]) {
  const int i0 = ea.tri_idx[3 * e], i1 = ea.tri_idx[3 * e + 1], i2 = ea.tri_idx[3 * e + 2];
  const double3 p0 = ld3(ea.x, i0), p1 = ld3(ea.x, i1), p2 = ld3(ea.x, i2);
  double dm[4];
  for (int k = 0; k < 4; ++k) dm[k] = ea.tri_dm[4 * e + k];
  w = ea.tri_w[e];
  const T a = (T)dm[0], b = (T)dm[1], c = (T)dm[2], d = (T)dm[3];
  g[0][0] = -(a + b); g[0][1] = -(c + d);
  g[1][0] = a; g[1][1] = c;
  g[2][0] = b; g[2][1] = d;
  if (Prec<T>::f32) {
    float3 D0, D1;
    membrane_delta_f32(p0, p1, p2, dm, ea.mm, D0, D1);
    D[0][0] = D0.x; D[0][1] = D0.y; D[0][2] = D0.z; D[1][0] = D1.x; D[1][1] = D1.y; D[1][2] = D1.z;
  } else {
    double3 D0, D1;
    membrane_delta_f64(p0, p1, p2, dm, ea.mm, D0, D1);
    D[0][0] = D0.x; D[0][1] = D0.y; D[0][2] = D0.z; D[1][0] = D1.x; D[1][1] = D1.y; D[1][2] = D1.z;
  }
}

// Corner k of triangle e: w g_k^T D (assembly.rs:176-205 minus A x).
template <class T>
__device__ __forceinline__ void tri_corner(T w, const T (&g)[3][2], const T (&D)[2][3], int k, T& ox, T& oy, T& oz) {
  const T g0 = k == 0 ? g[0][0] : (k == 1 ? g[1][0] : g[2][0]);
  const T g1 = k == 0 ? g[0][1] : (k == 1 ? g[1][1] : g[2][1]);
  ox = w * (g0 * D[0][0] + g1 * D[1][0]);
  oy = w * (g0 * D[0][1] + g1 * D[1][1]);
  oz = w * (g0 * D[0][2] + g1 * D[1][2]);
}

// Hinge e: its stencil and w_b (s^T target - s^T x).
template <class T>
__device__ __forceinline__ void hinge_delta(const ElemArgs<T>& ea, int e, T (&sg)[4], T& cx, T& cy, T& cz) {
  const int* q = ea.hinge_idx + 4 * (size_t)e;
  const double3 p0 = ld3(ea.x, q[0]), p1 = ld3(ea.x, q[1]), pa = ld3(ea.x, q[2]), pb = ld3(ea.x, q[3]);
  const T w = ea.hinge_w[e];
  // stencil s: (-1, -1, 1, 1) for dihedral springs, else the hinge's
  // Discrete Shells stencil
  sg[0] = T(-1); sg[1] = T(-1); sg[2] = T(1); sg[3] = T(1);
  if (ea.hinge_st)
    for (int k = 0; k < 4; ++k) sg[k] = ea.hinge_st[4 * (size_t)e + k];
  if (Prec<T>::f32) {
    const float3 c = bending_delta_f32(p0, p1, pa, pb, (float)ea.hinge_rest[e], (float)sg[2], (float)sg[3]);
    cx = c.x; cy = c.y; cz = c.z;
  } else {
    const double3 c = bending_delta_f64(p0, p1, pa, pb, (double)ea.hinge_rest[e], (double)sg[2], (double)sg[3]);
    cx = c.x; cy = c.y; cz = c.z;
  }
  cx *= w; cy *= w; cz *= w;
}

// Local step: one thread per triangle (blocks [0, tb)) or hinge (the rest).
// pd_solver.rs:300-367 (the loops) and the per-element RHS terms of
// assembly.rs:176-235. Each element corner's contribution goes to its own
// slot of the vertex-sorted slot array (a hinge's contribution enters with
// the stencil weight s[k]: (-1, -1, +1, +1) for dihedral springs,
// assembly.rs:210, or the hinge's Discrete Shells stencil, :222-236), so the RHS
// gather is a contiguous segmented sum with no atomics.
//
// Block order: workgroups go to the eight XCDs in turn, and XCD x takes the
// x-th eighth of the triangle chunks (256 elements each, sorted by their
// smallest vertex label) and the x-th eighth of the hinge chunks,
// interleaved in proportion (Bresenham: chunk j of the XCD's run is a
// triangle chunk when floor((j+1) tper / (tper+hper)) steps). Both eighths
// cover about the same vertices, so the slots of a vertex, written by its
// triangles and its hinges, are written by one XCD at about the same time
// and complete in its L2. The map is arithmetic: no load before the
// element indices.
struct LocalMap {
  int tper, hper;  // triangle / hinge chunks per XCD (the grid is 8 (tper + hper) blocks)
};
constexpr int kLocalXcds = 8;

// Element chunk of block b: >= 0 triangle chunk, < 0 hinge chunk ~c,
// kIdleChunk for none.
constexpr int kIdleChunk = -0x7fffffff - 1;
__device__ __forceinline__ int local_chunk(const LocalMap& m, int b, int tb, int hb) {
  const int x = b % kLocalXcds, j = b / kLocalXcds, run = m.tper + m.hper;
  if (j >= run) return kIdleChunk;
  const int tj = (int)((long long)j * m.tper / run), tn = (int)((long long)(j + 1) * m.tper / run);
  if (tn > tj) {
    const int c = x * m.tper + tj;
    return c < tb ? c : kIdleChunk;
  }
  const int c = x * m.hper + (j - tj);
  return c < hb ? ~c : kIdleChunk;
}

// The RHS slots of each local-step block (see vh::Setup::lv_ptr): chunk c's
// local vertices [lv_ptr[c], lv_ptr[c + 1]), local vertex j's slot lv_slot[j]
// and the end lv_end[j] of its corners in lc_pos (from the chunk's first
// corner), each corner given by its place in the block (element * K + k).
struct LocalPlan {
  const int* lv_ptr;
  const int* lv_slot;
  const unsigned short* lv_end;
  const unsigned short* lc_pos;
};

// One block per chunk of 256 triangles or hinges: each thread computes its
// element's corner contributions into LDS; then each local vertex's corners
// are summed in a fixed order (ascending place in the block) and written to
// its slot, one slot per (block, vertex) instead of one per corner. The
// plan entries of a thread's first local vertex are loaded with the
// element's indices.
template <class T>
__global__ void __launch_bounds__(256) k_local(int t, int h, LocalMap lm, ElemArgs<T> ea, LocalPlan lp,
                                               T* __restrict__ cslot, const StepDev* st) {
  if (st && st->done) return;
  const int tb = (t + 255) / 256;
  const int ch = local_chunk(lm, (int)blockIdx.x, tb, (h + 255) / 256);
  if (ch == kIdleChunk) return;  // uniform over the block
  __shared__ T cv[256 * 4 * 3];
  const bool is_tri = ch >= 0;
  const int c = is_tri ? ch : tb + ~ch;
  const int lv0 = lp.lv_ptr[c], lv1 = lp.lv_ptr[c + 1];
  const size_t cbase = is_tri ? (size_t)3 * 256 * ch : (size_t)3 * t + (size_t)4 * 256 * (~ch);
  const unsigned short* pos = lp.lc_pos + cbase;
  constexpr int kPre = 8;
  const int j0 = lv0 + (int)threadIdx.x;
  int sl0 = 0, a0 = 0, b0 = 0;
  if (j0 < lv1) {
    sl0 = lp.lv_slot[j0];
    a0 = j0 > lv0 ? lp.lv_end[j0 - 1] : 0;
    b0 = lp.lv_end[j0];
  }
  int pre[kPre];
#pragma unroll
  for (int u = 0; u < kPre; ++u) pre[u] = a0 + u < b0 ? pos[a0 + u] : 0;
  if (is_tri) {
    const int e = ch * 256 + (int)threadIdx.x;
    if (e < t) {
      T w, g[3][2], D[2][3];
      tri_delta(ea, e, w, g, D);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        T* o = cv + 3 * (3 * threadIdx.x + k);
        tri_corner(w, g, D, k, o[0], o[1], o[2]);
      }
    }
  } else {
    const int e = (~ch) * 256 + (int)threadIdx.x;
    if (e < h) {
      T sg[4], cx, cy, cz;
      hinge_delta(ea, e, sg, cx, cy, cz);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        T* o = cv + 3 * (4 * threadIdx.x + k);
        o[0] = sg[k] * cx; o[1] = sg[k] * cy; o[2] = sg[k] * cz;
      }
    }
  }
  __syncthreads();
  if (j0 < lv1) {
    T ax = 0, ay = 0, az = 0;
#pragma unroll
    for (int u = 0; u < kPre; ++u)
      if (a0 + u < b0) {
        const T* v = cv + 3 * pre[u];
        ax += v[0]; ay += v[1]; az += v[2];
      }
    for (int q = a0 + kPre; q < b0; ++q) {
      const T* v = cv + 3 * pos[q];
      ax += v[0]; ay += v[1]; az += v[2];
    }
    T* o = cslot + 3 * (size_t)sl0;
    o[0] = ax; o[1] = ay; o[2] = az;
  }
  // chunks that touch more than 256 vertices
  for (int j = j0 + 256; j < lv1; j += 256) {
    const int a = lp.lv_end[j - 1], b = lp.lv_end[j];
    T ax = 0, ay = 0, az = 0;
    for (int q = a; q < b; ++q) {
      const T* v = cv + 3 * pos[q];
      ax += v[0]; ay += v[1]; az += v[2];
    }
    T* o = cslot + 3 * (size_t)lp.lv_slot[j];
    o[0] = ax; o[1] = ay; o[2] = az;
  }
}

// Inputs of the RHS r = b - A x at one vertex: the inertia term
// m_i (pred_i / dt^2 - x_i / dt_A^2), an optional per-vertex external term
// (the contact step's -(mu g + lambda), assembly.rs:272-282) and the
// vertex's contiguous slots.
template <class T>
struct RhsSrc {
  const int* inc_ptr;
  const T* cslot;
  const double* mass;
  const double* pred;
  const double* x;
  double inv_dt2, inv_dtm2;
  const double* ext;  // [n][3] or null
};

template <class T>
__device__ __forceinline__ double3 inertia(const RhsSrc<T>& s, int i) {
  const double m = s.mass[i];
  double3 r = make_double3(m * (s.inv_dt2 * s.pred[3 * i + 0] - s.inv_dtm2 * s.x[3 * i + 0]),
                           m * (s.inv_dt2 * s.pred[3 * i + 1] - s.inv_dtm2 * s.x[3 * i + 1]),
                           m * (s.inv_dt2 * s.pred[3 * i + 2] - s.inv_dtm2 * s.x[3 * i + 2]));
  if (s.ext) { r.x += s.ext[3 * i + 0]; r.y += s.ext[3 * i + 1]; r.z += s.ext[3 * i + 2]; }
  return r;
}

template <class T>
__device__ __forceinline__ double3 rhs_from_slots(const RhsSrc<T>& s, int i) {
  T ax = 0, ay = 0, az = 0;
  const int pe = s.inc_ptr[i + 1];
  int p = s.inc_ptr[i];
#pragma unroll 4
  for (; p < pe; ++p) {
    const T* c = s.cslot + 3 * (size_t)p;
    ax += c[0]; ay += c[1]; az += c[2];
  }
  const double3 in = inertia(s, i);
  return make_double3(in.x + (double)ax, in.y + (double)ay, in.z + (double)az);
}

// Sum over the G lanes of a lane group (xor butterfly, every lane gets the
// total); the three components are interleaved so their shuffles overlap.
template <int G, class T>
__device__ __forceinline__ void sum3(T& x, T& y, T& z) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) {
    const T a = __shfl_xor(x, o, G), b = __shfl_xor(y, o, G), c = __shfl_xor(z, o, G);
    x += a; y += b; z += c;
  }
}

// Calls f(std::integral_constant<int, G>) for the runtime lane count g, one
// of the powers of two in [lo, 64].
template <int lo, class F>
__device__ __forceinline__ void with_lanes(int g, F&& f) {
  switch (g) {
    case 1: if constexpr (lo <= 1) f(std::integral_constant<int, 1>{}); break;
    case 2: if constexpr (lo <= 2) f(std::integral_constant<int, 2>{}); break;
    case 4: if constexpr (lo <= 4) f(std::integral_constant<int, 4>{}); break;
    case 8: f(std::integral_constant<int, 8>{}); break;
    case 16: f(std::integral_constant<int, 16>{}); break;
    case 32: f(std::integral_constant<int, 32>{}); break;
    case 64: f(std::integral_constant<int, 64>{}); break;
    default: break;
  }
}

// The slot sum of vertex i split over the L lanes of a group (all lanes of
// the group call it with the same i and get the total). Four slot loads per
// lane are issued together (clamped addresses, masked sums).
template <int L, class T>
__device__ __forceinline__ double3 rhs_group(const RhsSrc<T>& s, int i, int lane) {
  // the inertia inputs are loaded with the slot range, then U slots per
  // lane at a time (8 in FP32, 4 in FP64), each lane summing its slots in
  // order
  constexpr int U = sizeof(T) == 4 ? 8 : 4;
  T ax = 0, ay = 0, az = 0;
  const int p0 = s.inc_ptr[i], pe = s.inc_ptr[i + 1];
  const double3 in = inertia(s, i);
  for (int p = p0 + lane; p < pe; p += U * L) {
    T v[U][3];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int q = p + u * L;
      const T* c = s.cslot + 3 * (size_t)(q < pe ? q : p);
      v[u][0] = c[0]; v[u][1] = c[1]; v[u][2] = c[2];
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (p + u * L < pe) { ax += v[u][0]; ay += v[u][1]; az += v[u][2]; }
  }
  sum3<L>(ax, ay, az);
  return make_double3(in.x + (double)ax, in.y + (double)ay, in.z + (double)az);
}

// Stand-alone RHS gather (residual inspection entry point).
template <class T>
__global__ void __launch_bounds__(256) k_gather(int n, RhsSrc<T> s, double* __restrict__ r) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double3 v = rhs_from_slots(s, i);
  r[3 * i + 0] = v.x; r[3 * i + 1] = v.y; r[3 * i + 2] = v.z;
}

// Row kernels of the solve: a stored row (len entries, contiguous) times an
// LDS vector of xyz triples, split over the G lanes of a group; lane l owns
// entries l, l + G, l + 2G, ...

// Loads entries lane + i G (i < N) of a row into registers, zero past len.
// Every lane loads (clamped address), so the N loads are issued together.
template <int G, int N, class T>
__device__ __forceinline__ void load_row(T (&mv)[N], const T* __restrict__ row, int len, int lane) {
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const int k = lane + i * G;
    const T v = __builtin_nontemporal_load(row + (k < len ? k : 0));
    mv[i] = k < len ? v : T(0);
  }
}

// acc += sum_i mv[i] vec[lane + i G] for a row of len >= 1 entries (the
// index is clamped past len, where mv is zero).
template <int G, int N, class T>
__device__ __forceinline__ void fma_row(const T (&mv)[N], const T* __restrict__ vec, int len, int lane, T& ax, T& ay,
                                        T& az) {
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const int k0 = lane + i * G;
    const T* v = vec + 3 * (k0 < len ? k0 : len - 1);
    ax += mv[i] * v[0]; ay += mv[i] * v[1]; az += mv[i] * v[2];
  }
}

// Dot product of a row (len >= 1) whose first N G entries are in mv; longer
// rows continue 8 loads per lane at a time. Every lane gets the sums.
template <int G, int N, class T>
__device__ __forceinline__ void dot_row(const T (&mv)[N], const T* __restrict__ row, int len,
                                        const T* __restrict__ vec, int lane, T& sx, T& sy, T& sz) {
  T ax = 0, ay = 0, az = 0;
  fma_row<G, N>(mv, vec, len, lane, ax, ay, az);
  for (int base = N * G; base < len; base += 8 * G) {
    T m2[8];
    load_row<G, 8>(m2, row + base, len - base, lane);
    fma_row<G, 8>(m2, vec + 3 * base, len - base, lane, ax, ay, az);
  }
  sum3<G>(ax, ay, az);
  sx = ax; sy = ay; sz = az;
}

// Solve task descriptor: everything a task needs in one 48-byte record.
//   q0 = (col0, w, r, m_off)   supernode columns, off-diagonal rows, block offset
//   q1 = (b, e, g, u_off)      row range, lanes per row, update-row / v-slot base
//   q2 = (fm_off, K, 0, 0)     first front slot, slots per front position
struct TaskDev {
  int4 q0, q1, q2;
};

// Commit of one PD iteration for one vertex, pd_solver.rs:402-438: x +=
// delta, pinned -> prev, ground clamp; returns |delta|^2 and |x_old|^2.
struct CommitArgs {
  int on;           // commit in this launch
  int has_ground;
  double ground;
  double* x;
  const double* prev;
  const uint8_t* pinned;
  double* partials;  // per-block (diff, norm), when want_resid
  int want_resid;
  const double* acc;  // earlier refinement sweeps' correction, added when add_acc
  int add_acc;
};

__device__ __forceinline__ void commit_vertex(const CommitArgs& c, int i, double dx, double dy, double dz,
                                              double& df, double& nm) {
  const double ox = c.x[3 * i + 0], oy = c.x[3 * i + 1], oz = c.x[3 * i + 2];
  df += dx * dx + dy * dy + dz * dz;
  nm += ox * ox + oy * oy + oz * oz;
  double px = ox + dx, py = oy + dy, pz = oz + dz;
  if (c.pinned[i]) { px = c.prev[3 * i + 0]; py = c.prev[3 * i + 1]; pz = c.prev[3 * i + 2]; }
  else if (c.has_ground && py < c.ground) py = c.ground;
  c.x[3 * i + 0] = px; c.x[3 * i + 1] = py; c.x[3 * i + 2] = pz;
}

// Block sum of (df, nm) into partials[blockIdx.x] (fixed order).
__device__ inline void block_partials(double df, double nm, double* partials) {
  __shared__ double sd[256], sn[256];
  sd[threadIdx.x] = df; sn[threadIdx.x] = nm;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) { sd[threadIdx.x] += sd[threadIdx.x + s]; sn[threadIdx.x] += sn[threadIdx.x + s]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) { partials[2 * blockIdx.x] = sd[0]; partials[2 * blockIdx.x + 1] = sn[0]; }
}

template <class T>
struct BandArgs {
  const TaskDev* tasks;
  int ntasks;
  const int* uslot;
  const int* dpush_ptr;
  const int* dpush_idx;
  // RHS sources: the slots (read by the first launch of a sweep that starts
  // from them, template flag S, which also stores r) or r
  RhsSrc<T> src;
  double* r;
  // extra blocks after the tasks: RHS (forward) or commit (backward) of the
  // columns outside the deepest band
  const int* upper_cols;
  int nupper;
  // Task blocks: 0 = block b runs task b; else the tasks are cut into one
  // contiguous run of xper tasks per XCD (workgroups go to the eight XCDs
  // in turn), so the tasks of one supernode, which gather the same front,
  // share an L2; the grid then has 8 xper task blocks, the last few idle.
  int xper;
};

constexpr int kXcds = 8;

template <class T>
__device__ __forceinline__ int task_blocks(const BandArgs<T>& a) {
  return a.xper ? kXcds * a.xper : a.ntasks;
}
// The task of task block b, or -1 for an idle block of the padding.
template <class T>
__device__ __forceinline__ int task_of(const BandArgs<T>& a, int b) {
  if (!a.xper) return b;
  const int i = (b % kXcds) * a.xper + b / kXcds;
  return i < a.ntasks ? i : -1;
}

// Forward task (L y = r): rows [b, e) of supernode P's block
//   M_P = [L_PP^{-1}; L_IP L_PP^{-1}] (row-major, row i < w has i+1 entries).
// front[j] = r_P[j] - (sum of the slots the children pushed for position j),
// pass-through front[w + q] = slot sum only; then
//   rows i < w : y_P[i]            = sum_k M[i][k] t[k]
//   rows w + q : slot(U_P[q])      = sum_k M[w+q][k] t[k] + front[w + q]
// so each update row lands in its slot of the parent's front. Only the front
// entries the task's rows read are assembled.
// Each lane group of G lanes takes rows b + grp, b + grp + ngrp, ...; the
// first pref_rows<G>() of them (pref_loads<G>() entries per lane each) and
// their output slots are loaded before the front is assembled. Two rows of
// 8 loads: four rows cost more in occupancy than they save in front
// re-reads. Groups of 64 lanes (rows over 256 entries, one row per group
// in a task) take one row of 16 loads instead: the same registers, and a
// row of up to 1024 entries is in flight before the barrier.
template <int G>
constexpr int pref_rows() { return G == 64 ? 1 : 2; }
template <int G>
constexpr int pref_loads() { return G == 64 ? 16 : 8; }

__device__ __forceinline__ long long fwd_row_off(int row, int w, long long tri) {
  return row < w ? (long long)row * (row + 1) / 2 : tri + (long long)(row - w) * w;
}

// Output of forward row `row`: y (rows < w) or the update slot us of the
// parent's front (rows >= w); lane k < 3 stores component k.
template <class T>
__device__ __forceinline__ void fwd_store(int row, int w, int col0, int us, T sx, T sy, T sz, const T* front,
                                          T* __restrict__ y, T* __restrict__ F, int lane) {
  if (lane < 3) {
    const T s = lane == 0 ? sx : (lane == 1 ? sy : sz);
    if (row < w) y[3 * (size_t)(col0 + row) + lane] = s;
    else F[3 * (size_t)us + lane] = s + front[3 * row + lane];
  }
}

template <class T>
__device__ __forceinline__ void front_sum(const T* __restrict__ f, int K, T& ax, T& ay, T& az) {
  ax = 0; ay = 0; az = 0;
  if (K <= 0) return;
  // the first four slots loaded together (clamped), then summed in order
  T v[4][3];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const T* s = f + 3 * (q < K ? q : 0);
    v[q][0] = s[0]; v[q][1] = s[1]; v[q][2] = s[2];
  }
#pragma unroll
  for (int q = 0; q < 4; ++q)
    if (q < K) { ax += v[q][0]; ay += v[q][1]; az += v[q][2]; }
  for (int k = 4; k < K; ++k) { ax += f[3 * k]; ay += f[3 * k + 1]; az += f[3 * k + 2]; }
}

// Front assembly with every load issued before any is waited on. Loads in
// flight complete in issue order as far as the wait counter goes, so a
// front gathered after the block rows would wait for the rows at its first
// sum, and a slot-by-slot sum would wait once per slot. Here each thread
// loads, for its positions j = j0 + tid + p blockDim.x (p < P), up to four
// child slots (clamped addresses, masked sums) and, with R, r; the caller
// then issues its row loads, and front_finish sums in the order of
// front_sum: r - ((0 + s0) + s1 + ...), or the slot sum alone without R.
template <int P, class T>
struct FrontLoads {
  T f[P][4][3];
  double rb[P][3];
};

template <int P, bool R, class T>
__device__ __forceinline__ void front_issue(FrontLoads<P, T>& L, const double* __restrict__ r, int col0, int j0,
                                            int n, const T* __restrict__ Fp, int K) {
  if (n <= j0) return;  // no positions (uniform): nothing to load
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int j = j0 + (int)threadIdx.x + p * (int)blockDim.x;
    const int jj = j < n ? j : j0;
    if (K > 0) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const T* sl = Fp + 3 * ((size_t)jj * K + (q < K ? q : 0));
        L.f[p][q][0] = sl[0]; L.f[p][q][1] = sl[1]; L.f[p][q][2] = sl[2];
      }
    }
    if (R) {
      const double* rb = r + 3 * (size_t)(col0 + jj);
      L.rb[p][0] = rb[0]; L.rb[p][1] = rb[1]; L.rb[p][2] = rb[2];
    }
  }
}

template <int P, bool R, class T>
__device__ __forceinline__ void front_finish(const FrontLoads<P, T>& L, int j0, int n, const T* __restrict__ Fp,
                                             int K, T* front) {
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int j = j0 + (int)threadIdx.x + p * (int)blockDim.x;
    if (j < n) {
      T ax = 0, ay = 0, az = 0;
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (q < K) { ax += L.f[p][q][0]; ay += L.f[p][q][1]; az += L.f[p][q][2]; }
      for (int q = 4; q < K; ++q) {
        const T* sl = Fp + 3 * ((size_t)j * K + q);
        ax += sl[0]; ay += sl[1]; az += sl[2];
      }
      if (R) {
        front[3 * j + 0] = (T)L.rb[p][0] - ax;
        front[3 * j + 1] = (T)L.rb[p][1] - ay;
        front[3 * j + 2] = (T)L.rb[p][2] - az;
      } else {
        front[3 * j + 0] = ax; front[3 * j + 1] = ay; front[3 * j + 2] = az;
      }
    }
  }
}

// Positions gathered up front per thread: two in FP32; one in FP64, whose
// loads take twice the registers.
template <class T>
constexpr int front_pos() { return sizeof(T) == 4 ? 2 : 1; }

// Front entries [0, n) of a task: r (or, in the first launch of a sweep
// that starts from the RHS slots, the slot sums, also stored to r when
// store_r) minus the K child slots of each position.
template <bool S, class T>
__device__ __forceinline__ void assemble_front(const BandArgs<T>& a, int col0, int n, const T* __restrict__ Fp, int K,
                                               bool store_r, T* front) {
  const int tid = threadIdx.x;
  if constexpr (S) {
    // L lanes per column sum the column's RHS slots
    int L = 1;
    while (L < 8 && (int)blockDim.x / (2 * L) >= n) L *= 2;
    with_lanes<1>(L, [&](auto LC) {
      constexpr int Lc = decltype(LC)::value;
      const int ln = tid & (Lc - 1), gp = tid / Lc, ng = blockDim.x / Lc;
      for (int j = gp; j < n; j += ng) {
        const double3 rb = rhs_group<Lc>(a.src, col0 + j, ln);
        if (ln == 0) {
          if (store_r) { double* o = a.r + 3 * (size_t)(col0 + j); o[0] = rb.x; o[1] = rb.y; o[2] = rb.z; }
          T ax, ay, az;
          front_sum(Fp + 3 * (size_t)j * K, K, ax, ay, az);
          front[3 * j + 0] = (T)rb.x - ax;
          front[3 * j + 1] = (T)rb.y - ay;
          front[3 * j + 2] = (T)rb.z - az;
        }
      }
    });
  } else {
#pragma unroll 4
    for (int j = tid; j < n; j += blockDim.x) {
      T ax, ay, az;
      front_sum(Fp + 3 * (size_t)j * K, K, ax, ay, az);
      const double* rb = a.r + 3 * (size_t)(col0 + j);
      front[3 * j + 0] = (T)rb[0] - ax;
      front[3 * j + 1] = (T)rb[1] - ay;
      front[3 * j + 2] = (T)rb[2] - az;
    }
  }
}

template <int G, bool S, bool FB, class T>
__device__ void fwd_task(const TaskDev td, const BandArgs<T>& a, const T* __restrict__ M, T* __restrict__ y,
                         T* __restrict__ F, T* front) {
  constexpr int kPrefRows = pref_rows<G>();
  constexpr int kPrefLoads = pref_loads<G>();
  const int col0 = td.q0.x, w = td.q0.y, moff = td.q0.w;
  const int b = td.q1.x, e = td.q1.y, uoff = td.q1.w;
  const int K = td.q2.y;
  const T* Fp = F + 3 * (size_t)td.q2.x;
  const int tid = threadIdx.x;
  const int tlen = e <= w ? e : w;
  const int lane = tid & (G - 1), grp = tid / G, ngrp = blockDim.x / G;
  const T* Mp = M + moff;
  const long long tri = (long long)w * (w + 1) / 2;
  // The front's loads first (front_issue), then the block rows and output
  // slots: their addresses depend only on the task, so all of them are in
  // flight together, and the front's sums wait only for the front.
  // FB = false (the kernels compiled for more waves per SIMD): the front is
  // gathered after the rows, in fewer registers.
  constexpr int kFP = front_pos<T>();
  const int pt0 = b > w ? b : w;  // pass-through positions [pt0, e)
  FrontLoads<kFP, T> fl;
  FrontLoads<1, T> pl;
  if constexpr (!S && FB) front_issue<kFP, true>(fl, a.r, col0, 0, tlen, Fp, K);
  if constexpr (FB) front_issue<1, false>(pl, a.r, col0, pt0, e, Fp, K);
  T mv[kPrefRows][kPrefLoads];
  int us[kPrefRows];
#pragma unroll
  for (int p = 0; p < kPrefRows; ++p) {
    const int row = b + grp + p * ngrp;
    const int len = row < e ? (row < w ? row + 1 : w) : 0;
    load_row<G, kPrefLoads>(mv[p], Mp + fwd_row_off(row < e ? row : b, w, tri), len, lane);
    us[p] = row < e && row >= w ? a.uslot[uoff + row - w] : 0;
  }
  if constexpr (S || !FB) {
    assemble_front<S>(a, col0, tlen, Fp, K, e >= w, front);
  } else {
    front_finish<kFP, true>(fl, 0, tlen, Fp, K, front);
    for (int j = kFP * (int)blockDim.x + tid; j < tlen; j += blockDim.x) {
      T ax, ay, az;
      front_sum(Fp + 3 * (size_t)j * K, K, ax, ay, az);
      const double* rb = a.r + 3 * (size_t)(col0 + j);
      front[3 * j + 0] = (T)rb[0] - ax;
      front[3 * j + 1] = (T)rb[1] - ay;
      front[3 * j + 2] = (T)rb[2] - az;
    }
  }
  if constexpr (FB) front_finish<1, false>(pl, pt0, e, Fp, K, front);
  for (int j = pt0 + (FB ? (int)blockDim.x : 0) + tid; j < e; j += blockDim.x) {
    T ax, ay, az;
    front_sum(Fp + 3 * (size_t)j * K, K, ax, ay, az);
    front[3 * j + 0] = ax; front[3 * j + 1] = ay; front[3 * j + 2] = az;
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < kPrefRows; ++p) {
    const int row = b + grp + p * ngrp;
    if (row < e) {  // uniform within the group
      const int len = row < w ? row + 1 : w;
      T sx, sy, sz;
      dot_row<G, kPrefLoads>(mv[p], Mp + fwd_row_off(row, w, tri), len, front, lane, sx, sy, sz);
      fwd_store(row, w, col0, us[p], sx, sy, sz, front, y, F, lane);
    }
  }
  for (int row = b + grp + kPrefRows * ngrp; row < e; row += ngrp) {
    const int len = row < w ? row + 1 : w;
    const int u = row >= w ? a.uslot[uoff + row - w] : 0;
    T m1[kPrefLoads];
    load_row<G, kPrefLoads>(m1, Mp + fwd_row_off(row, w, tri), len, lane);
    T sx, sy, sz;
    dot_row<G, kPrefLoads>(m1, Mp + fwd_row_off(row, w, tri), len, front, lane, sx, sy, sz);
    fwd_store(row, w, col0, u, sx, sy, sz, front, y, F, lane);
  }
}

// Commit of column col with this sweep's d (plus earlier sweeps' acc).
__device__ __forceinline__ void commit_col(const CommitArgs& c, int col, double dx, double dy, double dz, double& df,
                                           double& nm) {
  if (c.add_acc) {
    dx += c.acc[3 * col + 0]; dy += c.acc[3 * col + 1]; dz += c.acc[3 * col + 2];
  }
  commit_vertex(c, col, dx, dy, dz, df, nm);
}

// Output metadata of one solved column, loaded together with the block rows
// so that the stores after the dot product wait on no further load: the
// range of its pushes and, when committing (C), component `lane` (< 3) of
// x, prev and acc.
// pt: the lane's first push target dpush_idx[q0 + lane] when push_pre
// loaded it, else -1.
struct ColPre {
  int q0, q1, pin, pt;
  double xo, po, ao;
};

template <bool C, class T>
__device__ __forceinline__ void col_pre(const BandArgs<T>& a, const CommitArgs& c, int col, int lane, ColPre& p) {
  p.q0 = a.dpush_ptr[col];
  p.q1 = a.dpush_ptr[col + 1];
  p.pt = -1;
  p.pin = 0; p.xo = 0; p.po = 0; p.ao = 0;
  if (C && lane < 3) {
    p.xo = c.x[3 * (size_t)col + lane];
    p.po = c.prev[3 * (size_t)col + lane];
    if (c.add_acc) p.ao = c.acc[3 * (size_t)col + lane];
    p.pin = c.pinned[col];
  }
}

// Stores of one solved column (every lane of the group holds sx, sy, sz):
// d, the negated pushes to the descendants' v slots spread over the g lanes,
// and, when committing, x += d (+ acc) with the pinned / ground rules of
// commit_vertex, component k by lane k.
template <bool C, class T>
__device__ __forceinline__ void col_store(const BandArgs<T>& a, const CommitArgs& c, int col, T sx, T sy, T sz,
                                          const ColPre& p, int lane, int g, T* __restrict__ V, T* __restrict__ d,
                                          double& df, double& nm) {
  const T s = lane == 0 ? sx : (lane == 1 ? sy : sz);
  if (lane < 3) d[3 * (size_t)col + lane] = s;
  for (int q = p.q0 + lane; q < p.q1; q += g) {
    T* o = V + 3 * (size_t)(q == p.q0 + lane && p.pt >= 0 ? p.pt : a.dpush_idx[q]);
    o[0] = -sx; o[1] = -sy; o[2] = -sz;
  }
  if (C && lane < 3) {
    const double dv = (double)s + p.ao;
    df += dv * dv;
    nm += p.xo * p.xo;
    double v = p.xo + dv;
    if (p.pin) v = p.po;
    else if (lane == 1 && c.has_ground && v < c.ground) v = c.ground;
    c.x[3 * (size_t)col + lane] = v;
  }
}

// Backward task (L^T d = y): columns [b, e) of P, read as rows of the
// transposed copy Mt (row j = column j of M_P, m - j entries). With
// v = [y_P ; V_P] where V_P[q] = -d[I_P[q]] was pushed by the owner of that
// column,
//   d_P[j] = sum_{k >= j} Linv[k][j] y_P[k] - sum_q W[q][j] d[I_P[q]]
//          = sum_i Mt[j][i] v[j + i].
// Each d_P[j] is stored, pushed (negated) to the v slots of the descendants
// that read it, and, when committing, added to x.
template <int G, bool C, bool FB, class T>
__device__ void bwd_task(const TaskDev td, const BandArgs<T>& a, const CommitArgs& c, const T* __restrict__ Mt,
                         const T* __restrict__ y, T* __restrict__ V, T* __restrict__ d, T* v, double& df, double& nm) {
  constexpr int kPrefRows = pref_rows<G>();
  constexpr int kPrefLoads = pref_loads<G>();
  const int col0 = td.q0.x, w = td.q0.y, r = td.q0.z, moff = td.q0.w;
  const int b = td.q1.x, e = td.q1.y, uoff = td.q1.w;
  const int m = w + r;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int lane = tid & (G - 1), grp = tid / G, ngrp = nt / G;
  const T* Tp = Mt + moff;
  auto row_ptr = [&](int j) { return Tp + (long long)j * m - (long long)j * (j - 1) / 2; };
  T mv[kPrefRows][kPrefLoads];
  ColPre cp[kPrefRows];
  if constexpr (FB) {
    // As in fwd_task: the small L2 loads (push ranges, commit values, v)
    // are issued before the block rows, so their uses wait only for them;
    // each lane's first push target is loaded before the barrier too.
    constexpr int PV = front_pos<T>();
#pragma unroll
    for (int p = 0; p < kPrefRows; ++p) {
      const int j = b + grp + p * ngrp;
      col_pre<C>(a, c, col0 + (j < e ? j : b), lane, cp[p]);
    }
    T vv[PV][3];
#pragma unroll
    for (int p = 0; p < PV; ++p) {
      const int k = b + tid + p * nt;
      const int kk = k < m ? k : b;
      const T* sv = kk < w ? y + 3 * (size_t)(col0 + kk) : V + 3 * (size_t)(uoff + kk - w);
      vv[p][0] = sv[0]; vv[p][1] = sv[1]; vv[p][2] = sv[2];
    }
#pragma unroll
    for (int p = 0; p < kPrefRows; ++p) {
      const int j = b + grp + p * ngrp;
      load_row<G, kPrefLoads>(mv[p], row_ptr(j < e ? j : b), j < e ? m - j : 0, lane);
    }
#pragma unroll
    for (int p = 0; p < kPrefRows; ++p) {
      const int q = cp[p].q0 + lane;
      if (q < cp[p].q1) cp[p].pt = a.dpush_idx[q];
    }
#pragma unroll
    for (int p = 0; p < PV; ++p) {
      const int k = b + tid + p * nt;
      if (k < m) { v[3 * k + 0] = vv[p][0]; v[3 * k + 1] = vv[p][1]; v[3 * k + 2] = vv[p][2]; }
    }
    for (int k = b + PV * nt + tid; k < m; k += nt) {
      const T* sv = k < w ? y + 3 * (size_t)(col0 + k) : V + 3 * (size_t)(uoff + k - w);
      v[3 * k + 0] = sv[0]; v[3 * k + 1] = sv[1]; v[3 * k + 2] = sv[2];
    }
  } else {
#pragma unroll
    for (int p = 0; p < kPrefRows; ++p) {
      const int j = b + grp + p * ngrp;
      const int jj = j < e ? j : b;
      load_row<G, kPrefLoads>(mv[p], row_ptr(jj), j < e ? m - j : 0, lane);
      col_pre<C>(a, c, col0 + jj, lane, cp[p]);
    }
#pragma unroll 4
    for (int k = b + tid; k < m; k += nt) {
      const T* s = k < w ? y + 3 * (size_t)(col0 + k) : V + 3 * (size_t)(uoff + k - w);
      v[3 * k + 0] = s[0]; v[3 * k + 1] = s[1]; v[3 * k + 2] = s[2];
    }
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < kPrefRows; ++p) {
    const int j = b + grp + p * ngrp;
    if (j < e) {  // uniform within the group
      T sx, sy, sz;
      dot_row<G, kPrefLoads>(mv[p], row_ptr(j), m - j, v + 3 * j, lane, sx, sy, sz);
      col_store<C>(a, c, col0 + j, sx, sy, sz, cp[p], lane, G, V, d, df, nm);
    }
  }
  for (int j = b + grp + kPrefRows * ngrp; j < e; j += ngrp) {
    ColPre q;
    col_pre<C>(a, c, col0 + j, lane, q);
    T m1[kPrefLoads];
    load_row<G, kPrefLoads>(m1, row_ptr(j), m - j, lane);
    T sx, sy, sz;
    dot_row<G, kPrefLoads>(m1, row_ptr(j), m - j, v + 3 * j, lane, sx, sy, sz);
    col_store<C>(a, c, col0 + j, sx, sy, sz, q, lane, G, V, d, df, nm);
  }
}

// One band of the forward sweep per launch. Blocks [0, ntasks) run tasks;
// in the first band of a sweep that starts from the slots, the remaining
// blocks compute (and store) r for the columns outside the deepest band.
template <class T, bool S, bool FB = true>
__device__ __forceinline__ void fwd_band(const BandArgs<T>& a, const T* __restrict__ M, T* __restrict__ y,
                                         T* __restrict__ F, const StepDev* st) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int ntb = task_blocks(a);
  if ((int)blockIdx.x < ntb) {
    const int ti = task_of(a, (int)blockIdx.x);
    if (ti < 0) return;
    const TaskDev td = a.tasks[ti];  // issued together with the flag load
    if (st && st->done) return;
    with_lanes<4>(td.q1.z, [&](auto GC) {
      fwd_task<decltype(GC)::value, S, FB, T>(td, a, M, y, F, reinterpret_cast<T*>(smem_raw));
    });
    return;
  }
  if (!S || (st && st->done)) return;
  // 4 lanes per column: blockDim/4 columns per extra block
  const int k = ((int)blockIdx.x - ntb) * (blockDim.x / 4) + (int)threadIdx.x / 4;
  const int lane = threadIdx.x & 3;
  if (k >= a.nupper) return;  // uniform within each 4-lane group
  const int col = a.upper_cols[k];
  const double3 rb = rhs_group<4>(a.src, col, lane);
  if (lane == 0) {
    double* o = a.r + 3 * (size_t)col;
    o[0] = rb.x; o[1] = rb.y; o[2] = rb.z;
  }
}

template <class T, bool S>
__global__ void __launch_bounds__(256) k_fwd_band(BandArgs<T> a, const T* __restrict__ M, T* __restrict__ y,
                                                  T* __restrict__ F, const StepDev* st) {
  fwd_band<T, S>(a, M, y, F, st);
}

// The FP32 forward band compiled for eight waves per SIMD (64 VGPRs, a few
// bytes of spill) instead of the compiler's own six. The host picks it for
// bands of many tasks, where occupancy pays, and not for small bands, where
// the spills' latency costs more.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8)))
k_fwd_band_w8(BandArgs<float> a, const float* __restrict__ M, float* __restrict__ y, float* __restrict__ F,
              const StepDev* st) {
  fwd_band<float, false, false>(a, M, y, F, st);
}

// The FP64 forward band for the same wide bands, at five waves per SIMD
// instead of four.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5)))
k_fwd_band_d5(BandArgs<double> a, const double* __restrict__ M, double* __restrict__ y, double* __restrict__ F,
              const StepDev* st) {
  fwd_band<double, false, false>(a, M, y, F, st);
}

// Forward band with each task's block rows staged in LDS (k_fwd_band_st).
// A task's entries are one contiguous run of M_P (rows [b, e) of the
// row-major trapezoid): every thread loads E of them, flat and coalesced,
// in the same round as the front (front_issue), writes them to LDS, and
// after the barrier the lane groups take the rows' dot products from LDS
// alone, however long and however many the rows. A task holds at most
// 256 E entries and 256 rows (the host cuts them so); lane l of a group
// sums entries l, l + G, ... and the group reduces as in fwd_task, so the
// sums are the same.
template <int E, class T>
__device__ void fwd_task_st(const TaskDev td, const BandArgs<T>& a, const T* __restrict__ M, T* __restrict__ y,
                            T* __restrict__ F, T* front, T* mat, int* usl) {
  const int col0 = td.q0.x, w = td.q0.y, moff = td.q0.w;
  const int b = td.q1.x, e = td.q1.y, uoff = td.q1.w;
  const int K = td.q2.y;
  const T* Fp = F + 3 * (size_t)td.q2.x;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int tlen = e <= w ? e : w;
  const long long tri = (long long)w * (w + 1) / 2;
  const long long o0 = fwd_row_off(b, w, tri);
  const int ne = (int)(fwd_row_off(e, w, tri) - o0);
  constexpr int kFP = front_pos<T>();
  const int pt0 = b > w ? b : w;  // pass-through positions [pt0, e)
  FrontLoads<kFP, T> fl;
  FrontLoads<1, T> pl;
  front_issue<kFP, true>(fl, a.r, col0, 0, tlen, Fp, K);
  front_issue<1, false>(pl, a.r, col0, pt0, e, Fp, K);
  T mv[E];
  const T* __restrict__ src = M + moff + o0;
  if (ne > 0) {
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const int k = tid + i * nt;
      mv[i] = __builtin_nontemporal_load(src + (k < ne ? k : 0));
    }
  }
  const int rt = b + tid;
  const int u = rt < e && rt >= w ? a.uslot[uoff + rt - w] : 0;
  front_finish<kFP, true>(fl, 0, tlen, Fp, K, front);
  for (int j = kFP * nt + tid; j < tlen; j += nt) {
    T ax, ay, az;
    front_sum(Fp + 3 * (size_t)j * K, K, ax, ay, az);
    const double* rb = a.r + 3 * (size_t)(col0 + j);
    front[3 * j + 0] = (T)rb[0] - ax;
    front[3 * j + 1] = (T)rb[1] - ay;
    front[3 * j + 2] = (T)rb[2] - az;
  }
  front_finish<1, false>(pl, pt0, e, Fp, K, front);
  for (int j = pt0 + nt + tid; j < e; j += nt) {
    T ax, ay, az;
    front_sum(Fp + 3 * (size_t)j * K, K, ax, ay, az);
    front[3 * j + 0] = ax; front[3 * j + 1] = ay; front[3 * j + 2] = az;
  }
  if (ne > 0) {
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const int k = tid + i * nt;
      if (k < ne) mat[k] = mv[i];
    }
  }
  if (rt < e) usl[tid] = u;
  __syncthreads();
  with_lanes<4>(td.q1.z, [&](auto GC) {
    constexpr int G = decltype(GC)::value;
    const int lane = tid & (G - 1), ngrp = nt / G;
    for (int row = b + tid / G; row < e; row += ngrp) {  // uniform within the group
      const int len = row < w ? row + 1 : w;
      const T* __restrict__ mr = mat + (fwd_row_off(row, w, tri) - o0);
      T ax = 0, ay = 0, az = 0;
#pragma unroll 4
      for (int k = lane; k < len; k += G) {
        const T m = mr[k];
        ax += m * front[3 * k]; ay += m * front[3 * k + 1]; az += m * front[3 * k + 2];
      }
      sum3<G>(ax, ay, az);
      fwd_store(row, w, col0, usl[row - b], ax, ay, az, front, y, F, lane);
    }
  });
}

// The first forward launch of a sweep that starts from the RHS slots, staged
// (k_fwd_band_st1). Its supernodes are the deepest, with no children (K =
// 0), and their columns' RHS slots are one contiguous run, [q2.z, q2.w) in
// the task record: every thread loads NS slot values of that run, the
// column starts inc_ptr[col0 .. col0 + tlen] and its column's inertia
// inputs, in the same round as E block entries (as fwd_task_st). After a
// barrier, L lanes per column sum the slots from LDS in the order of
// rhs_group (lane l: slots l, l + L, ... of the column, then the group
// reduction), add the inertia term, store r when the task covers rows past
// w, and form the front; after a second barrier the rows' dot products run
// from LDS. Tasks hold at most 255 columns, 256 rows, 256 E entries and
// 256 NS slot values (the host cuts them so).
template <int E, int NS, class T>
__device__ void fwd_task_st1(const TaskDev td, const BandArgs<T>& a, const T* __restrict__ M, T* __restrict__ y,
                             T* __restrict__ F, T* front, T* mat, int* usl, T* sl, int* ip, double* ine) {
  const int col0 = td.q0.x, w = td.q0.y, moff = td.q0.w;
  const int b = td.q1.x, e = td.q1.y, uoff = td.q1.w;
  const int sb = td.q2.z, nsv = 3 * (td.q2.w - td.q2.z);
  const int tid = threadIdx.x, nt = blockDim.x;
  const int tlen = e <= w ? e : w;
  const long long tri = (long long)w * (w + 1) / 2;
  const long long o0 = fwd_row_off(b, w, tri);
  const int ne = (int)(fwd_row_off(e, w, tri) - o0);
  const RhsSrc<T>& rs = a.src;
  T sv[NS];
  const T* __restrict__ ssrc = rs.cslot + 3 * (size_t)sb;
  if (nsv > 0) {
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      const int k = tid + i * nt;
      sv[i] = ssrc[k < nsv ? k : 0];
    }
  }
  const int ipv = tid <= tlen ? rs.inc_ptr[col0 + tid] : 0;
  double3 in = make_double3(0, 0, 0);
  if (tid < tlen) in = inertia(rs, col0 + tid);
  T mv[E];
  const T* __restrict__ src = M + moff + o0;
  if (ne > 0) {
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const int k = tid + i * nt;
      mv[i] = __builtin_nontemporal_load(src + (k < ne ? k : 0));
    }
  }
  const int rt = b + tid;
  const int u = rt < e && rt >= w ? a.uslot[uoff + rt - w] : 0;
  if (nsv > 0) {
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      const int k = tid + i * nt;
      if (k < nsv) sl[k] = sv[i];
    }
  }
  if (tid <= tlen) ip[tid] = ipv;
  if (tid < tlen) { ine[3 * tid] = in.x; ine[3 * tid + 1] = in.y; ine[3 * tid + 2] = in.z; }
  if (ne > 0) {
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const int k = tid + i * nt;
      if (k < ne) mat[k] = mv[i];
    }
  }
  if (rt < e) usl[tid] = u;
  // pass-through positions (rows past w): no children, so zero
  for (int j = (b > w ? b : w) + tid; j < e; j += nt) { front[3 * j] = 0; front[3 * j + 1] = 0; front[3 * j + 2] = 0; }
  __syncthreads();
  const bool store_r = e >= w;
  int L = 1;
  while (L < 8 && nt / (2 * L) >= tlen) L *= 2;
  with_lanes<1>(L, [&](auto LC) {
    constexpr int Lc = decltype(LC)::value;
    const int ln = tid & (Lc - 1), gp = tid / Lc, ng = nt / Lc;
    for (int j = gp; j < tlen; j += ng) {
      T ax = 0, ay = 0, az = 0;
      const int p0 = ip[j] - sb, pe = ip[j + 1] - sb;
      for (int p = p0 + ln; p < pe; p += Lc) { ax += sl[3 * p]; ay += sl[3 * p + 1]; az += sl[3 * p + 2]; }
      sum3<Lc>(ax, ay, az);
      if (ln == 0) {
        const double rx = ine[3 * j] + (double)ax, ry = ine[3 * j + 1] + (double)ay, rz = ine[3 * j + 2] + (double)az;
        if (store_r) { double* o = a.r + 3 * (size_t)(col0 + j); o[0] = rx; o[1] = ry; o[2] = rz; }
        front[3 * j + 0] = (T)rx; front[3 * j + 1] = (T)ry; front[3 * j + 2] = (T)rz;
      }
    }
  });
  __syncthreads();
  with_lanes<4>(td.q1.z, [&](auto GC) {
    constexpr int G = decltype(GC)::value;
    const int lane = tid & (G - 1), ngrp = nt / G;
    for (int row = b + tid / G; row < e; row += ngrp) {  // uniform within the group
      const int len = row < w ? row + 1 : w;
      const T* __restrict__ mr = mat + (fwd_row_off(row, w, tri) - o0);
      T ax = 0, ay = 0, az = 0;
#pragma unroll 4
      for (int k = lane; k < len; k += G) {
        const T m = mr[k];
        ax += m * front[3 * k]; ay += m * front[3 * k + 1]; az += m * front[3 * k + 2];
      }
      sum3<G>(ax, ay, az);
      fwd_store(row, w, col0, usl[row - b], ax, ay, az, front, y, F, lane);
    }
  });
}

// Dynamic LDS of k_fwd_band_st1: the front (3 fmax values), the staged
// entries (256 E), the slot values (256 NS), the output slot numbers (256
// ints), the column starts (256 ints), the inertia terms (3 x 256 doubles).
// Blocks past the task blocks form r for the columns outside the deepest
// band, as k_fwd_band<T, true>.
template <int E, int NS, class T>
__global__ void __launch_bounds__(256) k_fwd_band_st1(BandArgs<T> a, const T* __restrict__ M, T* __restrict__ y,
                                                      T* __restrict__ F, int fmax, const StepDev* st) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int ntb = task_blocks(a);
  if ((int)blockIdx.x < ntb) {
    const int ti = task_of(a, (int)blockIdx.x);
    if (ti < 0) return;
    const TaskDev td = a.tasks[ti];  // issued together with the flag load
    if (st && st->done) return;
    double* ine = reinterpret_cast<double*>(smem_raw);
    T* front = reinterpret_cast<T*>(ine + 3 * 256);
    T* mat = front + 3 * (size_t)fmax;
    T* sl = mat + 256 * E;
    int* usl = reinterpret_cast<int*>(sl + 256 * NS);
    int* ip = usl + 256;
    fwd_task_st1<E, NS, T>(td, a, M, y, F, front, mat, usl, sl, ip, ine);
    return;
  }
  if (st && st->done) return;
  const int k = ((int)blockIdx.x - ntb) * (blockDim.x / 4) + (int)threadIdx.x / 4;
  const int lane = threadIdx.x & 3;
  if (k >= a.nupper) return;  // uniform within each 4-lane group
  const int col = a.upper_cols[k];
  const double3 rb = rhs_group<4>(a.src, col, lane);
  if (lane == 0) {
    double* o = a.r + 3 * (size_t)col;
    o[0] = rb.x; o[1] = rb.y; o[2] = rb.z;
  }
}

// Dynamic LDS of k_fwd_band_st: the front (3 fmax values), then the staged
// entries (256 E), then the task's output slot numbers (256 ints).
template <int E, class T>
__global__ void __launch_bounds__(256) k_fwd_band_st(BandArgs<T> a, const T* __restrict__ M, T* __restrict__ y,
                                                     T* __restrict__ F, int fmax, const StepDev* st) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int ti = task_of(a, (int)blockIdx.x);
  if (ti < 0) return;
  const TaskDev td = a.tasks[ti];  // issued together with the flag load
  if (st && st->done) return;
  T* front = reinterpret_cast<T*>(smem_raw);
  T* mat = front + 3 * (size_t)fmax;
  int* usl = reinterpret_cast<int*>(mat + 256 * E);
  fwd_task_st<E, T>(td, a, M, y, F, front, mat, usl);
}

// One band of the backward sweep per launch. In the last band of an FP64
// solve the tasks also commit their columns and the extra blocks commit the
// columns outside the deepest band (solved by earlier launches).
template <class T, bool C, bool FB = true>
__device__ __forceinline__ void bwd_band(const BandArgs<T>& a, const CommitArgs& c, const T* __restrict__ Mt,
                                         const T* __restrict__ y, T* __restrict__ V, T* __restrict__ d,
                                         const StepDev* st) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double df = 0, nm = 0;
  const int ntb = task_blocks(a);
  if ((int)blockIdx.x < ntb) {
    const int ti = task_of(a, (int)blockIdx.x);
    if (ti >= 0) {  // an idle padding block still writes its (zero) partials
      const TaskDev td = a.tasks[ti];  // issued together with the flag load
      if (st && st->done) return;
      with_lanes<4>(td.q1.z, [&](auto GC) {
        bwd_task<decltype(GC)::value, C, FB, T>(td, a, c, Mt, y, V, d, reinterpret_cast<T*>(smem_raw), df, nm);
      });
    } else if (st && st->done) {
      return;
    }
  } else {
    if (st && st->done) return;
    const int k = ((int)blockIdx.x - ntb) * blockDim.x + threadIdx.x;
    if (C && k < a.nupper) {
      const int col = a.upper_cols[k];
      const T* s = d + 3 * (size_t)col;
      commit_col(c, col, (double)s[0], (double)s[1], (double)s[2], df, nm);
    }
  }
  if (C && c.want_resid) {
    __syncthreads();
    block_partials(df, nm, c.partials);
  }
}

template <class T, bool C>
__global__ void __launch_bounds__(256) k_bwd_band(BandArgs<T> a, CommitArgs c, const T* __restrict__ Mt,
                                                  const T* __restrict__ y, T* __restrict__ V, T* __restrict__ d,
                                                  const StepDev* st) {
  bwd_band<T, C>(a, c, Mt, y, V, d, st);
}

// The last backward band, committing, with each task's rows of Mt staged in
// LDS (k_bwd_band_stc), as k_fwd_band_st: the task's rows [b, e) of Mt
// are one contiguous run, loaded E per thread after v = [y_P ; V_P]
// (positions b .. m) and each row's push range and commit values (one row
// per thread, at most 256 rows); the dot products, pushes and commits then
// take their operands from LDS. Lane l of a group sums entries l, l + G,
// ... of its row as bwd_task does, so d and x are the same; only the
// residual partials are summed over another thread order.
template <int E, class T>
__device__ void bwd_task_stc(const TaskDev td, const BandArgs<T>& a, const CommitArgs& c, const T* __restrict__ Mt,
                             const T* __restrict__ y, T* __restrict__ V, T* __restrict__ d, T* v, T* mat, int* meta,
                             double* cmeta, double& df, double& nm) {
  const int col0 = td.q0.x, w = td.q0.y, r = td.q0.z, moff = td.q0.w;
  const int b = td.q1.x, e = td.q1.y, uoff = td.q1.w;
  const int m = w + r;
  const int tid = threadIdx.x, nt = blockDim.x;
  auto roff = [&](int j) { return (long long)j * m - (long long)j * (j - 1) / 2; };
  const long long o0 = roff(b);
  const int ne = (int)(roff(e) - o0);
  // row metadata of row b + tid: push range, pinned flag, x, prev, acc
  const int jr = b + tid;
  const bool hr = jr < e;
  const int cj = col0 + (hr ? jr : b);
  const int q0 = a.dpush_ptr[cj], q1 = a.dpush_ptr[cj + 1];
  const int pin = c.pinned[cj];
  double xo[3], po[3], ao[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    xo[k] = c.x[3 * (size_t)cj + k];
    po[k] = c.prev[3 * (size_t)cj + k];
    ao[k] = c.add_acc ? c.acc[3 * (size_t)cj + k] : 0.0;
  }
  T vv[2][3];
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    const int k = b + tid + p * nt;
    const int kk = k < m ? k : b;
    const T* sv = kk < w ? y + 3 * (size_t)(col0 + kk) : V + 3 * (size_t)(uoff + kk - w);
    vv[p][0] = sv[0]; vv[p][1] = sv[1]; vv[p][2] = sv[2];
  }
  T mv[E];
  const T* __restrict__ src = Mt + moff + o0;
  if (ne > 0) {
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const int k = tid + i * nt;
      mv[i] = __builtin_nontemporal_load(src + (k < ne ? k : 0));
    }
  }
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    const int k = b + tid + p * nt;
    if (k < m) { v[3 * k + 0] = vv[p][0]; v[3 * k + 1] = vv[p][1]; v[3 * k + 2] = vv[p][2]; }
  }
  for (int k = b + 2 * nt + tid; k < m; k += nt) {
    const T* sv = k < w ? y + 3 * (size_t)(col0 + k) : V + 3 * (size_t)(uoff + k - w);
    v[3 * k + 0] = sv[0]; v[3 * k + 1] = sv[1]; v[3 * k + 2] = sv[2];
  }
  if (ne > 0) {
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const int k = tid + i * nt;
      if (k < ne) mat[k] = mv[i];
    }
  }
  if (hr) {
    meta[3 * tid] = q0; meta[3 * tid + 1] = q1; meta[3 * tid + 2] = pin;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      cmeta[9 * tid + k] = xo[k]; cmeta[9 * tid + 3 + k] = po[k]; cmeta[9 * tid + 6 + k] = ao[k];
    }
  }
  __syncthreads();
  with_lanes<4>(td.q1.z, [&](auto GC) {
    constexpr int G = decltype(GC)::value;
    const int lane = tid & (G - 1), ngrp = nt / G;
    for (int j = b + tid / G; j < e; j += ngrp) {  // uniform within the group
      const int len = m - j;
      const T* __restrict__ mr = mat + (roff(j) - o0);
      const T* __restrict__ vj = v + 3 * j;
      T ax = 0, ay = 0, az = 0;
#pragma unroll 4
      for (int k = lane; k < len; k += G) {
        const T mk = mr[k];
        ax += mk * vj[3 * k]; ay += mk * vj[3 * k + 1]; az += mk * vj[3 * k + 2];
      }
      sum3<G>(ax, ay, az);
      const int t = j - b;
      ColPre cp;
      cp.q0 = meta[3 * t]; cp.q1 = meta[3 * t + 1]; cp.pin = meta[3 * t + 2]; cp.pt = -1;
      const int cl = lane < 3 ? lane : 0;
      cp.xo = cmeta[9 * t + cl]; cp.po = cmeta[9 * t + 3 + cl]; cp.ao = cmeta[9 * t + 6 + cl];
      col_store<true>(a, c, col0 + j, ax, ay, az, cp, lane, G, V, d, df, nm);
    }
  });
}

// Dynamic LDS of k_bwd_band_stc: the 256 rows' commit values (9 doubles
// each), v (3 fmax values), the staged entries (256 E), the rows' push
// ranges and pinned flags (3 x 256 ints). Blocks past the task blocks
// commit the columns outside the deepest band, and every block writes its
// residual partials when asked, as k_bwd_band<T, true>.
template <int E, class T>
__global__ void __launch_bounds__(256) k_bwd_band_stc(BandArgs<T> a, CommitArgs c, const T* __restrict__ Mt,
                                                      const T* __restrict__ y, T* __restrict__ V, T* __restrict__ d,
                                                      int fmax, const StepDev* st) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double df = 0, nm = 0;
  const int ntb = task_blocks(a);
  if ((int)blockIdx.x < ntb) {
    const int ti = task_of(a, (int)blockIdx.x);
    if (ti >= 0) {  // an idle padding block still writes its (zero) partials
      const TaskDev td = a.tasks[ti];  // issued together with the flag load
      if (st && st->done) return;
      double* cmeta = reinterpret_cast<double*>(smem_raw);
      T* v = reinterpret_cast<T*>(cmeta + 9 * 256);
      T* mat = v + 3 * (size_t)fmax;
      int* meta = reinterpret_cast<int*>(mat + 256 * E);
      bwd_task_stc<E, T>(td, a, c, Mt, y, V, d, v, mat, meta, cmeta, df, nm);
    } else if (st && st->done) {
      return;
    }
  } else {
    if (st && st->done) return;
    const int k = ((int)blockIdx.x - ntb) * blockDim.x + threadIdx.x;
    if (k < a.nupper) {
      const int col = a.upper_cols[k];
      const T* s = d + 3 * (size_t)col;
      commit_col(c, col, (double)s[0], (double)s[1], (double)s[2], df, nm);
    }
  }
  if (c.want_resid) {
    __syncthreads();
    block_partials(df, nm, c.partials);
  }
}

// The FP64 backward band (not the committing one) at five waves per SIMD,
// for bands of many tasks, as k_fwd_band_d5.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5)))
k_bwd_band_d5(BandArgs<double> a, CommitArgs c, const double* __restrict__ Mt, const double* __restrict__ y,
              double* __restrict__ V, double* __restrict__ d, const StepDev* st) {
  bwd_band<double, false, false>(a, c, Mt, y, V, d, st);
}

// Root task: a root supernode R has no off-diagonal rows, so its forward
// and backward sweeps collapse into d_R = (L_RR L_RR^T)^{-1} t_R with its
// dense symmetric inverse (row-major w x w at Sroot + td.q2.z, one per
// dissection tree); rows [b, e) of it. Same front assembly as a forward
// task, same stores as a backward task.
template <int G, bool S, bool C, class T>
__device__ void root_task(const TaskDev td, const BandArgs<T>& a, const CommitArgs& c, const T* __restrict__ Sroot_all,
                          T* __restrict__ F, T* __restrict__ V, T* __restrict__ d, T* front, double& df, double& nm) {
  constexpr int kPrefRows = pref_rows<G>();
  constexpr int kPrefLoads = pref_loads<G>();
  const T* __restrict__ Sroot = Sroot_all + (size_t)td.q2.z;
  const int col0 = td.q0.x, w = td.q0.y;
  const int b = td.q1.x, e = td.q1.y;
  const int K = td.q2.y;
  const T* Fp = F + 3 * (size_t)td.q2.x;
  const int tid = threadIdx.x;
  const int lane = tid & (G - 1), grp = tid / G, ngrp = blockDim.x / G;
  T mv[kPrefRows][kPrefLoads];
  ColPre cp[kPrefRows];
#pragma unroll
  for (int p = 0; p < kPrefRows; ++p) {
    const int row = b + grp + p * ngrp;
    const int rr = row < e ? row : b;
    load_row<G, kPrefLoads>(mv[p], Sroot + (size_t)rr * w, row < e ? w : 0, lane);
    col_pre<C>(a, c, col0 + rr, lane, cp[p]);
  }
  assemble_front<S>(a, col0, w, Fp, K, true, front);
  __syncthreads();
#pragma unroll
  for (int p = 0; p < kPrefRows; ++p) {
    const int row = b + grp + p * ngrp;
    if (row < e) {  // uniform within the group
      T sx, sy, sz;
      dot_row<G, kPrefLoads>(mv[p], Sroot + (size_t)row * w, w, front, lane, sx, sy, sz);
      col_store<C>(a, c, col0 + row, sx, sy, sz, cp[p], lane, G, V, d, df, nm);
    }
  }
  for (int row = b + grp + kPrefRows * ngrp; row < e; row += ngrp) {
    ColPre q;
    col_pre<C>(a, c, col0 + row, lane, q);
    T m1[kPrefLoads];
    load_row<G, kPrefLoads>(m1, Sroot + (size_t)row * w, w, lane);
    T sx, sy, sz;
    dot_row<G, kPrefLoads>(m1, Sroot + (size_t)row * w, w, front, lane, sx, sy, sz);
    col_store<C>(a, c, col0 + row, sx, sy, sz, q, lane, G, V, d, df, nm);
  }
}

template <class T, bool S, bool C>
__global__ void __launch_bounds__(256) k_root_band(BandArgs<T> a, CommitArgs c, const T* __restrict__ Sroot,
                                                   T* __restrict__ F, T* __restrict__ V, T* __restrict__ d,
                                                   const StepDev* st) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double df = 0, nm = 0;
  const int ti = (int)blockIdx.x < task_blocks(a) ? task_of(a, (int)blockIdx.x) : -1;
  if (ti >= 0) {  // an idle padding block still writes its (zero) partials
    const TaskDev td = a.tasks[ti];  // issued together with the flag load
    if (st && st->done) return;
    with_lanes<4>(td.q1.z, [&](auto GC) {
      root_task<decltype(GC)::value, S, C, T>(td, a, c, Sroot, F, V, d, reinterpret_cast<T*>(smem_raw), df, nm);
    });
  } else if (st && st->done) {
    return;
  }
  if (C && c.want_resid) {
    __syncthreads();
    block_partials(df, nm, c.partials);
  }
}

// The root band below at least one other band (no RHS slots, no commit),
// laid out for bandwidth: one row of the dense inverse per wave, its first
// 64 NL entries loaded in one round before the front is assembled, and
// blockDim / 64 rows per workgroup (the host gives each task at most that
// many rows), so each workgroup gathers the front once for that many rows.
// Task b covers rows [q1.x, q1.y) of the root in its record (q0.x col0,
// q0.y w, q2.x first front slot, q2.y slots per position, q2.z offset of
// the inverse in Sroot). The front and the products are taken eight
// entries at a time (scheduling barriers between the groups), which keeps
// the LDS reads of a group, not of the whole row, in registers. Prefetching
// a third position per thread (a 2295-wide front in one round of loads,
// with two or three slots per position instead of four) measured slower at
// 256x256: 12.4 -> 13.9 and 14.7 us per launch.
template <int NL, class T>
__global__ void __launch_bounds__(1024) k_root_dense(const TaskDev* __restrict__ tasks, BandArgs<T> a,
                                                     const T* __restrict__ Sroot_all, const T* __restrict__ F,
                                                     T* __restrict__ V, T* __restrict__ d, const StepDev* st) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* front = reinterpret_cast<T*>(smem_raw);
  const TaskDev td = tasks[blockIdx.x];
  if (st && st->done) return;
  const int col0 = td.q0.x, w = td.q0.y, b = td.q1.x, e = td.q1.y, K = td.q2.y;
  const T* __restrict__ S = Sroot_all + (size_t)td.q2.z;
  const T* __restrict__ Fp = F + 3 * (size_t)td.q2.x;
  const int tid = threadIdx.x, lane = tid & 63, row = b + (tid >> 6);
  const CommitArgs nc{};
  const bool live = row < e;
  const T* __restrict__ Srow = S + (size_t)(live ? row : b) * w;
  // front[j] = r[col0 + j] - (the K child slots of position j): its loads
  // first, then the row's (see front_issue)
  FrontLoads<2, T> fl;
  front_issue<2, true>(fl, a.r, col0, 0, w, Fp, K);
  T mv[NL];
  load_row<64, NL>(mv, Srow, live ? w : 0, lane);
  ColPre cp;
  col_pre<false>(a, nc, col0 + (live ? row : b), lane, cp);
  front_finish<2, true>(fl, 0, w, Fp, K, front);
  for (int j = 2 * (int)blockDim.x + tid; j < w; j += blockDim.x) {
    T ax, ay, az;
    front_sum(Fp + 3 * (size_t)j * K, K, ax, ay, az);
    const double* rb = a.r + 3 * (size_t)(col0 + j);
    front[3 * j + 0] = (T)rb[0] - ax;
    front[3 * j + 1] = (T)rb[1] - ay;
    front[3 * j + 2] = (T)rb[2] - az;
  }
  __syncthreads();
  if (!live) return;
  T ax = 0, ay = 0, az = 0;
#pragma unroll
  for (int c = 0; c < NL; c += 8) {
#pragma unroll
    for (int i = c; i < c + 8; ++i) {
      const int k = lane + i * 64;
      const T* v = front + 3 * (k < w ? k : w - 1);
      ax += mv[i] * v[0]; ay += mv[i] * v[1]; az += mv[i] * v[2];
    }
    __builtin_amdgcn_sched_barrier(0);
  }
  for (int base = NL * 64; base < w; base += 8 * 64) {
    T m2[8];
    load_row<64, 8>(m2, Srow + base, w - base, lane);
    fma_row<64, 8>(m2, front + 3 * base, w - base, lane, ax, ay, az);
  }
  sum3<64>(ax, ay, az);
  double df = 0, nm = 0;
  col_store<false>(a, nc, col0 + row, ax, ay, az, cp, lane, 64, V, d, df, nm);
}

// Iterative refinement between FP32 sweeps (FP64 residual): acc += d and,
// when update_r, r -= A d (A in FP64 CSR, permuted order; rows only read d,
// so the in-place update of r is race-free). The last sweep's d is
// committed together with acc by the last backward band.
__global__ void __launch_bounds__(256) k_ir_update(int n, const int* __restrict__ Ap, const int* __restrict__ Ai,
                                                   const double* __restrict__ Ax, const float* __restrict__ d,
                                                   double* __restrict__ r, double* __restrict__ acc, int first,
                                                   int update_r, const StepDev* st) {
  if (st && st->done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int k = 0; k < 3; ++k) {
    const double dv = (double)d[3 * i + k];
    acc[3 * i + k] = first ? dv : acc[3 * i + k] + dv;
  }
  if (update_r) {
    double sx = 0, sy = 0, sz = 0;
    for (int p = Ap[i]; p < Ap[i + 1]; ++p) {
      const double av = Ax[p];
      const float* dv = d + 3 * (size_t)Ai[p];
      sx += av * (double)dv[0]; sy += av * (double)dv[1]; sz += av * (double)dv[2];
    }
    r[3 * i + 0] -= sx; r[3 * i + 1] -= sy; r[3 * i + 2] -= sz;
  }
}

// Final residual of a PD iteration (pd_solver.rs:402-421, 440-444): sums the
// per-block (|delta|^2, |x|^2) partials in a fixed order.
// One block of 256 threads: the sums in a fixed order, the residual, and
// the early-exit flag.
__device__ __forceinline__ void resid_reduce_block(const double* __restrict__ partials, int count, double tol,
                                                   int iter, StepDev* st) {
  if (st->done) return;  // uniform over the block
  __shared__ double sd[256], sn[256];
  double a = 0, b = 0;
  for (int k = threadIdx.x; k < count; k += blockDim.x) { a += partials[2 * k]; b += partials[2 * k + 1]; }
  sd[threadIdx.x] = a; sn[threadIdx.x] = b;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) { sd[threadIdx.x] += sd[threadIdx.x + s]; sn[threadIdx.x] += sn[threadIdx.x + s]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double res = sn[0] > 1e-12 ? sqrt(sd[0] / sn[0]) : sqrt(sd[0]);
    st->residual = res;
    st->iterations = iter + 1;
    if (res < tol) st->done = 1;
  }
}
__global__ void __launch_bounds__(256) k_resid_reduce(const double* __restrict__ partials, int count, double tol,
                                                      int iter, StepDev* st) {
  resid_reduce_block(partials, count, tol, iter, st);
}

// Prediction, state.rs:145-169 and pd_solver.rs:274-283: prev = x;
// pred = x + dt v + dt^2 g (pinned: pred = x; ground clamp); x = pred.
__global__ void __launch_bounds__(256) k_predict(int n, double* __restrict__ x, const double* __restrict__ v,
                                                 double* __restrict__ prev, double* __restrict__ pred,
                                                 const uint8_t* __restrict__ pinned, double dt, double gx, double gy,
                                                 double gz, int has_ground, double ground, StepDev* st) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) { st->done = 0; st->iterations = 0; st->residual = DBL_MAX; st->counter = 0; }
  if (i >= n) return;
  const double dt2 = dt * dt;
  double p[3];
  for (int c = 0; c < 3; ++c) { p[c] = x[3 * i + c]; prev[3 * i + c] = p[c]; }
  if (!pinned[i]) {
    p[0] = p[0] + dt * v[3 * i + 0] + dt2 * gx;
    p[1] = p[1] + dt * v[3 * i + 1] + dt2 * gy;
    p[2] = p[2] + dt * v[3 * i + 2] + dt2 * gz;
    if (has_ground && p[1] < ground) p[1] = ground;
  }
  for (int c = 0; c < 3; ++c) { pred[3 * i + c] = p[c]; x[3 * i + c] = p[c]; }
}

// Finalize, pd_solver.rs:447-466 and state.rs:195-240.
// With rs.st set, one block more than the vertices need forms the last
// iteration's residual (resid_reduce_block), so the step needs no separate
// launch for it.
struct ResidArgs {
  const double* partials;
  int count;
  double tol;
  int iter;
  StepDev* st;
};
__global__ void __launch_bounds__(256) k_finalize(int n, const double* __restrict__ x, double* __restrict__ prev,
                                                  double* __restrict__ v, const uint8_t* __restrict__ pinned,
                                                  double inv_dt, int has_ground, double ground, double fdamp,
                                                  double frayl, ResidArgs rs) {
  if (rs.st && (int)blockIdx.x == (int)gridDim.x - 1) {
    resid_reduce_block(rs.partials, rs.count, rs.tol, rs.iter, rs.st);
    return;
  }
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double vel[3];
  for (int c = 0; c < 3; ++c) vel[c] = (x[3 * i + c] - prev[3 * i + c]) * inv_dt;
  if (has_ground && !pinned[i] && x[3 * i + 1] <= ground) {
    if (vel[1] < 0.0) vel[1] = 0.0;
    prev[3 * i + 1] = ground;
    vel[0] *= 0.5; vel[2] *= 0.5;
  }
  for (int c = 0; c < 3; ++c) vel[c] *= fdamp;
  if (frayl > 0.0 && !pinned[i])
    for (int c = 0; c < 3; ++c) vel[c] *= frayl;
  for (int c = 0; c < 3; ++c) v[3 * i + c] = vel[c];
}

}  // namespace vhk
