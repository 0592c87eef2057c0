// gfx950 kernels of the IPC contact path: LBVH broad phase, vertex-triangle
// and edge-edge barrier forces, continuous collision, and the per-vertex
// stages of the augmented-Lagrangian step.
//
// Reference: crates/vistio-contact/src/{barrier,distance_primitives,ccd,bvh,
// self_collision,exclusion,ipc_response,ground_plane,sphere,
// collision_pipeline}.rs and ProjectiveDynamicsSolver::step_with_ipc
// (crates/vistio-solver/src/pd_solver.rs:549-1157). Everything is FP64 except
// the broad phase, whose boxes are f32 (bvh.rs:15-62), so that the candidate
// pairs are exactly the overlapping pairs of the reference's f32 boxes.
//
// Broad phase: Karras' LBVH (Morton codes of box centroids, radix-sorted;
// internal nodes from common-prefix splits; node bounds as unions over their
// leaf ranges from a sparse table). Every leaf traverses the tree for leaves that
// follow it in Morton order, so each overlapping pair is emitted once. The
// reference's recursive median-split query drops the pairs whose left-subtree
// index is larger (bvh.rs:207); here the pair set is complete.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

// No a*b+c contraction in this file: the contact tests (distances, the CCD
// cubic's coefficients and roots, barycentric inside tests) take branch
// decisions on values with heavy cancellation, and must round like the
// reference's separate multiplies and adds. Restored at the end of the file.
#pragma clang fp contract(off)

namespace vhc {

// ---- small FP64 vector ---------------------------------------------------

struct D3 {
  double x, y, z;
};
__device__ __forceinline__ D3 d3(double x, double y, double z) { return D3{x, y, z}; }
__device__ __forceinline__ D3 operator+(D3 a, D3 b) { return d3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ D3 operator-(D3 a, D3 b) { return d3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ D3 operator-(D3 a) { return d3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ D3 operator*(D3 a, double s) { return d3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 cross(D3 a, D3 b) {
  return d3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double len2(D3 a) { return dot(a, a); }
__device__ __forceinline__ double clamp01(double v) { return fmin(fmax(v, 0.0), 1.0); }
__device__ __forceinline__ D3 ld(const double* x, int i) { return d3(x[3 * i], x[3 * i + 1], x[3 * i + 2]); }


// ---- barrier.rs ------------------------------------------------------------

// barrier.rs:40-53: db/dd of b(d) = -(d - dh)^2 ln(d/dh), d clamped to
// [1e-3 dh, dh].
__device__ __forceinline__ double barrier_gradient(double d, double dh) {
  const double dc = fmin(fmax(d, dh * 1e-3), dh);
  if (dc >= dh) return 0.0;
  const double diff = dc - dh;
  return -diff * (2.0 * log(dc / dh) + diff / dc);
}
// barrier.rs:53-63: d2b/dd2, 0 outside (0, dh).
__device__ __forceinline__ double barrier_hessian(double d, double dh) {
  if (d >= dh || d <= 0.0) return 0.0;
  const double diff = d - dh, inv = 1.0 / d;
  return -2.0 * log(d / dh) - 4.0 * diff * inv + diff * diff * inv * inv;
}
// barrier.rs:82-93
__device__ __forceinline__ double scaled_barrier_gradient(double d, double dh, double kappa) {
  const double r = kappa * barrier_gradient(d, dh);
  return isfinite(r) ? r : copysign(1e6, r);
}
// barrier.rs:140-163: thickness offset d - th^2 (non-positive: d = 1e-12).
__device__ __forceinline__ double scaled_barrier_gradient_thick(double d, double dh, double kappa, double th) {
  const double de = d - th * th;
  const double r = kappa * barrier_gradient(de <= 0.0 ? 1e-12 : de, dh);
  return isfinite(r) ? r : copysign(1e6, r);
}

// ---- distance_primitives.rs ------------------------------------------------

__device__ __forceinline__ double point_segment_dist2(D3 p, D3 a, D3 b) {
  const D3 ab = b - a, ap = p - a;
  const double l2 = len2(ab);
  if (l2 < 1e-20) return len2(ap);
  const double t = clamp01(dot(ap, ab) / l2);
  return len2(p - (a + ab * t));
}

// distance_primitives.rs:36-116 (Ericson's Voronoi-region closest point).
__device__ double pt_dist2(D3 p, D3 a, D3 b, D3 c) {
  const D3 ab = b - a, ac = c - a, ap = p - a, bp = p - b, cp = p - c;
  if (len2(cross(ab, ac)) < 1e-20) {
    const double e0 = point_segment_dist2(p, a, b), e1 = point_segment_dist2(p, b, c), e2 = point_segment_dist2(p, c, a);
    if (e0 <= e1 && e0 <= e2) return e0;
    return e1 <= e2 ? e1 : e2;
  }
  const double d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0 && d2 <= 0.0) return len2(ap);
  const double d3v = dot(ab, bp), d4 = dot(ac, bp);
  if (d3v >= 0.0 && d4 <= d3v) return len2(bp);
  const double d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0 && d5 <= d6) return len2(cp);
  const double vc = d1 * d4 - d3v * d2;
  if (vc <= 0.0 && d1 >= 0.0 && d3v <= 0.0) return len2(p - (a + ab * (d1 / (d1 - d3v))));
  const double va = d3v * d6 - d5 * d4;
  if (va <= 0.0 && (d4 - d3v) >= 0.0 && (d5 - d6) >= 0.0)
    return len2(p - (b + (c - b) * ((d4 - d3v) / ((d4 - d3v) + (d5 - d6)))));
  const double vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) return len2(p - (a + ac * (d2 / (d2 - d6))));
  const double den = 1.0 / (va + vb + vc);
  return len2(p - (a + ab * (vb * den) + ac * (vc * den)));
}

// distance_primitives.rs:124-177 (gradient through the plane projection).
__device__ void pt_grad(D3 p, D3 a, D3 b, D3 c, D3 g[4]) {
  g[0] = g[1] = g[2] = g[3] = d3(0, 0, 0);
  const D3 ab = b - a, ac = c - a;
  const D3 nrm = cross(ab, ac);
  const double a2 = len2(nrm);
  if (a2 < 1e-20) return;
  const D3 n = nrm * (1.0 / sqrt(a2));
  const D3 proj = p - n * dot(p - a, n);
  const D3 v2 = proj - a;
  const double d00 = dot(ab, ab), d01 = dot(ab, ac), d11 = dot(ac, ac), d20 = dot(v2, ab), d21 = dot(v2, ac);
  const double den = d00 * d11 - d01 * d01;
  if (fabs(den) < 1e-20) return;
  const double inv = 1.0 / den;
  const double v = (d11 * d20 - d01 * d21) * inv, w = (d00 * d21 - d01 * d20) * inv, u = 1.0 - v - w;
  const D3 diff = p - (a * u + b * v + c * w);
  g[0] = diff * 2.0;
  g[1] = -diff * (2.0 * u);
  g[2] = -diff * (2.0 * v);
  g[3] = -diff * (2.0 * w);
}

// distance_primitives.rs:185-246 (segment-segment, eps = 1e-10 branches).
__device__ double ee_dist2(D3 a0, D3 a1, D3 b0, D3 b1) {
  const D3 d1 = a1 - a0, d2 = b1 - b0, r = a0 - b0;
  const double a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
  const double eps = 1e-10;
  if (a < eps && e < eps) return len2(r);
  double s, t;
  if (a < eps) {
    s = 0.0; t = clamp01(f / e);
  } else {
    const double c = dot(d1, r);
    if (e < eps) {
      t = 0.0; s = clamp01(-c / a);
    } else {
      const double b = dot(d1, d2);
      const double den = a * e - b * b;
      s = fabs(den) > eps ? clamp01((b * f - c * e) / den) : 0.0;
      t = (b * s + f) / e;
      if (t < 0.0) { t = 0.0; s = clamp01(-c / a); }
      else if (t > 1.0) { t = 1.0; s = clamp01((b - c) / a); }
    }
  }
  return len2((a0 + d1 * s) - (b0 + d2 * t));
}

// distance_primitives.rs:251-303
__device__ void ee_grad(D3 a0, D3 a1, D3 b0, D3 b1, D3 g[4]) {
  const D3 d1 = a1 - a0, d2 = b1 - b0, r = a0 - b0;
  const double a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r), c = dot(d1, r), b = dot(d1, d2);
  const double eps = 1e-10;
  double s = 0.0, t = 0.0;
  if (!(a < eps || e < eps)) {
    const double den = a * e - b * b;
    s = fabs(den) > eps ? clamp01((b * f - c * e) / den) : 0.0;
    t = (b * s + f) / e;
    if (t < 0.0) { t = 0.0; s = clamp01(-c / a); }
    else if (t > 1.0) { t = 1.0; s = clamp01((b - c) / a); }
  }
  const D3 diff = (a0 + d1 * s) - (b0 + d2 * t);
  g[0] = diff * (2.0 * (1.0 - s));
  g[1] = diff * (2.0 * s);
  g[2] = -diff * (2.0 * (1.0 - t));
  g[3] = -diff * (2.0 * t);
}

// ---- ccd.rs ----------------------------------------------------------------

// ccd.rs:352-377
__device__ int quadratic_roots(double a, double b, double c, double out[3]) {
  const double eps = 1e-8;
  if (fabs(a) < eps) {
    if (fabs(b) < eps) return 0;
    out[0] = -c / b;
    return 1;
  }
  const double disc = b * b - 4.0 * a * c;
  if (disc < -eps) return 0;
  if (fabs(disc) <= eps) { out[0] = -b / (2.0 * a); return 1; }
  const double sq = sqrt(disc);
  out[0] = (-b + sq) / (2.0 * a);
  out[1] = (-b - sq) / (2.0 * a);
  return 2;
}

// ccd.rs:299-349 (Cardano; trigonometric form for three real roots)
__device__ int cubic_roots(double a, double b, double c, double d, double out[3]) {
  const double eps = 1e-8;
  if (fabs(a) < eps) return quadratic_roots(b, c, d, out);
  const double p = b / a, q = c / a, r = d / a;
  const double ad = q - p * p / 3.0;
  const double bd = r - p * q / 3.0 + 2.0 * p * p * p / 27.0;
  const double disc = bd * bd / 4.0 + ad * ad * ad / 27.0;
  const double shift = -p / 3.0;
  if (disc > eps) {
    const double sq = sqrt(disc);
    out[0] = cbrt(-bd / 2.0 + sq) + cbrt(-bd / 2.0 - sq) + shift;
    return 1;
  }
  if (fabs(disc) <= eps) {
    if (fabs(bd) < eps) { out[0] = shift; return 1; }
    const double u = cbrt(bd / 2.0);
    out[0] = -2.0 * u + shift;
    out[1] = u + shift;
    return 2;
  }
  const double rv = sqrt(-ad * ad * ad / 27.0);
  if (fabs(rv) < 1e-12) { out[0] = shift; return 1; }
  const double cv = fmin(fmax(-bd / (2.0 * rv), -1.0), 1.0);
  const double th = acos(cv);
  const double m = 2.0 * cbrt(rv);
  const double pi = 3.14159265358979323846;
  out[0] = m * cos(th / 3.0) + shift;
  out[1] = m * cos((th + 2.0 * pi) / 3.0) + shift;
  out[2] = m * cos((th + 4.0 * pi) / 3.0) + shift;
  return 3;
}

// ccd.rs:32-129: smallest safe TOI (0.9 t, at least 1e-6) of vertex p
// against the moving triangle (a, b, c), or 2 (no hit).
__device__ double vt_ccd(D3 p0, D3 p1, D3 a0, D3 a1, D3 b0, D3 b1, D3 c0, D3 c1) {
  const D3 dp = p1 - p0;
  const D3 ar = a0 - p0, br = b0 - p0, cr = c0 - p0;
  const D3 av = (a1 - a0) - dp, bv = (b1 - b0) - dp, cv = (c1 - c0) - dp;
  const D3 vv = cross(bv, cv), rv = cross(br, cv), vr = cross(bv, cr), rr = cross(br, cr);
  const double ca = dot(av, vv);
  const double cb = dot(ar, vv) + dot(av, rv + vr);
  const double cc = dot(ar, rv + vr) + dot(av, rr);
  const double cd = dot(ar, rr);
  double roots[3];
  const int nr = cubic_roots(ca, cb, cc, cd, roots);
  double best = 2.0;
  for (int k = 0; k < nr; ++k) {
    double t = roots[k];
    if (t < -1e-6 || t > 1.0 + 1e-6) continue;
    t = clamp01(t);
    const D3 pa = ar + av * t, pb = br + bv * t, pc = cr + cv * t;
    if (len2(cross(pb - pa, pc - pa)) < 1e-20) continue;
    const D3 v0 = pb - pa, v1 = pc - pa, v2 = -pa;
    const double d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1), d20 = dot(v2, v0), d21 = dot(v2, v1);
    const double den = d00 * d11 - d01 * d01;
    if (fabs(den) < 1e-20) continue;
    const double inv = 1.0 / den;
    const double u = (d11 * d20 - d01 * d21) * inv, w = (d00 * d21 - d01 * d20) * inv, v = 1.0 - u - w;
    if (v >= -0.01 && u >= -0.01 && w >= -0.01) best = fmin(best, fmax(t * 0.9, 1e-6));
  }
  return best;
}

// ccd.rs:138-229: smallest safe TOI (0.9 t, at least 1e-6) of edge (a, b)
// against edge (c, d), or 2 (no hit). The coplanarity cubic of
// (c - a) . ((b - a) x (d - c)); at each root in [-1e-6, 1 + 1e-6] the closest
// points' parameters are clamped to [0, 1] before the 0.05 tolerance test, as
// the reference does, so every root with non-degenerate, non-parallel edges
// is a hit.
__device__ double ee_ccd(D3 a0, D3 a1, D3 b0, D3 b1, D3 c0, D3 c1, D3 d0, D3 d1) {
  const D3 da = a1 - a0, db = b1 - b0, dc = c1 - c0, dd = d1 - d0;
  const D3 r0 = c0 - a0, rv = dc - da, e10 = b0 - a0, e1v = db - da, e20 = d0 - c0, e2v = dd - dc;
  const D3 vv = cross(e1v, e2v), rvx = cross(e10, e2v), vrx = cross(e1v, e20), rr = cross(e10, e20);
  const double ca = dot(rv, vv);
  const double cb = dot(r0, vv) + dot(rv, rvx + vrx);
  const double cc = dot(r0, rvx + vrx) + dot(rv, rr);
  const double cd = dot(r0, rr);
  double roots[3];
  const int nr = cubic_roots(ca, cb, cc, cd, roots);
  double best = 2.0;
  for (int k = 0; k < nr; ++k) {
    double t = roots[k];
    if (t < -1e-6 || t > 1.0 + 1e-6) continue;
    t = clamp01(t);
    const D3 at = a0 + da * t, bt = b0 + db * t, ct = c0 + dc * t, dtt = d0 + dd * t;
    const D3 e1 = bt - at, e2 = dtt - ct, r = ct - at;
    const double l1 = dot(e1, e1), l2 = dot(e2, e2);
    if (l1 < 1e-20 || l2 < 1e-20) continue;
    const double bb = dot(e1, e2), ff = dot(e2, r), c2 = dot(e1, r);
    const double den = l1 * l2 - bb * bb;
    if (fabs(den) < 1e-10) continue;
    const double sp = clamp01((bb * ff - c2 * l2) / den);
    const double up = clamp01((bb * sp + ff) / l2);
    if (sp >= -0.05 && sp <= 1.05 && up >= -0.05 && up <= 1.05) best = fmin(best, fmax(t * 0.9, 1e-6));
  }
  return best;
}

// ---- LBVH (bvh.rs semantics: f32 boxes, closed-interval overlap) ----------

struct Box {
  float lo[3], hi[3];
};

__device__ __forceinline__ bool overlaps(const Box& a, const Box& b) {
  return a.lo[0] <= b.hi[0] && a.hi[0] >= b.lo[0] && a.lo[1] <= b.hi[1] && a.hi[1] >= b.lo[1] &&
         a.lo[2] <= b.hi[2] && a.hi[2] >= b.lo[2];
}

// Order-preserving float <-> uint for atomic min/max.
__device__ __forceinline__ unsigned int f2o(float f) {
  const unsigned int u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float o2f(unsigned int u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

struct BvhDev {
  unsigned int scene[6];  // order-mapped lo xyz, hi xyz of the box centroids
  unsigned int npairs;    // pairs found (may exceed the capacity)
  unsigned int ncont;     // self contacts of the last detection (may exceed the entry buffer)
  unsigned int crossings; // triangle pairs that cross (k_tri_crossings)
  unsigned int walk_more; // a leaf has more pairs than k_bvh_walk1 keeps (k_bvh_more walks it again)
};

// bvh.rs:28-42: box of f32 positions, inflated by margin. Each block also
// writes the bounds of its box centroids (min xyz, max xyz) to part.
__global__ void __launch_bounds__(256) k_boxes(int t, const int* __restrict__ tri, const double* __restrict__ x,
                                               float margin, Box* __restrict__ box, float* __restrict__ part,
                                               const int* skip) {
  if (skip && *skip) return;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  float c[3] = {0, 0, 0};
  bool ok = e < t;
  if (ok) {
    Box b;
    const int v0 = tri[3 * e], v1 = tri[3 * e + 1], v2 = tri[3 * e + 2];
    for (int a = 0; a < 3; ++a) {
      const float p0 = (float)x[3 * v0 + a], p1 = (float)x[3 * v1 + a], p2 = (float)x[3 * v2 + a];
      b.lo[a] = fminf(fminf(p0, p1), p2) - margin;
      b.hi[a] = fmaxf(fmaxf(p0, p1), p2) + margin;
      c[a] = 0.5f * (b.lo[a] + b.hi[a]);
    }
    box[e] = b;
  }
  // block reduction of the centroid bounds, one atomic per block
  __shared__ float smin[3][256], smax[3][256];
  for (int a = 0; a < 3; ++a) {
    smin[a][threadIdx.x] = ok ? c[a] : FLT_MAX;
    smax[a][threadIdx.x] = ok ? c[a] : -FLT_MAX;
  }
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int a = 0; a < 3; ++a) {
        smin[a][threadIdx.x] = fminf(smin[a][threadIdx.x], smin[a][threadIdx.x + s]);
        smax[a][threadIdx.x] = fmaxf(smax[a][threadIdx.x], smax[a][threadIdx.x + s]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int a = 0; a < 3; ++a) {
      part[6 * blockIdx.x + a] = smin[a][0];
      part[6 * blockIdx.x + 3 + a] = smax[a][0];
    }
}
// Scene bounds from the per-block partials of k_boxes (one block).
__global__ void __launch_bounds__(256) k_scene_bounds(int nb, const float* __restrict__ part, BvhDev* bd,
                                                      const int* skip) {
  if (skip && *skip) return;
  __shared__ float smin[3][256], smax[3][256];
  float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  for (int b = threadIdx.x; b < nb; b += blockDim.x)
    for (int a = 0; a < 3; ++a) {
      lo[a] = fminf(lo[a], part[6 * b + a]);
      hi[a] = fmaxf(hi[a], part[6 * b + 3 + a]);
    }
  for (int a = 0; a < 3; ++a) { smin[a][threadIdx.x] = lo[a]; smax[a][threadIdx.x] = hi[a]; }
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int a = 0; a < 3; ++a) {
        smin[a][threadIdx.x] = fminf(smin[a][threadIdx.x], smin[a][threadIdx.x + s]);
        smax[a][threadIdx.x] = fmaxf(smax[a][threadIdx.x], smax[a][threadIdx.x + s]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int a = 0; a < 3; ++a) { bd->scene[a] = f2o(smin[a][0]); bd->scene[3 + a] = f2o(smax[a][0]); }
}

// Scene bounds to empty, pair count to 0.
__global__ void k_bvh_reset(BvhDev* bd, const int* skip) {
  if (threadIdx.x != 0 || (skip && *skip)) return;
  for (int a = 0; a < 3; ++a) { bd->scene[a] = f2o(FLT_MAX); bd->scene[3 + a] = f2o(-FLT_MAX); }
  bd->npairs = 0;
  bd->crossings = 0;
  bd->walk_more = 0;
}

__device__ __forceinline__ unsigned int expand_bits(unsigned int v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

// 30-bit Morton code of each box centroid in the scene bounds.
__global__ void __launch_bounds__(256) k_morton(int t, const Box* __restrict__ box, const BvhDev* bd,
                                                unsigned int* __restrict__ code, int* __restrict__ id,
                                                const int* skip) {
  if (skip && *skip) return;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= t) return;
  unsigned int q[3];
  for (int a = 0; a < 3; ++a) {
    const float lo = o2f(bd->scene[a]), hi = o2f(bd->scene[3 + a]);
    const float c = 0.5f * (box[e].lo[a] + box[e].hi[a]);
    const float ext = hi - lo;
    float u = ext > 0.0f ? (c - lo) / ext : 0.0f;
    u = fminf(fmaxf(u * 1024.0f, 0.0f), 1023.0f);
    q[a] = (unsigned int)u;
  }
  code[e] = (expand_bits(q[0]) << 2) | (expand_bits(q[1]) << 1) | expand_bits(q[2]);
  id[e] = e;
}

// Common-prefix length of sorted keys i and j (ties broken by index).
__device__ __forceinline__ int delta(const unsigned int* code, int t, int i, int j) {
  if (j < 0 || j >= t) return -1;
  const unsigned int a = code[i], b = code[j];
  if (a != b) return __clz(a ^ b);
  return 32 + __clz((unsigned int)(i ^ j));
}

// Karras (2012) internal node i of t - 1: its key range [lo, hi] and its
// left child (>= 0 internal node, < 0 leaf ~j).
__global__ void __launch_bounds__(256) k_karras(int t, const unsigned int* __restrict__ code, int* __restrict__ left,
                                                int* __restrict__ range_lo, int* __restrict__ range_hi,
                                                const int* skip) {
  if (skip && *skip) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= t - 1) return;
  const int d = (delta(code, t, i, i + 1) - delta(code, t, i, i - 1)) >= 0 ? 1 : -1;
  const int dmin = delta(code, t, i, i - d);
  int lmax = 2;
  while (delta(code, t, i, i + lmax * d) > dmin) lmax <<= 1;
  int l = 0;
  for (int s = lmax >> 1; s >= 1; s >>= 1)
    if (delta(code, t, i, i + (l + s) * d) > dmin) l += s;
  const int j = i + l * d;
  const int dnode = delta(code, t, i, j);
  int sp = 0;
  for (int s = (l + 1) >> 1;; s = (s + 1) >> 1) {
    if (delta(code, t, i, i + (sp + s) * d) > dnode) sp += s;
    if (s == 1) break;
  }
  const int g = i + sp * d + (d < 0 ? d : 0);
  const int lo = i < j ? i : j, hi = i < j ? j : i;
  left[i] = lo == g ? ~g : g;  // the right child (g + 1) is reached through escapes
  range_lo[i] = lo;
  range_hi[i] = hi;
}

// Node bounds without inter-thread communication. Internal node i covers
// the sorted leaves [lo, hi], so its box is the union of those leaf boxes:
// a sparse table st[k][j] = union of leaves j .. j + 2^k - 1 answers it with
// two overlapping windows of length 2^floor(log2(hi - lo + 1)).
__device__ __forceinline__ Box box_union(const Box& a, const Box& b) {
  Box m;
  for (int q = 0; q < 3; ++q) { m.lo[q] = fminf(a.lo[q], b.lo[q]); m.hi[q] = fmaxf(a.hi[q], b.hi[q]); }
  return m;
}
// Levels 0 .. levels - 1 (at most kStLowLevels + 1) in one launch: block b
// gathers the leaf boxes of [256 b, 256 b + 512) into LDS and forms each
// level of its 256 entries from the previous one there (level k at j needs
// level k - 1 at j and j + 2^(k-1), so a 256-entry halo covers 8 levels).
// Level 0 is the leaf boxes in Morton order; level k at j is the union of
// leaves [j, j + 2^k), clipped at t.
constexpr int kStLowLevels = 8;
__global__ void __launch_bounds__(256) k_st_low(int t, int levels, const Box* __restrict__ box,
                                                const int* __restrict__ id, Box* __restrict__ st, const int* skip) {
  if (skip && *skip) return;
  __shared__ Box w[512];
  const int base = blockIdx.x * 256, tid = threadIdx.x;
  for (int i = tid; i < 512; i += 256) {
    const int j = base + i;
    if (j < t) w[i] = box[id[j]];
  }
  if (base + tid < t) st[base + tid] = w[tid];
  __syncthreads();
  for (int k = 1; k < levels; ++k) {
    const int half = 1 << (k - 1), valid = 512 - ((1 << k) - 1);  // entries of level k this block can form
    Box nv[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int i = tid + 256 * u, j = base + i;
      if (i < valid && j < t) nv[u] = j + half < t ? box_union(w[i], w[i + half]) : w[i];
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int i = tid + 256 * u, j = base + i;
      if (i < valid && j < t) w[i] = nv[u];
    }
    __syncthreads();
    if (base + tid < t) st[(size_t)k * t + base + tid] = w[tid];
  }
}
// Wider ranges: cell c is leaves [256 c, 256 c + 256) (its box is level 8
// at 256 c), for the nc = t / 256 whole cells; coarse level m at c is the
// union of cells [c, c + 2^m) (clipped at nc), stored at coarse[m nc + c].
// One block forms all levels (each from the previous, a barrier apart), in
// LDS when the cells fit.
constexpr int kCoarseThreads = 1024;
constexpr int kCoarseLds = 2048;  // cells held in LDS (48 KB); more go through global memory
__global__ void __launch_bounds__(kCoarseThreads) k_st_coarse(int t, int nc, int levels, const Box* __restrict__ st8,
                                                             Box* __restrict__ coarse, const int* skip) {
  if (skip && *skip) return;
  if (nc <= kCoarseLds) {
    __shared__ Box w[kCoarseLds];
    for (int c = threadIdx.x; c < nc; c += blockDim.x) {
      const Box b = st8[(size_t)256 * c];
      w[c] = b;
      coarse[c] = b;
    }
    __syncthreads();
    for (int m = 1; m < levels; ++m) {
      const int half = 1 << (m - 1);
      Box nv[kCoarseLds / kCoarseThreads];
#pragma unroll
      for (int u = 0; u < kCoarseLds / kCoarseThreads; ++u) {
        const int c = threadIdx.x + u * kCoarseThreads;
        if (c < nc) nv[u] = c + half < nc ? box_union(w[c], w[c + half]) : w[c];
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < kCoarseLds / kCoarseThreads; ++u) {
        const int c = threadIdx.x + u * kCoarseThreads;
        if (c < nc) {
          w[c] = nv[u];
          coarse[(size_t)m * nc + c] = nv[u];
        }
      }
      __syncthreads();
    }
    return;
  }
  for (int c = threadIdx.x; c < nc; c += blockDim.x) coarse[c] = st8[(size_t)256 * c];
  __syncthreads();
  for (int m = 1; m < levels; ++m) {
    const int half = 1 << (m - 1);
    const Box* prev = coarse + (size_t)(m - 1) * nc;
    Box* cur = coarse + (size_t)m * nc;
    for (int c = threadIdx.x; c < nc; c += blockDim.x) cur[c] = c + half < nc ? box_union(prev[c], prev[c + half]) : prev[c];
    __syncthreads();
  }
}
// Traversal records. A subtree is skipped by jumping to its escape (rope):
// for a range ending at leaf hi that is the highest node starting at hi + 1,
// the right child of the split between hi and hi + 1. In a Karras tree a
// right child is internal node s exactly when its range starts at s, so the
// escape is internal node hi + 1 if range_lo[hi + 1] == hi + 1, else leaf
// hi + 1, and kEnd after the last leaf.
// Internal nodes and leaves share one 32-byte record, so a walk step is one
// load whichever it visits. A node's range ends at leaf hi with its escape
// encoding hi + 1 (kEnd: hi = t - 1), so hi needs no field of its own.
constexpr int kEnd = 0x7fffffff;
struct __align__(16) Rec {
  Box box;
  int a;       // node: left child (>= 0 internal node, < 0 leaf ~j); leaf: its triangle
  int escape;  // same encoding as left, or kEnd
};
__device__ __forceinline__ int escape_of(int hi, int t, const int* __restrict__ range_lo) {
  if (hi >= t - 1) return kEnd;
  return (hi + 1 <= t - 2 && range_lo[hi + 1] == hi + 1) ? hi + 1 : ~(hi + 1);
}
__global__ void __launch_bounds__(256) k_node_bounds(int t, const Box* __restrict__ st, const Box* __restrict__ coarse,
                                                     int nc, const int* __restrict__ id,
                                                     const int* __restrict__ left, const int* __restrict__ range_lo,
                                                     const int* __restrict__ range_hi, Rec* __restrict__ node,
                                                     Rec* __restrict__ leaf, const int* __restrict__ orig,
                                                     int* __restrict__ sorig, const int* skip) {
  if (skip && *skip) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < t) {
    Rec L;
    L.box = st[i];
    L.escape = escape_of(i, t, range_lo);
    L.a = id[i];
    leaf[i] = L;
    sorig[i] = orig[L.a];  // the caller's index of sorted leaf i
  }
  if (i >= t - 1) return;
  const int lo = range_lo[i], hi = range_hi[i];
  const int k = 31 - __clz((unsigned int)(hi - lo + 1));
  Rec N;
  if (k <= kStLowLevels) {
    // two windows of 2^k leaves that cover [lo, hi]
    const Box* lv = st + (size_t)k * (size_t)t;
    N.box = box_union(lv[lo], lv[hi - (1 << k) + 1]);
  } else {
    // at least 512 leaves: the 256-leaf windows at both ends, and the whole
    // cells [a, b) between them from the coarse table (two windows again)
    const Box* l8 = st + (size_t)kStLowLevels * (size_t)t;
    const int a = (lo + 255) >> 8, b = (hi + 1) >> 8;
    const int m = 31 - __clz((unsigned int)(b - a));
    const Box* cm = coarse + (size_t)m * nc;
    N.box = box_union(box_union(l8[lo], l8[hi - 255]), box_union(cm[a], cm[b - (1 << m)]));
  }
  N.a = left[i];
  N.escape = escape_of(hi, t, range_lo);
  node[i] = N;
}

// Every leaf k collects the leaves j > k (Morton order) whose boxes overlap
// its own: each unordered pair once. One walk counts each leaf's pairs and
// keeps them (k_bvh_walk1), an exclusive scan turns the counts into
// offsets, and the kept pairs are copied there (k_bvh_compact); a leaf
// with more pairs than the walk keeps is walked again and writes its pairs
// at its offset (k_bvh_more). No shared counter, so the list is the same
// every time. A pair is stored as (a, b) with the caller's triangle
// index orig[a] < orig[b]: the reference tests vertices of its first
// triangle against the second and edges of the first against the second
// (self_collision.rs:199-276), and for parallel edges the closest points,
// hence the gradient, depend on that order.
// Stackless: pre-order walk, from the node after leaf k, that enters a node
// when its box overlaps and its range reaches past k, and otherwise takes
// its escape; every step moves forward in pre-order, so the walk ends at
// kEnd.
// The LDS stage of a workgroup's walks, written out in whole lines: 32 KB,
// 4096 pairs, 16 per leaf. (With a stage shared by the workgroup's leaves
// and the pairs written at their final offsets, 16 KB left two thirds of
// the flat 512x512 drape's pairs to scattered 8-byte stores whose lines the
// walk's node misses evicted half written, 265 MB written for 103 MB of
// pairs; 32 KB, five workgroups per CU instead of eight, ran the drape
// 2.4 % faster; 48 KB was slower again, that walk 340 -> 357 us.)
constexpr int kPairStage = 4096;

// The walk of leaf k; with kEmit, sink(c, pair) takes the leaf's c-th pair.
template <bool kEmit, class Sink>
__device__ __forceinline__ unsigned int bvh_walk(int k, int t, const Rec* __restrict__ node, const Rec* __restrict__ leaf,
                                                 const int* __restrict__ sorig, Sink&& sink) {
  const Rec me = leaf[k];
  const int ome = kEmit ? sorig[k] : 0;
  unsigned int cnt = 0;
  // Start after leaf k in pre-order: the nodes there are the right subtrees
  // of k's ancestors, i.e. exactly the ranges past k, so the descent from the
  // root to k (which only re-enters k's ancestors) is skipped.
  int nd = me.escape;
  while (nd != kEnd) {
    const bool is_leaf = nd < 0;
    const int j = ~nd;
    const Rec* rp = is_leaf ? leaf + j : node + nd;  // one load for either kind
    const Rec R = *rp;
    const int oj = kEmit ? sorig[is_leaf ? j : k] : 0;  // loaded with the record, not after it
    const bool ov = overlaps(me.box, R.box);
    if (is_leaf) {
      if (j > k && ov) {
        if (kEmit) {
          const bool first = ome < oj;
          sink(cnt, make_int2(first ? me.a : R.a, first ? R.a : me.a));
        }
        ++cnt;
      }
      nd = R.escape;
    } else {
      // enter the node when its box overlaps and its range (ending at hi,
      // hi + 1 encoded in the escape) reaches past k
      const int e = R.escape;
      const int h1 = e == kEnd ? t : (e >= 0 ? e : ~e);
      nd = (h1 > k + 1 && ov) ? R.a : e;
    }
  }
  return cnt;
}
// Leaf chunk of block b for the walks: the chunks of 256 Morton-ordered
// leaves are cut into one contiguous run per XCD (workgroups go to the eight
// XCDs in turn; the grid is 8 * ceil(chunks / 8) blocks), so neighbouring
// walks, which visit the same nodes, share an L2. -1: an idle block.
__device__ __forceinline__ int leaf_chunk(int t) {
  constexpr int kX = 8;
  const int nch = (t + (int)blockDim.x - 1) / (int)blockDim.x;
  const int per = (nch + kX - 1) / kX;
  const int c = ((int)blockIdx.x % kX) * per + (int)blockIdx.x / kX;
  return c < nch ? c : -1;
}

// One walk per leaf for counts and pairs alike. k_bvh_walk1 counts each
// leaf's pairs (count) and keeps them: the first kLeafStage in the
// workgroup's LDS stage, written out whole as the workgroup's run of kept
// (leaf k's at kept[k kLeafStage ..]), the next c2 in leaf k's run of more
// (more[k c2 ..]). A leaf with still more sets bd->walk_more. After the scan
// of count, k_bvh_compact copies the kept runs to the pair list at the
// leaves' offsets, and k_bvh_more walks the leaves with more pairs again,
// writing theirs at their offsets. The list is the leaves' pairs in leaf
// order, each leaf's in walk order. On the flat 512 x 512 drape (about 25
// pairs per leaf) this replaces a count walk (about 250 us) by the copy.
constexpr int kLeafStage = kPairStage / 256;  // 16 pairs per leaf in LDS

__global__ void __launch_bounds__(256) k_bvh_walk1(int t, const Rec* __restrict__ node, const Rec* __restrict__ leaf,
                                                   const int* __restrict__ sorig, unsigned int* __restrict__ count,
                                                   int2* __restrict__ kept, int2* __restrict__ more, int c2,
                                                   BvhDev* bd, const int* skip) {
  if (skip && *skip) return;  // uniform over the grid
  __shared__ int2 stage[kPairStage];
  const int ch = leaf_chunk(t);
  if (ch < 0) return;  // uniform over the block
  const int k = ch * (int)blockDim.x + (int)threadIdx.x;
  if (k < t) {
    int2* mine = stage + threadIdx.x * kLeafStage;
    int2* over = more + (size_t)k * c2;
    const unsigned int n = t < 2 ? 0u : bvh_walk<true>(k, t, node, leaf, sorig, [&](unsigned int c, int2 pr) {
      if (c < (unsigned int)kLeafStage) mine[c] = pr;
      else if (c - kLeafStage < (unsigned int)c2) over[c - kLeafStage] = pr;
    });
    count[k] = n;
    if (n > (unsigned int)(kLeafStage + c2)) bd->walk_more = 1u;
  }
  __syncthreads();
  // the whole stage (entries past a leaf's count are never read)
  int2* out = kept + (size_t)ch * kPairStage;
  for (int q = threadIdx.x; q < kPairStage; q += blockDim.x) out[q] = stage[q];
}

// offset: exclusive scan of count. Block b copies the pairs of leaves
// [256 b, 256 b + 256) to [offset[256 b], ...): each output slot finds its
// leaf by a binary search over the block's offsets in LDS.
__global__ void __launch_bounds__(256) k_bvh_compact(int t, const unsigned int* __restrict__ count,
                                                     const unsigned int* __restrict__ offset,
                                                     const int2* __restrict__ kept, const int2* __restrict__ more,
                                                     int c2, int2* __restrict__ pairs, unsigned int cap, BvhDev* bd,
                                                     const int* skip) {
  if (skip && *skip) return;  // uniform over the grid
  __shared__ unsigned int so[257];
  const int k0 = (int)blockIdx.x * 256;
  if (k0 >= t) return;
  const int kn = min(256, t - k0);
  if ((int)threadIdx.x < kn) so[threadIdx.x] = offset[k0 + threadIdx.x];
  if (threadIdx.x == 0) {
    so[kn] = offset[k0 + kn - 1] + count[k0 + kn - 1];
    if (k0 + kn == t) bd->npairs = so[kn];
  }
  __syncthreads();
  const unsigned int e = min(so[kn], cap);
  for (unsigned int o = so[0] + threadIdx.x; o < e; o += blockDim.x) {
    int lo = 0, hi = kn - 1;  // the last leaf whose run starts at or before o
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (so[mid] <= o) lo = mid;
      else hi = mid - 1;
    }
    if (so[lo + 1] - so[lo] > (unsigned int)(kLeafStage + c2)) continue;  // k_bvh_more writes this leaf
    const unsigned int j = o - so[lo];
    const size_t k = (size_t)(k0 + lo);
    pairs[o] = j < (unsigned int)kLeafStage ? kept[k * kLeafStage + j] : more[k * c2 + (j - kLeafStage)];
  }
}

// The leaves with more pairs than k_bvh_walk1 keeps (kLeafStage + c2),
// walked again; each writes its pairs at its offset.
__global__ void __launch_bounds__(256) k_bvh_more(int t, const Rec* __restrict__ node, const Rec* __restrict__ leaf,
                                                  const int* __restrict__ sorig, const unsigned int* __restrict__ count,
                                                  const unsigned int* __restrict__ offset, int c2,
                                                  int2* __restrict__ pairs, unsigned int cap, const BvhDev* bd,
                                                  const int* skip) {
  if ((skip && *skip) || !bd->walk_more) return;  // uniform over the grid
  const int ch = leaf_chunk(t);
  if (ch < 0) return;
  const int k = ch * (int)blockDim.x + (int)threadIdx.x;
  if (k >= t || count[k] <= (unsigned int)(kLeafStage + c2)) return;
  const unsigned int base = offset[k];
  if (base >= cap) return;
  bvh_walk<true>(k, t, node, leaf, sorig, [&](unsigned int c, int2 pr) {
    const unsigned int slot = base + c;
    if (slot < cap) pairs[slot] = pr;
  });
}

// Grid-stride loop over candidate pairs [0, np), one contiguous share per
// XCD: workgroups go to the eight XCDs in turn (the grid is a multiple of 8
// blocks), and the pairs are in Morton order of their first triangle, so each
// XCD works on one region of the mesh and its L2 keeps that region's corners.
struct PairRange {
  unsigned int k, end, stride;
};
__device__ __forceinline__ PairRange pair_range(unsigned int np) {
  constexpr unsigned int kX = 8;
  const unsigned int x = blockIdx.x % kX, q = blockIdx.x / kX, nbx = gridDim.x / kX;
  const unsigned int per = (np + kX - 1) / kX;
  const unsigned int b = min(np, x * per), e = min(np, b + per);
  return PairRange{b + q * blockDim.x + threadIdx.x, e, nbx * blockDim.x};
}

// The candidate pairs of this block's share (pair_range) that pass `keep`,
// written in order to out[0, n) (compacted with a wave ballot; no global
// counter); n, the same in every thread, is returned. out must hold the
// most pairs one block visits (ccd_region on the host). Launch with 256
// threads per block.
template <class Keep>
__device__ __forceinline__ unsigned int cull_to_region(const int2* __restrict__ pairs, unsigned int np, Keep keep_fn,
                                                       int2* __restrict__ out) {
  __shared__ unsigned int wave_kept[4];
  const PairRange rg = pair_range(np);
  const unsigned int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  unsigned int n = 0;
  for (unsigned int base = rg.k - threadIdx.x; base < rg.end; base += rg.stride) {
    const unsigned int k = base + threadIdx.x;
    int2 pr = make_int2(0, 0);
    bool keep = false;
    if (k < rg.end) {
      pr = pairs[k];
      keep = keep_fn(pr);
    }
    const unsigned long long m = __ballot(keep);
    if (lane == 0) wave_kept[wave] = (unsigned int)__popcll(m);
    __syncthreads();
    unsigned int at = n, total = 0;
#pragma unroll
    for (unsigned int w = 0; w < 4; ++w) {
      at += w < wave ? wave_kept[w] : 0u;
      total += wave_kept[w];
    }
    if (keep) out[at + (unsigned int)__popcll(m & ((1ull << lane) - 1ull))] = pr;
    n += total;
    __syncthreads();  // wave_kept is rewritten by the next round
  }
  return n;
}

// ---- IPC step state --------------------------------------------------------

// "No time of impact": 2.0, above every real step fraction; the minimum is
// taken with atomicMin on the bits (non-negative doubles order like their bits).
constexpr unsigned long long kNoToi = 0x4000000000000000ull;

struct IpcDev {
  double mu;                    // AL penalty
  double alpha;                 // step fraction of the current PD iteration
  double resid;                 // displacement residual of the current PD iteration
  unsigned long long toi;       // min TOI as double bits, kNoToi when none
  unsigned long long viol;      // max violation as double bits (non-negative)
  unsigned long long active;    // active contacts of the last detection
  unsigned long long self_contacts;
  int al_done, pd_done;
  int al_iters, iterations;
  int converged, overflow;  // overflow: the candidate pairs exceeded the buffer (sticky)
  double last_alpha;
};

// Per-detection inputs of the scene (collision_pipeline.rs:207-351).
struct Scene {
  int has_ground;
  double ground;
  int has_sphere;
  double cx, cy, cz, radius;
  int self_collision;
  double thickness;
  double dh, kappa;
};

// Per-triangle bounds for the pair cull of k_ccd_pairs: a float centre c, the
// largest distance r of a corner from c, a float reference displacement g
// (the corners' mean displacement) and the largest distance s of a corner's
// displacement from g, r and s rounded up. For a pair (A, B) every relative
// position in vt_ccd (a corner of one triangle minus a corner of the other)
// is then at most D = |c_A - c_B| + r_A + r_B long, and every relative motion
// (the difference of two corners' displacements) at most
// s_A + s_B + |g_A - g_B|: the cubic's leading coefficients depend on the
// relative motion only, which is small where the cloth moves as a piece.
struct CcdTri {
  float c[3], r, g[3], s;
};

__device__ __forceinline__ float round_up(double v) { return (float)(v * (1.0 + 1e-6)) + 1e-30f; }

// Bounding sphere of each triangle at x (CcdTri with g = 0, s = 0), for the
// pair cull of the self-contact tests.
__global__ void __launch_bounds__(256) k_tri_spheres(int t, const int* __restrict__ tri, const double* __restrict__ x,
                                                     CcdTri* __restrict__ out, const int* skip) {
  if (skip && *skip) return;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= t) return;
  const D3 a = ld(x, tri[3 * e]), b = ld(x, tri[3 * e + 1]), c = ld(x, tri[3 * e + 2]);
  CcdTri o;
  o.c[0] = (float)((a.x + b.x + c.x) / 3.0);
  o.c[1] = (float)((a.y + b.y + c.y) / 3.0);
  o.c[2] = (float)((a.z + b.z + c.z) / 3.0);
  const D3 cc = d3(o.c[0], o.c[1], o.c[2]);
  o.r = round_up(sqrt(fmax(len2(a - cc), fmax(len2(b - cc), len2(c - cc)))));
  o.g[0] = o.g[1] = o.g[2] = 0.0f;
  o.s = 0.0f;
  out[e] = o;
}

// self_collision.rs:199-276 with ipc_response.rs:58-219: the vertex-triangle
// (6) and edge-edge (9) tests of each candidate pair (self_pair_contacts),
// their barrier gradients summed per vertex in a fixed order (k_self_count
// .. k_vgather below).
//
// Pair cull (exact): no two points of A and B are closer than the gap of
// their bounding spheres (CcdTri centre and radius at x, k_tri_spheres). When
// that gap exceeds sqrt(dh) with a relative 1e-6 to spare, every distance the
// tests compute is beyond the barrier's reach, so the pair adds nothing and
// is skipped before its exclusion lookups.
//
// k_self_cull applies the cull at full occupancy and writes each block's
// survivors to its region of surv (as k_ccd_cull); the test kernels run on
// block b's survivors, every lane on a pair that needs them. Tested where
// they lay, most lanes of a wave sat idle (1002 us at 4.4 M candidate pairs);
// queued in LDS inside the test kernel, 512 us.
__global__ void __launch_bounds__(256) k_self_cull(const int2* __restrict__ pairs, const unsigned int* npairs_p,
                                                   unsigned int cap, Scene sc, const CcdTri* __restrict__ sph,
                                                   int2* __restrict__ surv, unsigned int reg,
                                                   unsigned int* __restrict__ scnt, const int* skip) {
  if (skip && *skip) return;
  const unsigned int np = *npairs_p < cap ? *npairs_p : cap;
  const double dh = sc.dh;
  const unsigned int n = cull_to_region(pairs, np, [&](const int2 pr) {
    const CcdTri SA = sph[pr.x], SB = sph[pr.y];
    const double gx = (double)SA.c[0] - (double)SB.c[0], gy = (double)SA.c[1] - (double)SB.c[1],
                 gz = (double)SA.c[2] - (double)SB.c[2];
    const double gap = sqrt(gx * gx + gy * gy + gz * gz) * (1.0 - 1e-12) - (double)SA.r - (double)SB.r;
    return !(gap > 0.0 && gap * gap > dh * (1.0 + 1e-6));
  }, surv + (size_t)blockIdx.x * reg);
  if (threadIdx.x == 0) scnt[blockIdx.x] = n;
}

// Squared gap between the axis-aligned boxes of two segments or of a point
// and a triangle (0 where they overlap): a lower bound of the squared
// distance between any two of their points.
__device__ __forceinline__ double box_gap2(D3 p0, D3 p1, D3 p2, D3 q0, D3 q1, D3 q2, D3 q3) {
  const double lx = fmin(fmin(p0.x, p1.x), p2.x), hx = fmax(fmax(p0.x, p1.x), p2.x);
  const double ly = fmin(fmin(p0.y, p1.y), p2.y), hy = fmax(fmax(p0.y, p1.y), p2.y);
  const double lz = fmin(fmin(p0.z, p1.z), p2.z), hz = fmax(fmax(p0.z, p1.z), p2.z);
  const double mx = fmin(fmin(q0.x, q1.x), fmin(q2.x, q3.x)), Mx = fmax(fmax(q0.x, q1.x), fmax(q2.x, q3.x));
  const double my = fmin(fmin(q0.y, q1.y), fmin(q2.y, q3.y)), My = fmax(fmax(q0.y, q1.y), fmax(q2.y, q3.y));
  const double mz = fmin(fmin(q0.z, q1.z), fmin(q2.z, q3.z)), Mz = fmax(fmax(q0.z, q1.z), fmax(q2.z, q3.z));
  const double gx = fmax(0.0, fmax(lx - Mx, mx - hx)), gy = fmax(0.0, fmax(ly - My, my - hy)),
               gz = fmax(0.0, fmax(lz - Mz, mz - hz));
  return gx * gx + gy * gy + gz * gz;
}

// The tests of one candidate pair that survived the cull, in the reference's
// order: on_contact(d2) for each active contact; with kWant, also
// emit(vertices[4], gradients[4], scaled barrier gradient), the
// contributions ipc_response.rs:170-219 adds to the four vertices.
template <bool kWant, class OnContact, class Emit>
__device__ __forceinline__ void self_pair_contacts(const int2 pr, const int* __restrict__ tri,
                                                   const double* __restrict__ x, const int* __restrict__ xptr,
                                                   const int* __restrict__ xidx, const Scene& sc, OnContact&& on_contact,
                                                   Emit&& emit) {
  const double dh = sc.dh;
  auto bgrad = [&](double d2) {
    return sc.thickness > 0.0 ? scaled_barrier_gradient_thick(d2, dh, sc.kappa, sc.thickness)
                              : scaled_barrier_gradient(d2, dh, sc.kappa);
  };
    const int A[3] = {tri[3 * pr.x], tri[3 * pr.x + 1], tri[3 * pr.x + 2]};
    const int B[3] = {tri[3 * pr.y], tri[3 * pr.y + 1], tri[3 * pr.y + 2]};
    // The six corners, loaded once (with the exclusion lists below) for all
    // fifteen tests.
    const D3 pA0 = ld(x, A[0]), pA1 = ld(x, A[1]), pA2 = ld(x, A[2]);
    const D3 pB0 = ld(x, B[0]), pB1 = ld(x, B[1]), pB2 = ld(x, B[2]);
    // Every exclusion test of the pair is between a vertex of A and one of
    // B, and the relation is symmetric (within `depth` rings): bit 3i + j
    // holds excluded(A[i], B[j]).
    // One pass over each of A's three exclusion lists tests it against all
    // of B's corners (the set of matches of nine separate searches).
    // The first kExPref entries of the three lists are loaded together; the
    // rest of a longer list follows one entry at a time.
    constexpr int kExPref = 8;
    int lo[3], hi[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) { lo[a] = xptr[A[a]]; hi[a] = xptr[A[a] + 1]; }
    int w[3][kExPref];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int j = 0; j < kExPref; ++j) w[a][j] = lo[a] + j < hi[a] ? xidx[lo[a] + j] : -1;
    unsigned int ex = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
      for (int j = 0; j < kExPref; ++j)
        ex |= ((w[a][j] == B[0]) ? 1u : 0u) << (3 * a) | ((w[a][j] == B[1]) ? 1u : 0u) << (3 * a + 1) |
              ((w[a][j] == B[2]) ? 1u : 0u) << (3 * a + 2);
      for (int q = lo[a] + kExPref; q < hi[a]; ++q) {
        const int v = xidx[q];
        ex |= ((v == B[0]) ? 1u : 0u) << (3 * a) | ((v == B[1]) ? 1u : 0u) << (3 * a + 1) |
              ((v == B[2]) ? 1u : 0u) << (3 * a + 2);
      }
    }
    if (ex == 0x1ffu) return;  // every test of the pair is excluded
    auto pick = [](const int* P, int q) { return q == 0 ? P[0] : (q == 1 ? P[1] : P[2]); };
    // corner q of a triangle, selected component by component (no indexed array)
    auto sel3 = [](int q, D3 u0, D3 u1, D3 u2) {
      return d3(q == 0 ? u0.x : (q == 1 ? u1.x : u2.x), q == 0 ? u0.y : (q == 1 ? u1.y : u2.y),
                q == 0 ? u0.z : (q == 1 ? u1.z : u2.z));
    };
    const D3 PA[3] = {pA0, pA1, pA2}, PB[3] = {pB0, pB1, pB2};  // indexed with constants only
    // Which of the fifteen tests run: bit u < 6 for vertex-triangle test u (A's
    // vertices against B, then B's against A), bit 6 + e for edge-edge test e.
    // A test is dropped when it shares a vertex, when the exclusion rule
    // removes it, or by the exact pre-test: boxes further apart than sqrt(dh)
    // (1e-6 relative to spare) mean d2 >= dh, no contact. The distances then
    // run only for the tests left, each lane over its own, so a wave pays for
    // its busiest lane rather than for all fifteen.
    const double reach = dh * (1.0 + 1e-6);
    unsigned int todo = 0;
#pragma unroll
    for (int u = 0; u < 6; ++u) {
      const bool fa = u < 3;
      const int q = fa ? u : u - 3;
      const int v = fa ? A[q] : B[q];
      const int* T = fa ? B : A;
      if (v == T[0] || v == T[1] || v == T[2]) continue;
      // excluded(v, T[*]): row q of ex for a vertex of A, column q for one of B
      const unsigned int row = fa ? (ex >> (3 * q)) & 7u : (ex >> q) & 0x49u;
      if (row) continue;
      const D3 p = fa ? PA[q] : PB[q];
      const D3* TP = fa ? PB : PA;
      if (box_gap2(p, p, p, TP[0], TP[1], TP[2], TP[2]) > reach) continue;
      todo |= 1u << u;
    }
#pragma unroll
    for (int e = 0; e < 9; ++e) {
      const int i = e / 3, j = e - 3 * i;
      const int i1 = i == 2 ? 0 : i + 1, j1 = j == 2 ? 0 : j + 1;
      if (A[i] == B[j] || A[i] == B[j1] || A[i1] == B[j] || A[i1] == B[j1]) continue;
      const unsigned int need = (1u << (3 * i + j)) | (1u << (3 * i + j1)) | (1u << (3 * i1 + j)) |
                                (1u << (3 * i1 + j1));
      if ((ex & need) == need) continue;
      if (box_gap2(PA[i], PA[i1], PA[i1], PB[j], PB[j1], PB[j1], PB[j1]) > reach) continue;
      todo |= 1u << (6 + e);
    }
    // vertex-triangle tests in order (A's vertices against B first)
    for (unsigned int vt = todo & 0x3fu; vt; vt &= vt - 1) {
      const int u = __builtin_ctz(vt);
      const bool fa = u < 3;
      const int q = fa ? u : u - 3;
      const int v = fa ? pick(A, q) : pick(B, q);
      const int t0 = fa ? B[0] : A[0], t1 = fa ? B[1] : A[1], t2 = fa ? B[2] : A[2];
      const D3 p = fa ? sel3(q, pA0, pA1, pA2) : sel3(q, pB0, pB1, pB2);
      const D3 a = fa ? pB0 : pA0, b = fa ? pB1 : pA1, c = fa ? pB2 : pA2;
      const double d2 = pt_dist2(p, a, b, c);
      if (d2 < dh && d2 > 0.0) {
        if constexpr (kWant) {
          D3 g[4];
          pt_grad(p, a, b, c, g);
          const int vs[4] = {v, t0, t1, t2};
          emit(vs, g, bgrad(d2));
        }
        on_contact(d2);
      }
    }
    // edge-edge tests in order: edge i of A (A[i], A[i+1]) against edge j of B
    for (unsigned int ee = todo >> 6; ee; ee &= ee - 1) {
      const int e = __builtin_ctz(ee);
      const int i = e / 3, j = e - 3 * i;
      const int i1 = i == 2 ? 0 : i + 1, j1 = j == 2 ? 0 : j + 1;
      const int a0 = pick(A, i), a1 = pick(A, i1), b0 = pick(B, j), b1 = pick(B, j1);
      const D3 pa0 = sel3(i, pA0, pA1, pA2), pa1 = sel3(i1, pA0, pA1, pA2), pb0 = sel3(j, pB0, pB1, pB2),
               pb1 = sel3(j1, pB0, pB1, pB2);
      const double d2 = ee_dist2(pa0, pa1, pb0, pb1);
      if (d2 < dh && d2 > 0.0) {
        if constexpr (kWant) {
          D3 g[4];
          ee_grad(pa0, pa1, pb0, pb1, g);
          const int vs[4] = {a0, a1, b0, b1};
          emit(vs, g, bgrad(d2));
        }
        on_contact(d2);
      }
    }
}

// The self-contact gradients, summed in a fixed order (no floating-point
// atomics, so a detection gives the same bits every time):
//   k_self_count  the contacts of each surviving pair, their offsets within
//                 block b's survivors (cof) and the block's total (btot);
//                 the contact counters and the violation flag
//   k_self_base   the blocks' offsets and the total (BvhDev::ncont)
//   k_self_emit   each contact's four (vertex, gradient) entries at its
//                 offset, in survivor and test order
//   k_vcount / k_vscatter  the entries grouped by vertex (integer counts)
//   k_vgather     each vertex's sum over its entries in ascending entry order
// Entries past the buffer (ecap contacts) are dropped and flag an overflow,
// as a pair overflow does; the host grows the buffer and repeats.
//
// The tests load the pair's six corners once and evaluate only the
// distances their exact pre-tests leave (self_pair_contacts); at the
// 512 x 512 drape that took k_self_count from 426 to 297 us, and four waves
// per SIMD (128 VGPRs, 12 bytes of spill) instead of the compiler's three to
// 262 us. k_self_emit runs the tests only for survivors with contacts.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_self_count(const unsigned int* npairs_p, unsigned int cap,
                                                    const int2* __restrict__ surv, unsigned int reg,
                                                    const unsigned int* __restrict__ scnt, const int* __restrict__ tri,
                                                    const double* __restrict__ x, const int* __restrict__ xptr,
                                                    const int* __restrict__ xidx, Scene sc,
                                                    unsigned int* __restrict__ cof, unsigned int* __restrict__ btot,
                                                    IpcDev* ip, const int* skip) {
  if (skip && *skip) return;
  if (blockIdx.x == 0 && threadIdx.x == 0 && *npairs_p > cap) ip->overflow = 1;
  __shared__ unsigned int wsum[4];
  const unsigned int n = scnt[blockIdx.x];
  const int2* in = surv + (size_t)blockIdx.x * reg;
  unsigned int* out = cof + (size_t)blockIdx.x * reg;
  const unsigned int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  unsigned int carry = 0;
  unsigned long long total = 0;
  bool viol = false;
  for (unsigned int base = 0; base < n; base += blockDim.x) {
    const unsigned int k = base + threadIdx.x;
    unsigned int c = 0;
    if (k < n)
      self_pair_contacts<false>(in[k], tri, x, xptr, xidx, sc, [&](double d2) {
        ++c;
        if (d2 <= 1e-12) viol = true;
      }, [](const int*, const D3*, double) {});
    // exclusive scan of c over the block, in survivor order
    unsigned int incl = c;
    for (int o = 1; o < 64; o <<= 1) {
      const unsigned int v = __shfl_up(incl, o);
      if ((int)lane >= o) incl += v;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    unsigned int before = 0, all = 0;
#pragma unroll
    for (unsigned int w = 0; w < 4; ++w) {
      before += w < wave ? wsum[w] : 0u;
      all += wsum[w];
    }
    if (k < n) out[k] = carry + before + incl - c;
    carry += all;
    total += c;
    __syncthreads();  // wsum is rewritten by the next round
  }
  if (threadIdx.x == 0) btot[blockIdx.x] = carry;
  for (int o = 32; o > 0; o >>= 1) total += __shfl_xor(total, o);
  const bool any_viol = __ballot(viol) != 0;
  if (lane == 0) {
    if (total) {
      atomicAdd(&ip->self_contacts, total);
      atomicAdd(&ip->active, total);
    }
    if (any_viol) atomicMax(&ip->viol, (unsigned long long)__double_as_longlong(1e-3));
  }
}

// One block: exclusive scan of the nb block totals into base, the total into
// bd->ncont; a total past ecap contacts flags an overflow.
__global__ void __launch_bounds__(256) k_self_base(int nb, const unsigned int* __restrict__ btot,
                                                   unsigned int* __restrict__ base, unsigned int ecap, BvhDev* bd,
                                                   IpcDev* ip, const int* skip) {
  if (skip && *skip) return;
  __shared__ unsigned int wsum[4];
  const unsigned int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  unsigned int carry = 0;
  for (int b0 = 0; b0 < nb; b0 += blockDim.x) {
    const int b = b0 + threadIdx.x;
    const unsigned int c = b < nb ? btot[b] : 0u;
    unsigned int incl = c;
    for (int o = 1; o < 64; o <<= 1) {
      const unsigned int v = __shfl_up(incl, o);
      if ((int)lane >= o) incl += v;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    unsigned int before = 0, all = 0;
#pragma unroll
    for (unsigned int w = 0; w < 4; ++w) {
      before += w < wave ? wsum[w] : 0u;
      all += wsum[w];
    }
    if (b < nb) base[b] = carry + before + incl - c;
    carry += all;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    bd->ncont = carry;
    if (carry > ecap) ip->overflow = 1;
  }
}

__global__ void __launch_bounds__(256) k_self_emit(const int2* __restrict__ surv, unsigned int reg, const unsigned int* __restrict__ scnt,
            const unsigned int* __restrict__ cof, const unsigned int* __restrict__ base, const int* __restrict__ tri,
            const double* __restrict__ x, const int* __restrict__ xptr, const int* __restrict__ xidx, Scene sc,
            unsigned int ecap, int* __restrict__ ev, double* __restrict__ eg, const int* skip) {
  if (skip && *skip) return;
  const unsigned int n = scnt[blockIdx.x];
  const int2* in = surv + (size_t)blockIdx.x * reg;
  const unsigned int* off = cof + (size_t)blockIdx.x * reg;
  const unsigned int b0 = base[blockIdx.x];
  const unsigned int bend = blockIdx.x + 1 < gridDim.x ? base[blockIdx.x + 1] : 0xffffffffu;
  for (unsigned int k = threadIdx.x; k < n; k += blockDim.x) {
    // the pair's contact count is the gap to the next offset: most survivors
    // have none, and their lanes skip the tests
    const unsigned int next = k + 1 < n ? b0 + off[k + 1] : bend;
    unsigned int c = b0 + off[k];
    if (next == c) continue;
    self_pair_contacts<true>(in[k], tri, x, xptr, xidx, sc, [](double) {}, [&](const int* vs, const D3* g, double bg) {
      if (c < ecap) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const size_t e = 4 * (size_t)c + q;
          ev[e] = vs[q];
          eg[3 * e + 0] = bg * g[q].x;
          eg[3 * e + 1] = bg * g[q].y;
          eg[3 * e + 2] = bg * g[q].z;
        }
      }
      ++c;
    });
  }
}

__global__ void __launch_bounds__(256) k_fill_u(int count, unsigned int* __restrict__ p, const int* skip) {
  if (skip && *skip) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0u;
}

// The entries of each vertex: counts (vcnt[n] stays 0, so the exclusive scan
// of n + 1 counts ends with the number of entries).
__global__ void __launch_bounds__(256) k_vcount(const BvhDev* bd, unsigned int ecap, const int* __restrict__ ev,
                                                unsigned int* __restrict__ vcnt, const int* skip) {
  if (skip && *skip) return;
  const unsigned int ne = 4u * (bd->ncont < ecap ? bd->ncont : ecap);
  for (unsigned int e = blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += gridDim.x * blockDim.x)
    atomicAdd(&vcnt[ev[e]], 1u);
}

// Entry e into its vertex's segment [vptr[v], vptr[v + 1]), in any order.
__global__ void __launch_bounds__(256) k_vscatter(const BvhDev* bd, unsigned int ecap, const int* __restrict__ ev,
                                                  const unsigned int* __restrict__ vptr, unsigned int* __restrict__ vfill,
                                                  unsigned int* __restrict__ vidx, const int* skip) {
  if (skip && *skip) return;
  const unsigned int ne = 4u * (bd->ncont < ecap ? bd->ncont : ecap);
  for (unsigned int e = blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += gridDim.x * blockDim.x) {
    const int v = ev[e];
    vidx[vptr[v] + atomicAdd(&vfill[v], 1u)] = e;
  }
}

// g[v] = the sum of vertex v's entries in ascending entry order (the
// segment's entries are picked smallest first), 0 without entries.
__global__ void __launch_bounds__(256) k_vgather(int n, const unsigned int* __restrict__ vptr,
                                                 const unsigned int* __restrict__ vidx, const double* __restrict__ eg,
                                                 double* __restrict__ g, const int* skip) {
  if (skip && *skip) return;
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const unsigned int b = vptr[v], e = vptr[v + 1];
  double sx = 0.0, sy = 0.0, sz = 0.0;
  unsigned int last = 0;
  for (unsigned int done = 0; done < e - b; ++done) {
    unsigned int pick = 0xffffffffu;
    for (unsigned int q = b; q < e; ++q) {
      const unsigned int id = vidx[q];
      if ((done == 0 || id > last) && id < pick) pick = id;
    }
    sx += eg[3 * (size_t)pick + 0];
    sy += eg[3 * (size_t)pick + 1];
    sz += eg[3 * (size_t)pick + 2];
    last = pick;
  }
  g[3 * v + 0] = sx; g[3 * v + 1] = sy; g[3 * v + 2] = sz;
}

// Per-vertex half of collision_pipeline.rs:207-351: self-contact normal from
// the self gradient, ground and sphere barriers (ground_plane.rs:68-110,
// sphere.rs:102-176), normal normalisation; then the solver's bookkeeping of
// pd_solver.rs:640-664 / 966-978 (ever-in-contact, last normals) and, at the
// start of an AL iteration (lambda_reset), the multiplier reset of
// pd_solver.rs:656-663.
__global__ void __launch_bounds__(256) k_vertex_contacts(int n, const double* __restrict__ x, Scene sc,
                                                         const double* __restrict__ gself, double* __restrict__ grad,
                                                         double* __restrict__ nrm, uint8_t* __restrict__ inc,
                                                         uint8_t* __restrict__ ever, double* __restrict__ fnrm,
                                                         double* __restrict__ lambda, int lambda_reset, IpcDev* ip,
                                                         double* __restrict__ hess, const int* skip) {
  if (skip && *skip) return;
  // A detection whose candidate pairs overflowed the buffer writes nothing:
  // the host grows the buffer and runs the detection again.
  if (ip->overflow) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double gx = 0, gy = 0, gz = 0, nx = 0, ny = 0, nz = 0, hd = 0;
  bool in = false;
  unsigned long long act = 0;
  double viol = 0;
  if (sc.self_collision) {
    const double sx = gself[3 * i], sy = gself[3 * i + 1], sz = gself[3 * i + 2];
    gx += sx; gy += sy; gz += sz;
    const double g2 = sx * sx + sy * sy + sz * sz;
    if (g2 > 1e-12) {
      const double gl = sqrt(g2);
      nx += sx / gl; ny += sy / gl; nz += sz / gl;
      in = true;
    }
  }
  const double px = x[3 * i], py = x[3 * i + 1], pz = x[3 * i + 2];
  const double dh = sc.dh, kappa = sc.kappa;
  if (sc.has_ground) {
    const double ds = py - sc.ground;
    if (ds <= 0.0) {
      viol = fmax(viol, -ds);
      const double dcl = fmax(-ds, 1e-6);
      ++act;
      gy += scaled_barrier_gradient(dcl * dcl, dh, kappa) * 2.0 * dcl;
    } else if (ds * ds < dh) {
      ++act;
      gy += scaled_barrier_gradient(ds * ds, dh, kappa) * 2.0 * ds;
    }
    if (ds * ds < dh || ds <= 0.0) {
      ny += 1.0;
      in = true;
      const double de = fmax(ds, 1e-6), d2 = de * de;  // collision_pipeline.rs:272-276
      hd += fabs(kappa * barrier_hessian(d2, dh) * 4.0 * d2);
    }
  }
  if (sc.has_sphere) {
    const double dx = px - sc.cx, dy = py - sc.cy, dz = pz - sc.cz;
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    const double ds = r - sc.radius;
    if (ds <= 0.0) {
      viol = fmax(viol, -ds);
      const double dcl = fmax(-ds, 1e-6);
      ++act;
      const double bg = scaled_barrier_gradient(dcl * dcl, dh, kappa);
      if (r > 1e-6) {
        const double f = bg * 2.0 * dcl;
        gx += f * (dx / r); gy += f * (dy / r); gz += f * (dz / r);
      } else {
        gy += bg * 2.0 * dcl;
      }
    } else if (ds * ds < dh) {
      ++act;
      const double f = scaled_barrier_gradient(ds * ds, dh, kappa) * 2.0 * ds;
      gx += f * (dx / r); gy += f * (dy / r); gz += f * (dz / r);
    }
    if (ds * ds < dh || ds <= 0.0) {
      if (r > 1e-6) { nx += dx / r; ny += dy / r; nz += dz / r; }
      in = true;
      const double de = fmax(ds, 1e-6), d2 = de * de;  // collision_pipeline.rs:301-305
      hd += fabs(kappa * barrier_hessian(d2, dh) * 4.0 * d2);
    }
  }
  if (hess) hess[i] = hd;
  if (in) {
    const double l2 = nx * nx + ny * ny + nz * nz;
    if (l2 > 1e-12) {
      const double inv = 1.0 / sqrt(l2);
      nx *= inv; ny *= inv; nz *= inv;
    }
  }
  grad[3 * i] = gx; grad[3 * i + 1] = gy; grad[3 * i + 2] = gz;
  nrm[3 * i] = nx; nrm[3 * i + 1] = ny; nrm[3 * i + 2] = nz;
  inc[i] = in ? 1 : 0;
  if (in) {
    ever[i] = 1;
    fnrm[3 * i] = nx; fnrm[3 * i + 1] = ny; fnrm[3 * i + 2] = nz;
  } else if (lambda_reset) {
    lambda[3 * i] = 0; lambda[3 * i + 1] = 0; lambda[3 * i + 2] = 0;
  }
  if (act) atomicAdd(&ip->active, act);
  if (viol > 0) atomicMax(&ip->viol, (unsigned long long)__double_as_longlong(viol));
}

// ---- per-step kernels of step_with_ipc --------------------------------------

// Start of the step (pd_solver.rs:566-609): save_previous, the CFL-like
// velocity clamp |v| <= 0.5 sqrt(d_hat) / dt for free vertices, predict
// (state.rs:145-169, ground clamp when the scene has a ground), x = pred.
// Also resets the step's counters.
__global__ void __launch_bounds__(256) k_ipc_begin(int n, double* __restrict__ x, double* __restrict__ v,
                                                   double* __restrict__ prev, double* __restrict__ pred,
                                                   const uint8_t* __restrict__ pinned, uint8_t* __restrict__ ever,
                                                   double dt, double gx, double gy, double gz, double max_vel,
                                                   int has_ground, double ground, double mu0, IpcDev* ip,
                                                   vhk::StepDev* st) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    ip->mu = mu0; ip->alpha = 1.0; ip->resid = DBL_MAX; ip->al_done = 0; ip->pd_done = 0;
    ip->al_iters = 1; ip->iterations = 0; ip->converged = 0; ip->last_alpha = 1.0; ip->toi = kNoToi;
    ip->viol = 0; ip->active = 0; ip->self_contacts = 0;
    st->done = 0; st->iterations = 0; st->residual = DBL_MAX;
  }
  if (i >= n) return;
  ever[i] = 0;
  double px = x[3 * i], py = x[3 * i + 1], pz = x[3 * i + 2];
  prev[3 * i] = px; prev[3 * i + 1] = py; prev[3 * i + 2] = pz;
  if (pinned[i]) {
    pred[3 * i] = px; pred[3 * i + 1] = py; pred[3 * i + 2] = pz;
    return;
  }
  double vx = v[3 * i], vy = v[3 * i + 1], vz = v[3 * i + 2];
  const double v2 = vx * vx + vy * vy + vz * vz;
  if (v2 > max_vel * max_vel) {
    const double s = max_vel / sqrt(v2);
    vx *= s; vy *= s; vz *= s;
    v[3 * i] = vx; v[3 * i + 1] = vy; v[3 * i + 2] = vz;
  }
  const double dt2 = dt * dt;
  px += dt * vx + dt2 * gx;
  py += dt * vy + dt2 * gy;
  pz += dt * vz + dt2 * gz;
  if (has_ground && py < ground) py = ground;
  pred[3 * i] = px; pred[3 * i + 1] = py; pred[3 * i + 2] = pz;
  x[3 * i] = px; x[3 * i + 1] = py; x[3 * i + 2] = pz;
}

// Counters of the detection that follows.
__global__ void k_ipc_detect_begin(IpcDev* ip) {
  if (threadIdx.x != 0 || ip->al_done) return;
  ip->viol = 0; ip->active = 0; ip->self_contacts = 0;
}

// Zero fill (instead of memset, so the sequence stays capturable).
__global__ void __launch_bounds__(256) k_zero_d(int count, double* __restrict__ p, const int* skip) {
  if (skip && *skip) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0;
}

// assemble_barrier_rhs (assembly.rs:272-282): the per-vertex RHS term
// ext = -(mu g + lambda), frozen for the AL iteration.
__global__ void __launch_bounds__(256) k_ipc_ext(int n, const double* __restrict__ grad,
                                                 const double* __restrict__ lambda, double* __restrict__ ext,
                                                 const IpcDev* ip) {
  if (ip->al_done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * n) return;
  ext[i] = -(ip->mu * grad[i] + lambda[i]);
}

// Per-block partial sums of one PD iteration (pd_solver.rs:873-891, 911-922):
// |delta|^2, |x|^2 and, over free vertices, g . delta.
__global__ void __launch_bounds__(256) k_ipc_sums(int n, const double* __restrict__ x, const double* __restrict__ dl,
                                                  const double* __restrict__ grad, const uint8_t* __restrict__ pinned,
                                                  double* __restrict__ partials, const IpcDev* ip) {
  if (ip->al_done || ip->pd_done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double a = 0, b = 0, c = 0;
  if (i < n) {
    const double dx = dl[3 * i], dy = dl[3 * i + 1], dz = dl[3 * i + 2];
    const double ox = x[3 * i], oy = x[3 * i + 1], oz = x[3 * i + 2];
    a = dx * dx + dy * dy + dz * dz;
    b = ox * ox + oy * oy + oz * oz;
    if (!pinned[i]) c = grad[3 * i] * dx + grad[3 * i + 1] * dy + grad[3 * i + 2] * dz;
  }
  __shared__ double sa[256], sb[256], sc2[256];
  sa[threadIdx.x] = a; sb[threadIdx.x] = b; sc2[threadIdx.x] = c;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sa[threadIdx.x] += sa[threadIdx.x + s]; sb[threadIdx.x] += sb[threadIdx.x + s];
      sc2[threadIdx.x] += sc2[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partials[3 * blockIdx.x] = sa[0]; partials[3 * blockIdx.x + 1] = sb[0]; partials[3 * blockIdx.x + 2] = sc2[0];
  }
}

__device__ __forceinline__ void toi_min(IpcDev* ip, double t) {
  atomicMin(&ip->toi, (unsigned long long)__double_as_longlong(t));
}


__global__ void __launch_bounds__(256) k_ccd_tri_bounds(int t, const int* __restrict__ tri,
                                                        const double* __restrict__ x, const double* __restrict__ dl,
                                                        CcdTri* __restrict__ out, const IpcDev* ip) {
  if (ip->al_done || ip->pd_done) return;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= t) return;
  const D3 a = ld(x, tri[3 * e]), b = ld(x, tri[3 * e + 1]), c = ld(x, tri[3 * e + 2]);
  const float cx = (float)((a.x + b.x + c.x) / 3.0), cy = (float)((a.y + b.y + c.y) / 3.0),
              cz = (float)((a.z + b.z + c.z) / 3.0);
  const D3 cc = d3(cx, cy, cz);
  const double r2 = fmax(len2(a - cc), fmax(len2(b - cc), len2(c - cc)));
  const D3 da = ld(dl, tri[3 * e]), db = ld(dl, tri[3 * e + 1]), dc = ld(dl, tri[3 * e + 2]);
  const float gx = (float)((da.x + db.x + dc.x) / 3.0), gy = (float)((da.y + db.y + dc.y) / 3.0),
              gz = (float)((da.z + db.z + dc.z) / 3.0);
  const D3 g = d3(gx, gy, gz);
  const double s2 = fmax(len2(da - g), fmax(len2(db - g), len2(dc - g)));
  CcdTri o;
  o.c[0] = cx; o.c[1] = cy; o.c[2] = cz;
  o.r = round_up(sqrt(r2));
  o.g[0] = gx; o.g[1] = gy; o.g[2] = gz;
  o.s = round_up(sqrt(s2));
  out[e] = o;
}

// Self part of compute_ccd_step (ccd.rs:238-292) over the candidate pairs of
// the last detection, from x to x + delta; with EE, also edge_edge_ccd
// (ccd.rs:138-229) on the nine edge pairs of each triangle pair.
//
// Pair cull (exact): the cubic of a vertex-triangle test has coefficients
// |ca| <= W^3, |cb| <= 3 D W^2, |cc| <= 3 D^2 W, with D and the relative
// motion bound W = s_A + s_B + |g_A - g_B| from CcdTri. An edge-edge cubic
// has the same bounds with W = max(2 s_A, 2 s_B, s_A + s_B + |g_A - g_B|),
// since it also holds the motion of one edge's end relative to its other end
// (|db - da| <= 2 s_A); the cull then uses that W, which also bounds the
// vertex-triangle tests. When all three are below half of the solver's 1e-8
// threshold, cubic_roots falls through quadratic_roots to the linear case and
// finds no root for any test of the pair (the computed coefficients differ
// from the bounded ones by a few ulps, and the relative motions the tests
// form from x and x + delta by a few ulps of the coordinates), so the pair
// cannot report a hit and is skipped.
//
// k_ccd_cull applies the cull; block b writes the survivors of its share of
// the pairs (pair_range) to its own region surv[b reg, (b + 1) reg) and their
// count to scnt[b]. reg bounds the pairs one block visits (ccd_region on the
// host, from the same grid and capacity). k_ccd_pairs then runs the tests on
// the survivors only, block b on region b. The cull runs at full occupancy;
// the 205-VGPR tests run at two waves per SIMD with every lane on a pair
// that needs them. (Testing the survivors where they lie left most lanes of
// a wave idle; queueing them in LDS inside the test kernel cost more in
// barriers at two waves per SIMD: 106.6 -> 137.4 us at 4.4 M pairs.)
__global__ void __launch_bounds__(256) k_ccd_cull(const int2* __restrict__ pairs, const unsigned int* npairs_p,
                                                  unsigned int cap, const CcdTri* __restrict__ tb_,
                                                  int2* __restrict__ surv, unsigned int reg,
                                                  unsigned int* __restrict__ scnt, const IpcDev* ip, int ee) {
  if (ip->al_done || ip->pd_done) return;
  const unsigned int np = *npairs_p < cap ? *npairs_p : cap;
  const unsigned int n = cull_to_region(pairs, np, [&](const int2 pr) {
    const CcdTri A = tb_[pr.x], B = tb_[pr.y];
    const double gx = (double)A.c[0] - (double)B.c[0], gy = (double)A.c[1] - (double)B.c[1],
                 gz = (double)A.c[2] - (double)B.c[2];
    const double D = sqrt(gx * gx + gy * gy + gz * gz) * (1.0 + 1e-12) + (double)A.r + (double)B.r;
    const double hx = (double)A.g[0] - (double)B.g[0], hy = (double)A.g[1] - (double)B.g[1],
                 hz = (double)A.g[2] - (double)B.g[2];
    const double Wv = sqrt(hx * hx + hy * hy + hz * hz) * (1.0 + 1e-12) + (double)A.s + (double)B.s;
    const double W = ee ? fmax(Wv, 2.0 * fmax((double)A.s, (double)B.s)) : Wv;
    constexpr double kHalfEps = 0.5e-8;
    return !(W * W * W < kHalfEps && 3.0 * D * W * W < kHalfEps && 3.0 * D * D * W < kHalfEps);
  }, surv + (size_t)blockIdx.x * reg);
  if (threadIdx.x == 0) scnt[blockIdx.x] = n;
}

template <bool EE>
__global__ void __launch_bounds__(256) k_ccd_pairs(const int2* __restrict__ surv, unsigned int reg,
                                                   const unsigned int* __restrict__ scnt, const int* __restrict__ tri,
                                                   const double* __restrict__ x, const double* __restrict__ dl,
                                                   IpcDev* ip) {
  if (ip->al_done || ip->pd_done) return;
  const unsigned int n = scnt[blockIdx.x];
  const int2* in = surv + (size_t)blockIdx.x * reg;
  double best = 2.0;
  for (unsigned int k = threadIdx.x; k < n; k += blockDim.x) {
    const int2 pr = in[k];
    // The six corners (A then B) with their start and end positions, loaded
    // once per pair: measured faster than reloading them per test, despite
    // the register cost of the unrolled tests.
    const int V[6] = {tri[3 * pr.x], tri[3 * pr.x + 1], tri[3 * pr.x + 2],
                      tri[3 * pr.y], tri[3 * pr.y + 1], tri[3 * pr.y + 2]};
    D3 P0[6], P1[6];
#pragma unroll
    for (int u = 0; u < 6; ++u) {
      P0[u] = ld(x, V[u]);
      P1[u] = P0[u] + ld(dl, V[u]);
    }
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      const int tb = side == 0 ? 3 : 0;  // the triangle's first corner in V
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const int vi = 3 * side + q;
        if (V[vi] == V[tb] || V[vi] == V[tb + 1] || V[vi] == V[tb + 2]) continue;
        best = fmin(best, vt_ccd(P0[vi], P1[vi], P0[tb], P1[tb], P0[tb + 1], P1[tb + 1], P0[tb + 2], P1[tb + 2]));
      }
    }
    if constexpr (EE) {
      // edges (0,1), (1,2), (2,0) of A against those of B, skipping pairs
      // that share a vertex (unrolled: the corner arrays stay in registers)
#pragma unroll
      for (int ea = 0; ea < 3; ++ea)
#pragma unroll
        for (int eb = 0; eb < 3; ++eb) {
          const int a0 = ea, a1 = ea == 2 ? 0 : ea + 1, b0 = 3 + eb, b1 = eb == 2 ? 3 : 4 + eb;
          if (V[a0] == V[b0] || V[a0] == V[b1] || V[a1] == V[b0] || V[a1] == V[b1]) continue;
          best = fmin(best, ee_ccd(P0[a0], P1[a0], P0[a1], P1[a1], P0[b0], P1[b0], P0[b1], P1[b1]));
        }
    }
  }
  for (int o = 32; o > 0; o >>= 1) best = fmin(best, __shfl_xor(best, o));
  if ((threadIdx.x & 63) == 0 && best <= 1.0) toi_min(ip, fmax(best, 1e-6));
}

// Ground and sphere parts of compute_ccd_step (ground_plane.rs:112-134,
// sphere.rs:178-224), from x to x + delta.
__global__ void __launch_bounds__(256) k_ccd_vertices(int n, const double* __restrict__ x,
                                                      const double* __restrict__ dl, Scene sc, IpcDev* ip) {
  if (ip->al_done || ip->pd_done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double best = 2.0;
  const double x0 = x[3 * i], y0 = x[3 * i + 1], z0 = x[3 * i + 2];
  const double x1 = x0 + dl[3 * i], y1 = y0 + dl[3 * i + 1], z1 = z0 + dl[3 * i + 2];
  if (sc.has_ground && y0 > sc.ground && y1 < sc.ground) {
    const double t = (sc.ground - y0) / (y1 - y0);
    if (t >= 0.0 && t <= 1.0) best = fmin(best, fmax(t * 0.9, 1e-6));
  }
  if (sc.has_sphere) {
    const double r2 = sc.radius * sc.radius;
    const double ax = x0 - sc.cx, ay = y0 - sc.cy, az = z0 - sc.cz;
    if (!(ax * ax + ay * ay + az * az < r2)) {
      const double bx = x1 - sc.cx, by = y1 - sc.cy, bz = z1 - sc.cz;
      const double vx = bx - ax, vy = by - ay, vz = bz - az;
      const double a = vx * vx + vy * vy + vz * vz;
      const double b = 2.0 * (ax * vx + ay * vy + az * vz);
      const double c = ax * ax + ay * ay + az * az - r2;
      if (a > 1e-8) {
        const double disc = b * b - 4.0 * a * c;
        if (disc >= 0.0) {
          const double t = (-b - sqrt(disc)) / (2.0 * a);
          if (t >= 0.0 && t <= 1.0) best = fmin(best, fmax(t * 0.9, 1e-6));
        }
      }
    }
  }
  if (best <= 1.0) toi_min(ip, best);
}

// The step fraction of one PD iteration (pd_solver.rs:873-930): residual,
// CCD limit with the 0.8 safety factor, and the barrier backtracking.
__global__ void __launch_bounds__(256) k_ipc_alpha(const double* __restrict__ partials, int blocks, double tol,
                                                   IpcDev* ip) {
  if (ip->al_done || ip->pd_done) return;
  __shared__ double sa[256], sb[256], sc2[256];
  double a = 0, b = 0, c = 0;
  for (int k = threadIdx.x; k < blocks; k += blockDim.x) {
    a += partials[3 * k]; b += partials[3 * k + 1]; c += partials[3 * k + 2];
  }
  sa[threadIdx.x] = a; sb[threadIdx.x] = b; sc2[threadIdx.x] = c;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sa[threadIdx.x] += sa[threadIdx.x + s]; sb[threadIdx.x] += sb[threadIdx.x + s];
      sc2[threadIdx.x] += sc2[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  const double diff = sa[0], norm = sb[0], gdd = sc2[0];
  const double res = norm > 1e-12 ? sqrt(diff / norm) : sqrt(diff);
  double alpha = fmin(1.0, __longlong_as_double((long long)ip->toi));
  alpha = fmax(alpha, 1e-6);
  if (alpha < 1.0) alpha *= 0.8;
  alpha = fmin(alpha, 1.0);
  for (int bt = 0; bt < 4; ++bt) {
    if (alpha < 0.1) break;
    if (gdd * ip->mu > 1e-4 * diff * 100.0) alpha *= 0.5;
    else break;
  }
  ip->alpha = alpha;
  ip->last_alpha = alpha;
  ip->resid = res;
  ip->iterations += 1;
  ip->toi = kNoToi;
  if (res < tol) ip->pd_done = 2;  // set after this iteration's update
}

// x += alpha delta for free vertices (pd_solver.rs:934-950).
__global__ void __launch_bounds__(256) k_ipc_update(int n, double* __restrict__ x, const double* __restrict__ dl,
                                                    const uint8_t* __restrict__ pinned, IpcDev* ip) {
  if (ip->al_done || ip->pd_done == 1) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || pinned[i]) return;
  const double a = ip->alpha;
  x[3 * i] += a * dl[3 * i];
  x[3 * i + 1] += a * dl[3 * i + 1];
  x[3 * i + 2] += a * dl[3 * i + 2];
}

// Turns a converged PD loop (pd_done = 2 after its last update) into a stop
// of the remaining PD iterations of this AL iteration.
__global__ void k_ipc_pd_latch(IpcDev* ip, vhk::StepDev* st) {
  if (threadIdx.x == 0 && ip->pd_done == 2) { ip->pd_done = 1; st->done = 1; }
}

// End of an AL iteration (pd_solver.rs:996-1021), after the detection at
// the new positions. Unless the violation is under al_tolerance: lambda +=
// mu g on contact vertices and, when another AL iteration follows
// (reset), lambda = 0 off contact (the reset at the start of that
// iteration, pd_solver.rs:656-663, which reuses this detection).
__global__ void __launch_bounds__(256) k_ipc_al_end(int n, const double* __restrict__ grad,
                                                    const uint8_t* __restrict__ inc, double* __restrict__ lambda,
                                                    double al_tol, int reset, const IpcDev* ip) {
  if (ip->al_done || ip->overflow) return;
  if (__longlong_as_double((long long)ip->viol) < al_tol) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (inc[i]) {
    const double mu = ip->mu;
    lambda[3 * i] += mu * grad[3 * i];
    lambda[3 * i + 1] += mu * grad[3 * i + 1];
    lambda[3 * i + 2] += mu * grad[3 * i + 2];
  } else if (reset) {
    lambda[3 * i] = 0; lambda[3 * i + 1] = 0; lambda[3 * i + 2] = 0;
  }
}

// The AL decision (one thread, after k_ipc_al_end): converged -> stop;
// else mu *= growth, and the next AL iteration starts unless this was the
// last one. The PD-loop flags are re-armed for that iteration. Like
// k_ipc_al_end it does nothing after a pair overflow: the host then grows
// the buffer and runs the detection, k_ipc_al_end and this again.
__global__ void k_ipc_al_commit(double al_tol, double growth, int last, IpcDev* ip, vhk::StepDev* st) {
  if (threadIdx.x != 0 || ip->al_done || ip->overflow) return;
  if (__longlong_as_double((long long)ip->viol) < al_tol) {
    ip->al_done = 1; ip->converged = 1; st->done = 1;
    return;
  }
  ip->mu *= growth;
  if (last) { ip->al_done = 1; st->done = 1; return; }
  ip->al_iters += 1;
  ip->pd_done = 0;
  st->done = 0;
}

// End of the step (pd_solver.rs:1026-1135): velocities, the inelastic contact
// filter with Coulomb friction and (adaptive) contact damping for vertices
// that were ever in contact, the 5 % global damping when the last detection
// has contacts, ground velocities, damping and Rayleigh damping.
struct FinalArgs {
  double inv_dt, mu_f, cdamp, cdamp_max, v_thresh, one_minus_damp, frayl, ground;
  int adaptive, has_ground;
};
__global__ void __launch_bounds__(256) k_ipc_finalize(int n, const double* __restrict__ x, double* __restrict__ prev,
                                                      double* __restrict__ v, const uint8_t* __restrict__ pinned,
                                                      const uint8_t* __restrict__ ever,
                                                      const double* __restrict__ fnrm, FinalArgs f,
                                                      const IpcDev* ip) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double vx = (x[3 * i] - prev[3 * i]) * f.inv_dt;
  double vy = (x[3 * i + 1] - prev[3 * i + 1]) * f.inv_dt;
  double vz = (x[3 * i + 2] - prev[3 * i + 2]) * f.inv_dt;
  const bool free_v = !pinned[i];
  if (free_v && ever[i]) {
    const double nx = fnrm[3 * i], ny = fnrm[3 * i + 1], nz = fnrm[3 * i + 2];
    if (nx * nx + ny * ny + nz * nz >= 0.5) {
      const double vn = vx * nx + vy * ny + vz * nz;
      vx -= vn * nx; vy -= vn * ny; vz -= vn * nz;
      const double vnm = fabs(vn);
      const double vt = sqrt(vx * vx + vy * vy + vz * vz);
      if (vt > 1e-8) {
        const double fi = f.mu_f * vnm;
        if (fi >= vt) { vx = 0; vy = 0; vz = 0; }
        else { const double s = 1.0 - fi / vt; vx *= s; vy *= s; vz *= s; }
      }
      double dmp = f.cdamp;
      if (f.adaptive) {
        const double vm = sqrt(vx * vx + vy * vy + vz * vz);
        const double t = fmin(fmax(1.0 - vm / fmax(f.v_thresh, 1e-6), 0.0), 1.0);
        dmp = f.cdamp + t * (f.cdamp_max - f.cdamp);
      }
      vx *= 1.0 - dmp; vy *= 1.0 - dmp; vz *= 1.0 - dmp;
    }
  }
  if (ip->active > 0) { vx *= 0.95; vy *= 0.95; vz *= 0.95; }
  if (f.has_ground && free_v && x[3 * i + 1] <= f.ground) {
    if (vy < 0) vy = 0;
    prev[3 * i + 1] = f.ground;
    vx *= 0.5; vz *= 0.5;
  }
  vx *= f.one_minus_damp; vy *= f.one_minus_damp; vz *= f.one_minus_damp;
  if (free_v && f.frayl > 0) { vx *= f.frayl; vy *= f.frayl; vz *= f.frayl; }
  v[3 * i] = vx; v[3 * i + 1] = vy; v[3 * i + 2] = vz;
}

// ---- intersection-freeness -------------------------------------------------

// Sign of the volume of the tetrahedron (a, b, c, d): -1, 0 or 1. A volume
// below 1e-10 of |b-a| |c-a| |d-a| counts as 0: for (nearly) coplanar points
// the computed volume is rounding noise with a random sign.
__device__ __forceinline__ int orient(D3 a, D3 b, D3 c, D3 d) {
  const D3 u = b - a, w = c - a, z = d - a;
  const double v = dot(cross(u, w), z);
  if (v * v <= 1e-20 * (len2(u) * len2(w) * len2(z))) return 0;
  return (v > 0.0) - (v < 0.0);
}
// Segment pq properly crosses triangle abc: p and q strictly on opposite
// sides of its plane, and the line pq passes strictly inside the triangle
// (the three volumes along pq have one strict sign).
__device__ __forceinline__ bool seg_crosses_tri(D3 p, D3 q, D3 a, D3 b, D3 c) {
  const int s1 = orient(a, b, c, p), s2 = orient(a, b, c, q);
  if (s1 * s2 >= 0) return false;  // same side, or an end on the plane
  const int t1 = orient(p, q, a, b), t2 = orient(p, q, b, c), t3 = orient(p, q, c, a);
  return t1 != 0 && t1 == t2 && t2 == t3;
}
// Two triangles with no shared vertex cross when an edge of one crosses
// the other. Touching (zero volumes) and coplanar overlap are not counted.
__device__ __forceinline__ bool tris_cross(const D3 P[3], const D3 Q[3]) {
  for (int e = 0; e < 3; ++e) {
    if (seg_crosses_tri(P[e], P[(e + 1) % 3], Q[0], Q[1], Q[2])) return true;
    if (seg_crosses_tri(Q[e], Q[(e + 1) % 3], P[0], P[1], P[2])) return true;
  }
  return false;
}

// Counts the candidate pairs (boxes with margin 0) whose triangles share no
// vertex and cross: the intersection-freeness check of a contact scene.
__global__ void __launch_bounds__(256) k_tri_crossings(const int2* __restrict__ pairs, const unsigned int* npairs_p,
                                                       unsigned int cap, const int* __restrict__ tri,
                                                       const double* __restrict__ x, BvhDev* bd) {
  const unsigned int np = *npairs_p < cap ? *npairs_p : cap;
  unsigned int hits = 0;
  for (PairRange pr_ = pair_range(np); pr_.k < pr_.end; pr_.k += pr_.stride) {
    const unsigned int k = pr_.k;
    const int2 pr = pairs[k];
    const int A[3] = {tri[3 * pr.x], tri[3 * pr.x + 1], tri[3 * pr.x + 2]};
    const int B[3] = {tri[3 * pr.y], tri[3 * pr.y + 1], tri[3 * pr.y + 2]};
    bool shared = false;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) shared |= A[i] == B[j];
    if (!shared) {
      const D3 P[3] = {ld(x, A[0]), ld(x, A[1]), ld(x, A[2])};
      const D3 Q[3] = {ld(x, B[0]), ld(x, B[1]), ld(x, B[2])};
      hits += tris_cross(P, Q) ? 1u : 0u;
    }
  }
  for (int o = 32; o > 0; o >>= 1) hits += __shfl_xor(hits, o);
  if ((threadIdx.x & 63) == 0 && hits) atomicAdd(&bd->crossings, hits);
}

}  // namespace vhc

#pragma clang fp contract(fast)
