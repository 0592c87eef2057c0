// The pre-IPC narrow phase of the reference's collision pipeline
// (VertexTriangleTest, crates/vistio-contact/src/vertex_triangle.rs) on the
// GPU, in f32 with the reference's operation order and no FMA contraction,
// so each contact carries the reference's bits.
//
// The reference loops over every triangle for every candidate (a = vertex
// vi, b = vertex vj) and keeps the triangles that contain vj and not vi. Here
// each candidate reads vj's incident triangles from a vertex -> triangle CSR
// (ascending triangle order, as the reference's loop meets them): one pass
// counts each candidate's contacts, a scan turns the counts into offsets, a
// second pass writes the contacts there, in candidate then triangle order.
#pragma once
#include <hip/hip_runtime.h>

#include "vistio_hip.h"

#pragma clang fp contract(off)

namespace vhn {

struct F3 {
  float x, y, z;
};
__device__ __forceinline__ F3 sub(F3 a, F3 b) { return F3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// glam's Vec3::cross: (y z' - y' z, z x' - z' x, x y' - x' y)
__device__ __forceinline__ F3 cross(F3 a, F3 b) {
  return F3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}

// point_triangle_proximity (vertex_triangle.rs:88-152).
__device__ __forceinline__ bool proximity(F3 p, F3 a, F3 b, F3 c, float thickness, vh_contact_pair& o) {
  const F3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
  F3 n = cross(ab, ac);
  const float area2 = sqrtf(dot(n, n));
  if (area2 < 1e-10f) return false;  // degenerate triangle
  n = F3{n.x / area2, n.y / area2, n.z / area2};
  const float sd = dot(ap, n);  // signed distance to the triangle's plane
  const float ad = fabsf(sd);
  if (ad > thickness) return false;
  const F3 proj = F3{p.x - n.x * sd, p.y - n.y * sd, p.z - n.z * sd};
  const F3 app = sub(proj, a);
  const float d00 = dot(ab, ab), d01 = dot(ab, ac), d11 = dot(ac, ac), d20 = dot(app, ab), d21 = dot(app, ac);
  const float den = d00 * d11 - d01 * d01;
  if (fabsf(den) < 1e-10f) return false;
  const float inv = 1.0f / den;
  const float v = (d11 * d20 - d01 * d21) * inv;
  const float w = (d00 * d21 - d01 * d20) * inv;
  const float u = 1.0f - v - w;
  const float tol = -0.01f;
  if (u < tol || v < tol || w < tol) return false;
  const bool up = sd >= 0.0f;
  o.distance = -ad;
  o.normal[0] = up ? n.x : -n.x;
  o.normal[1] = up ? n.y : -n.y;
  o.normal[2] = up ? n.z : -n.z;
  o.barycentric[0] = u;
  o.barycentric[1] = v;
  o.barycentric[2] = w;
  o.contact_type = 0;
  o.is_self = 1;
  return true;
}

struct NarrowArgs {
  const float *px, *py, *pz;
  const unsigned int* tri;   // 3t
  const int* vptr;           // n + 1: vertex -> incident triangles
  const int* vtri;           // ascending per vertex
  const unsigned int *ca, *cb;
  long long nc;
  float thickness;
};

// kEmit = false: cnt[k] = candidate k's contacts. kEmit = true: writes them
// from off[k] (contacts past cap are counted, not written).
template <bool kEmit>
__global__ void __launch_bounds__(256) k_narrow_vt(NarrowArgs a, unsigned int* __restrict__ cnt,
                                                   const unsigned int* __restrict__ off,
                                                   vh_contact_pair* __restrict__ out, unsigned long long cap) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.nc) return;
  const unsigned int vi = a.ca[k], vj = a.cb[k];
  const F3 p = F3{a.px[vi], a.py[vi], a.pz[vi]};
  unsigned long long at = kEmit ? off[k] : 0ull;
  unsigned int c = 0;
  for (int q = a.vptr[vj]; q < a.vptr[vj + 1]; ++q) {
    const int e = a.vtri[q];
    const unsigned int A = a.tri[3 * (size_t)e], B = a.tri[3 * (size_t)e + 1], C = a.tri[3 * (size_t)e + 2];
    if (vi == A || vi == B || vi == C) continue;
    vh_contact_pair o;
    if (!proximity(p, F3{a.px[A], a.py[A], a.pz[A]}, F3{a.px[B], a.py[B], a.pz[B]}, F3{a.px[C], a.py[C], a.pz[C]},
                   a.thickness, o))
      continue;
    if (kEmit) {
      if (at < cap) {
        o.indices[0] = vi; o.indices[1] = A; o.indices[2] = B; o.indices[3] = C;
        out[at] = o;
      }
      ++at;
    }
    ++c;
  }
  if (!kEmit) cnt[k] = c;
}

}  // namespace vhn

#pragma clang fp contract(fast)
