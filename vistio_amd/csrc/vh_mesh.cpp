// Mesh helpers of the C ABI: the reference's quad grid generator and the
// interior edges of its Topology, in a canonical order.
//   quad_grid        crates/vistio-mesh/src/generators.rs:28-100
//   Topology::build  crates/vistio-mesh/src/topology.rs:58-120 (+ find_wing_vertex :158-168)
// Topology::build iterates a randomly seeded HashMap, so its interior-edge
// order differs per process; here interior edges are sorted by (v0, v1),
// with tri_a < tri_b as the reference pushes triangles in index order.
#include <algorithm>
#include <cstdint>
#include <vector>

#include "vistio_hip.h"

extern "C" {

vh_status vh_mesh_quad_grid(uint32_t cols, uint32_t rows, float width, float height, float* px, float* py,
                            float* pz, uint32_t* indices) {
  if (cols == 0 || rows == 0 || !px || !py || !pz || !indices) return VH_E_INVALID_ARG;
  const uint32_t vx = cols + 1, vy = rows + 1;
  const float half_w = width / 2.0f, half_h = height / 2.0f;
  for (uint32_t j = 0; j < vy; ++j)
    for (uint32_t i = 0; i < vx; ++i) {
      const float u = (float)i / (float)cols;
      const float v = (float)j / (float)rows;
      const uint32_t k = j * vx + i;
      px[k] = -half_w + u * width;
      py[k] = half_h - v * height;
      pz[k] = 0.0f;
    }
  uint32_t* o = indices;
  for (uint32_t j = 0; j < rows; ++j)
    for (uint32_t i = 0; i < cols; ++i) {
      const uint32_t tl = j * vx + i, tr = tl + 1, bl = tl + vx, br = bl + 1;
      if ((i + j) % 2 == 0) {
        *o++ = tl; *o++ = bl; *o++ = tr;
        *o++ = tr; *o++ = bl; *o++ = br;
      } else {
        *o++ = tl; *o++ = bl; *o++ = br;
        *o++ = tl; *o++ = br; *o++ = tr;
      }
    }
  return VH_OK;
}

vh_status vh_mesh_interior_edges(uint32_t n, uint32_t t, const uint32_t* indices, uint32_t* out, uint32_t cap,
                                 uint32_t* count) {
  if (!count || (t > 0 && !indices)) return VH_E_INVALID_ARG;
  struct E { uint64_t key; uint32_t tri, wing; };
  std::vector<E> es;
  es.reserve(3 * (size_t)t);
  for (uint32_t f = 0; f < t; ++f) {
    const uint32_t a = indices[3 * f], b = indices[3 * f + 1], c = indices[3 * f + 2];
    if (a >= n || b >= n || c >= n) return VH_E_INVALID_MESH;
    const uint32_t ev[3][3] = {{a, b, c}, {b, c, a}, {c, a, b}};
    for (const auto& x : ev) {
      const uint64_t lo = std::min(x[0], x[1]), hi = std::max(x[0], x[1]);
      es.push_back({lo * n + hi, f, x[2]});
    }
  }
  std::sort(es.begin(), es.end(), [](const E& p, const E& q) { return p.key < q.key || (p.key == q.key && p.tri < q.tri); });
  uint32_t cnt = 0;
  for (size_t k = 0; k < es.size();) {
    size_t j = k;
    while (j < es.size() && es[j].key == es[k].key) ++j;
    if (j - k == 2) {
      if (out && cnt < cap) {
        uint32_t* o = out + 6 * (size_t)cnt;
        o[0] = (uint32_t)(es[k].key / n); o[1] = (uint32_t)(es[k].key % n);
        o[2] = es[k].wing; o[3] = es[k + 1].wing;
        o[4] = es[k].tri; o[5] = es[k + 1].tri;
      }
      ++cnt;
    }
    k = j;
  }
  *count = cnt;
  return (out && cnt > cap) ? VH_E_INVALID_ARG : VH_OK;
}

}  // extern "C"
