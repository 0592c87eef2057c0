// Test-only CPU model of the HIP solve schedule (tests/ infrastructure).
// Runs the product's host set-up (vistio_amd/csrc/vh_setup.cpp) and replays
// the forward/backward task lists exactly as k_fwd / k_bwd do, sequentially
// in double, so the set-up's supernode blocks, extend-add maps and task
// schedule can be checked against an independent solve without a GPU.
#include <cstdint>
#include <cstring>
#include <vector>

#include "vh_setup.hpp"

extern "C" int emul_solve(int n, const float* px, const float* py, const float* pz, int t, const uint32_t* tri,
                          int h, const uint32_t* hinge, const uint8_t* pinned, double k_mem, double k_bend,
                          double density, double dt_matrix, int leaf, const double* b, double* x, long long* stats) {
  vh::SetupInput in;
  in.n = n; in.t = t; in.h = h; in.px = px; in.py = py; in.pz = pz; in.tri = tri; in.hinge = hinge;
  in.pinned = pinned; in.k_membrane = k_mem; in.k_bending = k_bend; in.density = density;
  in.dt_matrix = dt_matrix; in.leaf_size = leaf;
  vh::Setup S;
  if (!vh::build_setup(in, S)) return S.status ? S.status : -1;
  std::vector<double> r(3 * (size_t)n), y(3 * (size_t)n), d(3 * (size_t)n), U(3 * (size_t)S.u_total + 3);
  for (int k = 0; k < n; ++k)
    for (int c = 0; c < 3; ++c) r[3 * k + c] = b[(size_t)c * n + S.perm[k]];
  for (int l = 0; l <= S.max_depth; ++l)
    for (int q = S.fwd_level_off[l]; q < S.fwd_level_off[l + 1]; ++q) {
      const vh::Task& T = S.fwd_tasks[q];
      const vh::Supernode& P = S.sn[T.sn];
      const int w = P.w, m = P.w + P.r;
      std::vector<double> front(3 * (size_t)m, 0.0);
      for (int c = 0; c < P.nchild; ++c) {
        const vh::Supernode& C = S.sn[S.children[P.child_off + c]];
        for (int qq = 0; qq < C.r; ++qq) {
          const int pos = S.emap[C.map_off + qq];
          for (int k = 0; k < 3; ++k) front[3 * pos + k] += U[3 * (size_t)(C.u_off + qq) + k];
        }
      }
      for (int j = 0; j < w; ++j)
        for (int k = 0; k < 3; ++k) front[3 * j + k] = r[3 * (size_t)(P.col0 + j) + k] - front[3 * j + k];
      const double* Mp = &S.M[P.m_off];
      const long long tri_n = (long long)w * (w + 1) / 2;
      for (int row = T.b; row < T.e; ++row) {
        const int len = row < w ? row + 1 : w;
        const double* Mr = Mp + (row < w ? (long long)row * (row + 1) / 2 : tri_n + (long long)(row - w) * w);
        double s[3] = {0, 0, 0};
        for (int kk = 0; kk < len; ++kk)
          for (int k = 0; k < 3; ++k) s[k] += Mr[kk] * front[3 * kk + k];
        for (int k = 0; k < 3; ++k) {
          if (row < w) y[3 * (size_t)(P.col0 + row) + k] = s[k];
          else U[3 * (size_t)(P.u_off + row - w) + k] = s[k] + front[3 * row + k];
        }
      }
    }
  for (int l = 0; l <= S.max_depth; ++l)
    for (int q = S.bwd_level_off[l]; q < S.bwd_level_off[l + 1]; ++q) {
      const vh::Task& T = S.bwd_tasks[q];
      const vh::Supernode& P = S.sn[T.sn];
      const int w = P.w, m = P.w + P.r;
      std::vector<double> v(3 * (size_t)m);
      for (int kk = 0; kk < m; ++kk)
        for (int k = 0; k < 3; ++k)
          v[3 * kk + k] = kk < w ? y[3 * (size_t)(P.col0 + kk) + k] : -d[3 * (size_t)S.rowidx[P.rows_off + kk - w] + k];
      const double* Mp = &S.M[P.m_off];
      const long long tri_n = (long long)w * (w + 1) / 2;
      for (int j = T.b; j < T.e; ++j) {
        double s[3] = {0, 0, 0};
        for (int kk = j; kk < w; ++kk)
          for (int k = 0; k < 3; ++k) s[k] += Mp[(long long)kk * (kk + 1) / 2 + j] * v[3 * kk + k];
        for (int qq = 0; qq < P.r; ++qq)
          for (int k = 0; k < 3; ++k) s[k] += Mp[tri_n + (long long)qq * w + j] * v[3 * (w + qq) + k];
        for (int k = 0; k < 3; ++k) d[3 * (size_t)(P.col0 + j) + k] = s[k];
      }
    }
  for (int k = 0; k < n; ++k)
    for (int c = 0; c < 3; ++c) x[(size_t)c * n + S.perm[k]] = d[3 * k + c];
  if (stats) {
    stats[0] = (long long)S.sn.size(); stats[1] = S.max_depth + 1; stats[2] = S.nnz_factor;
    stats[3] = (long long)S.Ai.size(); stats[4] = S.max_front;
    stats[5] = (long long)S.fwd_tasks.size(); stats[6] = (long long)S.bwd_tasks.size();
  }
  return 0;
}
