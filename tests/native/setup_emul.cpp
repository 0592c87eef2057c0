// Test-only CPU model of the HIP solve schedule (tests/ infrastructure).
// Runs the product's host set-up (vistio_amd/csrc/vh_setup.cpp) and replays
// the forward, root and backward task lists as k_fwd_band / k_root_band /
// k_bwd_band do, sequentially in double, so the set-up's supernode blocks,
// push slots, root inverse and task schedule can be checked against an
// independent solve without a GPU.
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <vector>

#include "vh_setup.hpp"

bool host_factor(const vh::Setup& S, std::vector<double>& M, std::vector<double>& Mt, std::vector<double>& Sroot);

extern "C" int emul_solve(int n, const float* px, const float* py, const float* pz, int t, const uint32_t* tri,
                          int h, const uint32_t* hinge, const uint8_t* pinned, double k_mem, double k_bend,
                          double density, double dt_matrix, int leaf, int max_band_levels, const double* b,
                          double* x, long long* stats) {
  vh::SetupInput in;
  in.n = n; in.t = t; in.h = h; in.px = px; in.py = py; in.pz = pz; in.tri = tri; in.hinge = hinge;
  in.pinned = pinned; in.k_membrane = k_mem; in.k_bending = k_bend; in.density = density;
  in.dt_matrix = dt_matrix; in.leaf_size = leaf;
  if (max_band_levels > 0) in.max_band_levels = max_band_levels;
  vh::Setup S;
  if (!vh::build_setup(in, S)) return S.status ? S.status : -1;
  std::vector<double> Mv, Mtv, Srv;
  if (!host_factor(S, Mv, Mtv, Srv)) return -2;
  std::vector<double> r(3 * (size_t)n), y(3 * (size_t)n), d(3 * (size_t)n), U(3 * (size_t)S.u_total + 3);
  for (int k = 0; k < n; ++k)
    for (int c = 0; c < 3; ++c) r[3 * k + c] = b[(size_t)c * n + S.perm[k]];
  // Replay the band schedule task by task (forward bands deepest first,
  // backward bands root first), with the push-slot data flow of the GPU
  // kernels: children write update rows into the parent's front slots,
  // backward tasks push -d into the descendants' v slots.
  const int nb = S.max_depth + 1;
  const int nfwd = (int)S.fwd_tasks.size();
  std::vector<double> F(3 * (size_t)S.fslot_total + 3, 0.0), V(3 * (size_t)S.u_total + 3, 0.0);
  for (int l = 0; l < nb; ++l)
    for (int t = S.fwd_level_off[l]; t < S.fwd_level_off[l + 1]; ++t) {
      const vh::Task& T = S.ptasks[t];
      const vh::Supernode& P = S.sn[T.sn];
      if (P.depth != S.max_depth - l) return -10;
      if ((T.kind == 2) != (l == nb - 1) || (T.kind == 2) != (P.parent < 0)) return -12;
      const int w = P.w, m = P.w + P.r;
      std::vector<double> front(3 * (size_t)m, 0.0);
      for (int j = 0; j < m; ++j)
        for (int q = 0; q < P.fslot_k; ++q)
          for (int k = 0; k < 3; ++k) front[3 * j + k] += F[3 * ((size_t)P.fm_off + (size_t)j * P.fslot_k + q) + k];
      for (int j = 0; j < w; ++j)
        for (int k = 0; k < 3; ++k) front[3 * j + k] = r[3 * (size_t)(P.col0 + j) + k] - front[3 * j + k];
      if (T.kind == 2) {  // root: d = Sroot front, pushed to the descendants
        if (P.r != 0 || P.s_off < 0 || Srv.size() < (size_t)P.s_off + (size_t)w * w) return -13;
        for (int row = T.b; row < T.e; ++row) {
          const double* Sr = &Srv[(size_t)P.s_off + (size_t)row * w];
          double s3[3] = {0, 0, 0};
          for (int kk = 0; kk < w; ++kk)
            for (int k = 0; k < 3; ++k) s3[k] += Sr[kk] * front[3 * kk + k];
          const int col = P.col0 + row;
          for (int k = 0; k < 3; ++k) d[3 * (size_t)col + k] = s3[k];
          for (int q = S.dpush_ptr[col]; q < S.dpush_ptr[col + 1]; ++q)
            for (int k = 0; k < 3; ++k) V[3 * (size_t)S.dpush_idx[q] + k] = -s3[k];
        }
        continue;
      }
      const double* Mp = &Mv[P.m_off];
      const long long tri_n = (long long)w * (w + 1) / 2;
      for (int row = T.b; row < T.e; ++row) {
        const int len = row < w ? row + 1 : w;
        const double* Mr = Mp + (row < w ? (long long)row * (row + 1) / 2 : tri_n + (long long)(row - w) * w);
        double s3[3] = {0, 0, 0};
        for (int kk = 0; kk < len; ++kk)
          for (int k = 0; k < 3; ++k) s3[k] += Mr[kk] * front[3 * kk + k];
        for (int k = 0; k < 3; ++k) {
          if (row < w) y[3 * (size_t)(P.col0 + row) + k] = s3[k];
          else F[3 * (size_t)S.uslot[P.u_off + row - w] + k] = s3[k] + front[3 * row + k];
        }
      }
    }
  for (int l = 0; l < nb; ++l)
    for (int t = S.bwd_level_off[l]; t < S.bwd_level_off[l + 1]; ++t) {
      const vh::Task& T = S.ptasks[nfwd + t];
      const vh::Supernode& P = S.sn[T.sn];
      if (P.depth != l || l == 0 || T.kind != 1) return -11;
      const int w = P.w, m = P.w + P.r;
      std::vector<double> v(3 * (size_t)m);
      for (int kk = T.b; kk < m; ++kk)
        for (int k = 0; k < 3; ++k)
          v[3 * kk + k] = kk < w ? y[3 * (size_t)(P.col0 + kk) + k] : V[3 * (size_t)(P.u_off + kk - w) + k];
      const double* Tp = &Mtv[P.m_off];
      for (int j = T.b; j < T.e; ++j) {
        const double* row = Tp + (long long)j * m - (long long)j * (j - 1) / 2;
        double s3[3] = {0, 0, 0};
        for (int i = 0; i < m - j; ++i)
          for (int k = 0; k < 3; ++k) s3[k] += row[i] * v[3 * (j + i) + k];
        const int col = P.col0 + j;
        for (int k = 0; k < 3; ++k) d[3 * (size_t)col + k] = s3[k];
        for (int q = S.dpush_ptr[col]; q < S.dpush_ptr[col + 1]; ++q)
          for (int k = 0; k < 3; ++k) V[3 * (size_t)S.dpush_idx[q] + k] = -s3[k];
      }
    }
  for (int k = 0; k < n; ++k)
    for (int c = 0; c < 3; ++c) x[(size_t)c * n + S.perm[k]] = d[3 * k + c];
  if (stats) {
    stats[0] = (long long)S.sn.size(); stats[1] = S.max_depth + 1; stats[2] = S.nnz_factor;
    stats[3] = (long long)S.Ai.size(); stats[4] = S.max_front;
    stats[5] = (long long)S.fwd_tasks.size(); stats[6] = (long long)S.bwd_tasks.size();
    stats[7] = (long long)S.band_start.size();
  }
  return 0;
}


// 1 when the device's 32-bit offsets hold these totals (the set-up's own check).
extern "C" int emul_index_fits(long long m_total, long long sroot_entries, long long front_slots,
                               long long rhs_values) {
  return vh::device_index_overflow(m_total, sroot_entries, front_slots, rhs_values).empty() ? 1 : 0;
}

// The set-up of vh_pd_init_ex's defaults (value_bytes 8 or 4, the band cost
// model's defaults, per-element materials optional) and the host factor of
// it: sizes[0] = m_total, sizes[1] = sroot_total; M / Mt / Sr are written
// when not null (m_total, m_total, sroot_total doubles). Returns 0, or the
// set-up's status, or -2 when the factor fails.
extern "C" int emul_factor(int n, const float* px, const float* py, const float* pz, int t, const uint32_t* tri,
                           int h, const uint32_t* hinge, const uint8_t* pinned, double k_mem, double k_bend,
                           double density, const double* km, const double* kb, const double* dens, int value_bytes,
                           int max_band_levels, double* M, double* Mt, double* Sr, long long* sizes) {
  vh::SetupInput in;
  in.n = n; in.t = t; in.h = h; in.px = px; in.py = py; in.pz = pz; in.tri = tri; in.hinge = hinge;
  in.pinned = pinned; in.k_membrane = k_mem; in.k_bending = k_bend; in.density = density;
  in.k_membrane_tri = km; in.k_bending_hinge = kb; in.density_tri = dens;
  in.value_bytes = value_bytes;
  if (max_band_levels > 0) in.max_band_levels = max_band_levels;
  vh::Setup S;
  if (!vh::build_setup(in, S)) return S.status ? S.status : -1;
  sizes[0] = S.m_total;
  sizes[1] = S.sroot_total;
  if (!M) return 0;
  std::vector<double> Mv, Mtv, Srv;
  if (!host_factor(S, Mv, Mtv, Srv)) return -2;
  std::copy(Mv.begin(), Mv.end(), M);
  std::copy(Mtv.begin(), Mtv.end(), Mt);
  std::copy(Srv.begin(), Srv.end(), Sr);
  return 0;
}

// The local step's RHS slot plan (vh::Setup::lv_ptr .. lc_pos), replayed as
// k_local uses it: every corner of every chunk is summed exactly once, into
// a slot of its own vertex; every slot is written by exactly one (chunk,
// vertex); a vertex's slots follow chunk order; a local vertex's corners are
// in ascending place in the block. Corner (element e, k) carries the value
// 1 + (index of the corner in 3t + 4h) * 0.5 (exact in double), so the plan's
// per-vertex sums equal the direct per-corner sums exactly. stats[0] = slots,
// stats[1] = corners, stats[2] = most local vertices of one chunk. Returns 0,
// or a negative code naming the first broken property.
extern "C" int emul_local_plan(int n, const float* px, const float* py, const float* pz, int t, const uint32_t* tri,
                               int h, const uint32_t* hinge, const uint8_t* pinned, int leaf, long long* stats) {
  vh::SetupInput in;
  in.n = n; in.t = t; in.h = h; in.px = px; in.py = py; in.pz = pz; in.tri = tri; in.hinge = hinge;
  in.pinned = pinned; in.k_membrane = 1e3; in.k_bending = 1e-4; in.density = 0.15;
  in.dt_matrix = 1.0 / 60.0; in.leaf_size = leaf;
  vh::Setup S;
  if (!vh::build_setup(in, S)) return S.status ? S.status : -1;
  const int tb = (t + 255) / 256, hb = (h + 255) / 256, nch = tb + hb;
  if ((int)S.lv_ptr.size() != nch + 1 || S.lv_ptr[0] != 0) return -2;
  const int64_t nslot = S.inc_ptr[n];
  if ((int64_t)S.lv_slot.size() != S.lv_ptr[nch] || nslot != S.lv_ptr[nch]) return -3;
  std::vector<int> owner(nslot, -1);  // vertex of each slot (from inc_ptr)
  for (int v = 0; v < n; ++v)
    for (int s = S.inc_ptr[v]; s < S.inc_ptr[v + 1]; ++s) owner[s] = v;
  std::vector<int> slot_chunk(nslot, -1);
  std::vector<double> got(nslot, 0.0), want_v(n, 0.0), got_v(n, 0.0);
  long long most = 0;
  for (int c = 0; c < nch; ++c) {
    const bool tr = c < tb;
    const int K = tr ? 3 : 4;
    const int64_t e0 = (int64_t)(tr ? c : c - tb) * 256;
    const int64_t ne = std::min<int64_t>(256, (tr ? t : h) - e0);
    const int32_t* idx = tr ? S.tri_idx.data() : S.hinge_idx.data();
    const size_t cb = tr ? 3 * (size_t)e0 : 3 * (size_t)t + 4 * (size_t)e0;
    std::vector<int> seen((size_t)ne * K, 0);
    int prev_end = 0;
    most = std::max<long long>(most, S.lv_ptr[c + 1] - S.lv_ptr[c]);
    for (int j = S.lv_ptr[c]; j < S.lv_ptr[c + 1]; ++j) {
      const int a = prev_end, b = S.lv_end[j];
      if (b <= a) return -4;
      prev_end = b;
      const int s = S.lv_slot[j];
      if (s < 0 || s >= nslot || slot_chunk[s] != -1) return -5;
      slot_chunk[s] = c;
      const int v = owner[s];
      int last = -1;
      for (int q = a; q < b; ++q) {
        const int p = S.lc_pos[cb + q];
        if (p <= last || p >= ne * K) return -6;
        last = p;
        seen[p]++;
        const size_t corner = cb + (size_t)p;  // == e0 * K + p
        if (idx[(size_t)e0 * K + p] != v) return -7;
        got[s] += 1.0 + 0.5 * (double)corner;
      }
    }
    if (prev_end != ne * K) return -8;
    for (int x : seen)
      if (x != 1) return -9;
  }
  for (int v = 0; v < n; ++v) {
    for (int s = S.inc_ptr[v]; s < S.inc_ptr[v + 1]; ++s) {
      if (slot_chunk[s] < 0) return -10;
      if (s > S.inc_ptr[v] && slot_chunk[s] <= slot_chunk[s - 1]) return -11;
      got_v[v] += got[s];
    }
  }
  for (int64_t k = 0; k < 3 * (int64_t)t; ++k) want_v[S.tri_idx[k]] += 1.0 + 0.5 * (double)k;
  for (int64_t k = 0; k < 4 * (int64_t)h; ++k) want_v[S.hinge_idx[k]] += 1.0 + 0.5 * (double)(3 * (int64_t)t + k);
  for (int v = 0; v < n; ++v)
    if (got_v[v] != want_v[v]) return -12;
  stats[0] = nslot;
  stats[1] = 3 * (long long)t + 4 * (long long)h;
  stats[2] = most;
  return 0;
}
