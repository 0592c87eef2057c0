// Host-side set-up for the HIP PD path. See vh_setup.hpp for the contract.
//
// The rest-state formulas restate the reference's (cited per function) in
// f64. The factorization is this library's own: nested dissection by
// recursive coordinate bisection, one supernode per dissection-tree node, a
// multifrontal dense LL^T per supernode, and the partitioned-inverse blocks
// M_P = [L_PP^{-1}; L_IP L_PP^{-1}] that turn both triangular sweeps of the
// GPU solve into independent GEMVs per supernode.
#include "vh_setup.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <functional>
#include <numeric>
#include <thread>

#include "vistio_hip.h"

namespace vh {
namespace {

struct D3 {
  double x, y, z;
};
inline D3 sub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 scale(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline D3 divs(D3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3 cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double len(D3 a) { return std::sqrt(dot(a, a)); }

// ord (the identity on entry) reordered stably by key(e) in [0, nkeys): a
// counting sort, the order std::stable_sort gives.
template <class K>
void stable_order_by_key(std::vector<int>& ord, int nkeys, K key) {
  std::vector<int> at((size_t)nkeys + 1, 0);
  for (int e : ord) at[(size_t)key(e) + 1]++;
  for (int k = 0; k < nkeys; ++k) at[k + 1] += at[k];
  std::vector<int> out(ord.size());
  for (int e : ord) out[(size_t)at[key(e)]++] = e;
  ord.swap(out);
}

// Host threads for `count` independent items, at least min_per items each,
// at most 16.
int host_threads(int64_t count, int64_t min_per) {
  const int64_t hw = std::max(1u, std::thread::hardware_concurrency());
  return (int)std::max<int64_t>(1, std::min<int64_t>(std::min<int64_t>(hw, 16), count / std::max<int64_t>(1, min_per)));
}
// Runs f(begin, end, k) over [0, count) cut into nt contiguous ranges, range
// k on thread k (in the calling thread when nt == 1).
template <class F>
void parallel_ranges(int64_t count, int nt, F&& f) {
  if (nt <= 1) { f((int64_t)0, count, 0); return; }
  std::vector<std::thread> th;
  th.reserve(nt);
  for (int k = 0; k < nt; ++k) th.emplace_back([&f, count, nt, k] { f(count * k / nt, count * (k + 1) / nt, k); });
  for (auto& x : th) x.join();
}

// compute_dm_inv — crates/vistio-solver/src/element.rs:228-263.
void dm_inverse(D3 e1, D3 e2, double out[4]) {
  out[0] = 1; out[1] = 0; out[2] = 0; out[3] = 1;
  const double l1 = len(e1);
  if (l1 < 1e-10) return;
  const D3 u = divs(e1, l1);
  const D3 nrm = cross(e1, e2);
  if (len(nrm) < 1e-10) return;
  const D3 vv = cross(nrm, u);
  const D3 v = scale(vv, 1.0 / len(vv));
  const double dm00 = dot(e1, u), dm10 = dot(e1, v), dm01 = dot(e2, u), dm11 = dot(e2, v);
  const double det = dm00 * dm11 - dm01 * dm10;
  if (std::fabs(det) < 1e-10) return;
  const double inv = 1.0 / det;
  out[0] = dm11 * inv;
  out[1] = -dm10 * inv;
  out[2] = -dm01 * inv;
  out[3] = dm00 * inv;
}

// compute_dihedral_angle — crates/vistio-solver/src/bending.rs:56-91.
double dihedral(D3 p0, D3 p1, D3 pa, D3 pb) {
  const D3 edge = sub(p1, p0);
  const double el = len(edge);
  if (el < 1e-10) return M_PI;
  const D3 e = divs(edge, el);
  const D3 ta = sub(pa, p0), tb = sub(pb, p0);
  D3 qa = sub(ta, scale(e, dot(ta, e)));
  D3 qb = sub(tb, scale(e, dot(tb, e)));
  const double la = len(qa), lb = len(qb);
  if (la < 1e-10 || lb < 1e-10) return M_PI;
  qa = divs(qa, la);
  qb = divs(qb, lb);
  double c = dot(qa, qb);
  c = std::min(1.0, std::max(-1.0, c));
  return std::acos(c);
}

// Nested dissection by recursive coordinate bisection. The vertex separator
// is the one-sided boundary of the median cut on the side where it is
// smaller. Nodes are appended in postorder (left, right, separator).
struct NdNode {
  std::vector<int> verts;
  int left = -1, right = -1;
};

struct NdBuilder {
  const std::vector<int32_t>& Ap;
  const std::vector<int32_t>& Ai;
  const double* X;
  const double* Y;
  const double* Z;
  int leaf;
  int* mark;  // per vertex; builders of disjoint components may share it
  int stamp = 0;
  std::vector<NdNode> nodes;

  int build(std::vector<int>& verts) {
    if ((int)verts.size() <= leaf) {
      NdNode nd;
      nd.verts = verts;
      nodes.push_back(std::move(nd));
      return (int)nodes.size() - 1;
    }
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
    const int L = ++stamp, R = ++stamp, S = ++stamp;
    for (size_t i = 0; i < s.size(); ++i) mark[s[i]] = i < half ? L : R;
    std::vector<int> bl, br;
    for (int v : s) {
      const int other = mark[v] == L ? R : L;
      for (int p = Ap[v]; p < Ap[v + 1]; ++p)
        if (mark[Ai[p]] == other) { (mark[v] == L ? bl : br).push_back(v); break; }
    }
    std::vector<int>& sep = bl.size() <= br.size() ? bl : br;
    for (int v : sep) mark[v] = S;
    std::vector<int> left, right;
    for (int v : s) {
      if (mark[v] == L) left.push_back(v);
      else if (mark[v] == R) right.push_back(v);
    }
    if (left.empty() || right.empty()) {  // cannot split: keep as one dense leaf
      NdNode nd;
      nd.verts = verts;
      nodes.push_back(std::move(nd));
      return (int)nodes.size() - 1;
    }
    std::vector<int> sepv(sep);
    const int a = build(left);
    const int b = build(right);
    NdNode nd;
    nd.verts = std::move(sepv);
    nd.left = a;
    nd.right = b;
    nodes.push_back(std::move(nd));
    return (int)nodes.size() - 1;
  }
};

}  // namespace

bool build_setup(const SetupInput& in, Setup& S) {
  const int n = in.n, t = in.t, h = in.h;
  S.n = n; S.t = t; S.h = h;
  if (n <= 0 || t < 0 || h < 0 || !in.px || !in.py || !in.pz || !in.pinned || (t > 0 && !in.tri) ||
      (h > 0 && !in.hinge)) {
    S.status = VH_E_INVALID_MESH; S.err = "empty mesh or null input array"; return false;
  }
  for (int64_t k = 0; k < 3 * (int64_t)t; ++k)
    if (in.tri[k] >= (uint32_t)n) { S.status = VH_E_INVALID_MESH; S.err = "triangle index out of range"; return false; }
  for (int64_t k = 0; k < 4 * (int64_t)h; ++k)
    if (in.hinge[k] >= (uint32_t)n) { S.status = VH_E_INVALID_MESH; S.err = "hinge index out of range"; return false; }
  if (!(in.dt_matrix > 0)) { S.status = VH_E_INVALID_CONFIG; S.err = "dt_matrix must be > 0"; return false; }

  std::vector<double> X(n), Y(n), Z(n);
  for (int i = 0; i < n; ++i) { X[i] = in.px[i]; Y[i] = in.py[i]; Z[i] = in.pz[i]; }
  auto P = [&](int i) { return D3{X[i], Y[i], Z[i]}; };

  // Rest data, element.rs:54-91 (weight = k * area, element.rs:77).
  std::vector<double> dm(4 * (size_t)t), area(t), tw(t);
  for (int e = 0; e < t; ++e) {
    const int i0 = in.tri[3 * e], i1 = in.tri[3 * e + 1], i2 = in.tri[3 * e + 2];
    const D3 e1 = sub(P(i1), P(i0)), e2 = sub(P(i2), P(i0));
    area[e] = 0.5 * len(cross(e1, e2));
    dm_inverse(e1, e2, &dm[4 * (size_t)e]);
    tw[e] = (in.k_membrane_tri ? in.k_membrane_tri[e] : in.k_membrane) * area[e];
  }
  // Lumped masses, pd_solver.rs:1219-1242.
  std::vector<double> mass(n, 0.0);
  for (int e = 0; e < t; ++e) {
    const double third = (area[e] * (in.density_tri ? in.density_tri[e] : in.density)) / 3.0;
    for (int k = 0; k < 3; ++k) mass[in.tri[3 * e + k]] += third;
  }
  for (int i = 0; i < n; ++i) {
    if (in.pinned[i]) mass[i] = 1e8;
    else if (mass[i] < 1e-8) mass[i] = 1e-8;
  }
  S.mass_orig = mass;
  // Bending rest data, bending.rs:98-148 (snap to pi / 0 within 1e-3).
  const bool ds = in.bending_kind == 2;
  std::vector<double> hrest(h), hw(h), hst(4 * (size_t)h);
  for (int e = 0; e < h; ++e) {
    const uint32_t* q = in.hinge + 4 * (size_t)e;
    double a = dihedral(P(q[0]), P(q[1]), P(q[2]), P(q[3]));
    if (std::fabs(a - M_PI) < 1e-3) a = M_PI;
    else if (std::fabs(a) < 1e-3) a = 0.0;
    hrest[e] = a;
    hw[e] = (in.k_bending_hinge ? in.k_bending_hinge[e] : in.k_bending) * len(sub(P(q[1]), P(q[0])));
    double* st = &hst[4 * (size_t)e];
    st[0] = -1.0; st[1] = -1.0; st[2] = 1.0; st[3] = 1.0;  // assembly.rs:97-101
    if (!ds) continue;
    // DiscreteShellsBendingData::from_topology (discrete_shells.rs:106-195):
    // cotangent stencil from the dihedral angle's gradient, normalised to
    // unit length; a zero-length edge has no element.
    const D3 p0 = P(q[0]), p1 = P(q[1]), pa = P(q[2]), pb = P(q[3]);
    const D3 edge = sub(p1, p0);
    const double el = len(edge);
    if (el < 1e-10) { hw[e] = 0.0; continue; }
    const double area_a = 0.5 * len(cross(edge, sub(pa, p0))), area_b = 0.5 * len(cross(edge, sub(pb, p0)));
    const double inv_e2 = 1.0 / (el * el);
    const double c_wa = el / (2.0 * std::max(area_a, 1e-8)), c_wb = el / (2.0 * std::max(area_b, 1e-8));
    const double c_0 = -(dot(sub(p1, pa), edge) * inv_e2) * c_wa - (dot(sub(p1, pb), edge) * inv_e2) * c_wb;
    const double c_1 = -(dot(sub(pa, p0), edge) * inv_e2) * c_wa - (dot(sub(pb, p0), edge) * inv_e2) * c_wb;
    const double nrm = std::sqrt(c_0 * c_0 + c_1 * c_1 + c_wa * c_wa + c_wb * c_wb);
    if (nrm > 1e-10) {
      st[0] = c_0 / nrm; st[1] = c_1 / nrm; st[2] = c_wa / nrm; st[3] = c_wb / nrm;
    } else {
      st[0] = -0.5; st[1] = -0.5; st[2] = 0.5; st[3] = 0.5;
    }
    // from_topology_with_anisotropic_material (discrete_shells.rs:206-268):
    // weight *= (k_weft cos^2 + k_warp sin^2) / k_avg, the angle taken between
    // the edge's XZ projection and X (warp).
    const double avg_raw = 0.5 * (in.bend_warp + in.bend_weft);
    if (avg_raw > 1e-8) {
      const D3 dir = D3{edge.x / el, edge.y / el, edge.z / el};
      const double hl = std::sqrt(dir.x * dir.x + dir.z * dir.z);
      if (hl >= 1e-10) {
        const double c = std::fabs(dir.x / hl), c2 = c * c, s2 = 1.0 - c2;
        hw[e] *= (in.bend_weft * c2 + in.bend_warp * s2) / avg_raw;
      }
    }
  }

  // System matrix A = M/h^2 + sum w g g^T + sum w_b s s^T, assembly.rs:44-134,
  // with the same |value| > 1e-12 filter on each triplet. Row by row, rows
  // on several threads: a row's candidate columns are the vertices of its
  // elements, and its triplets are added in the order they are generated
  // (the mass term, then the triangles, then the hinges, each in element
  // order), so every entry is the same sum in the same order as one pass
  // over all triplets; an entry no triplet passed the filter for is left out.
  const double inv_dt2 = 1.0 / (in.dt_matrix * in.dt_matrix);
  // vertex -> incident triangles, then hinges, ascending, each element once
  std::vector<int32_t> itp((size_t)n + 1, 0), ihp((size_t)n + 1, 0), itl, ihl;
  {
    auto first = [](const uint32_t* q, int a) {
      for (int b = 0; b < a; ++b)
        if (q[b] == q[a]) return false;
      return true;
    };
    for (int e = 0; e < t; ++e)
      for (int a = 0; a < 3; ++a)
        if (first(in.tri + 3 * (size_t)e, a)) itp[in.tri[3 * (size_t)e + a] + 1]++;
    for (int e = 0; e < h; ++e)
      for (int a = 0; a < 4; ++a)
        if (first(in.hinge + 4 * (size_t)e, a)) ihp[in.hinge[4 * (size_t)e + a] + 1]++;
    for (int i = 0; i < n; ++i) { itp[i + 1] += itp[i]; ihp[i + 1] += ihp[i]; }
    itl.resize((size_t)itp[n]); ihl.resize((size_t)ihp[n]);
    std::vector<int32_t> at(itp.begin(), itp.end() - 1), ah(ihp.begin(), ihp.end() - 1);
    for (int e = 0; e < t; ++e)
      for (int a = 0; a < 3; ++a)
        if (first(in.tri + 3 * (size_t)e, a)) itl[(size_t)at[in.tri[3 * (size_t)e + a]]++] = e;
    for (int e = 0; e < h; ++e)
      for (int a = 0; a < 4; ++a)
        if (first(in.hinge + 4 * (size_t)e, a)) ihl[(size_t)ah[in.hinge[4 * (size_t)e + a]]++] = e;
  }
  std::vector<int32_t> Ap(n + 1, 0), Ai;
  std::vector<double> Ax;
  {
    const int nth = host_threads(n, 8192);
    std::vector<std::vector<int32_t>> tAi(nth);
    std::vector<std::vector<double>> tAx(nth);
    parallel_ranges(n, nth, [&](int64_t rb, int64_t re, int k) {
      std::vector<int32_t> cols;
      std::vector<double> val;
      std::vector<char> kept;
      std::vector<int32_t>& oi = tAi[k];
      std::vector<double>& ox = tAx[k];
      for (int r = (int)rb; r < (int)re; ++r) {
        cols.assign(1, r);
        for (int p = itp[r]; p < itp[r + 1]; ++p)
          for (int a = 0; a < 3; ++a) cols.push_back((int32_t)in.tri[3 * (size_t)itl[p] + a]);
        for (int p = ihp[r]; p < ihp[r + 1]; ++p)
          for (int a = 0; a < 4; ++a) cols.push_back((int32_t)in.hinge[4 * (size_t)ihl[p] + a]);
        std::sort(cols.begin(), cols.end());
        cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
        val.assign(cols.size(), 0.0);
        kept.assign(cols.size(), 0);
        auto add = [&](int32_t c, double v) {
          const size_t q = (size_t)(std::lower_bound(cols.begin(), cols.end(), c) - cols.begin());
          val[q] += v;
          kept[q] = 1;
        };
        add(r, mass[r] * inv_dt2);
        for (int p = itp[r]; p < itp[r + 1]; ++p) {
          const int e = itl[p];
          const double* d = &dm[4 * (size_t)e];
          const double g[3][2] = {{-(d[0] + d[1]), -(d[2] + d[3])}, {d[0], d[2]}, {d[1], d[3]}};
          for (int a = 0; a < 3; ++a) {
            if ((int)in.tri[3 * (size_t)e + a] != r) continue;
            for (int b = 0; b < 3; ++b) {
              const double v = g[a][0] * g[b][0] + g[a][1] * g[b][1];
              if (std::fabs(v) > 1e-12) add((int32_t)in.tri[3 * (size_t)e + b], tw[e] * v);
            }
          }
        }
        for (int p = ihp[r]; p < ihp[r + 1]; ++p) {
          const int e = ihl[p];
          const double* st = &hst[4 * (size_t)e];
          for (int a = 0; a < 4; ++a) {
            if ((int)in.hinge[4 * (size_t)e + a] != r) continue;
            for (int b = 0; b < 4; ++b) {
              const double v = hw[e] * st[a] * st[b];
              if (std::fabs(v) > 1e-12) add((int32_t)in.hinge[4 * (size_t)e + b], v);
            }
          }
        }
        int cnt = 0;
        for (size_t q = 0; q < cols.size(); ++q)
          if (kept[q]) { oi.push_back(cols[q]); ox.push_back(val[q]); ++cnt; }
        Ap[r + 1] = cnt;
      }
    });
    for (int i = 0; i < n; ++i) Ap[i + 1] += Ap[i];
    Ai.reserve((size_t)Ap[n]);
    Ax.reserve((size_t)Ap[n]);
    for (int k = 0; k < nth; ++k) {  // the ranges are contiguous and in row order
      Ai.insert(Ai.end(), tAi[k].begin(), tAi[k].end());
      Ax.insert(Ax.end(), tAx[k].begin(), tAx[k].end());
    }
  }

  // Nested dissection, one tree per connected component (a forest for
  // several garments; their blocks never couple, so each tree has its own
  // root at depth 0 and the levels of all trees run in the same launches).
  // The components' trees are built on several threads (largest first, each
  // thread taking the next), then concatenated in component order, as one
  // builder would have appended them.
  std::vector<int> nd_mark(n, 0);
  NdBuilder nb{Ap, Ai, X.data(), Y.data(), Z.data(), std::max(1, in.leaf_size), nd_mark.data()};
  std::vector<int> roots;
  {
    std::vector<int> comp(n, -1), queue;
    std::vector<std::vector<int>> members;
    for (int s0 = 0; s0 < n; ++s0) {
      if (comp[s0] >= 0) continue;
      const int id = (int)members.size();
      members.emplace_back();
      queue.assign(1, s0);
      comp[s0] = id;
      for (size_t q = 0; q < queue.size(); ++q) {
        const int v = queue[q];
        members[id].push_back(v);
        for (int p = Ap[v]; p < Ap[v + 1]; ++p)
          if (comp[Ai[p]] < 0) { comp[Ai[p]] = id; queue.push_back(Ai[p]); }
      }
    }
    const int nc = (int)members.size();
    std::vector<int> by_size(nc);
    std::iota(by_size.begin(), by_size.end(), 0);
    std::stable_sort(by_size.begin(), by_size.end(),
                     [&](int x, int y) { return members[x].size() > members[y].size(); });
    std::vector<NdBuilder> sub(nc, NdBuilder{Ap, Ai, X.data(), Y.data(), Z.data(), nb.leaf, nd_mark.data()});
    std::vector<int> sub_root(nc, -1);
    std::atomic<int> next{0};
    const int nth = host_threads(nc, 1);
    parallel_ranges(nth, nth, [&](int64_t, int64_t, int) {
      for (int q = next++; q < nc; q = next++) {
        const int c = by_size[q];
        std::sort(members[c].begin(), members[c].end());
        sub_root[c] = sub[c].build(members[c]);
      }
    });
    for (int c = 0; c < nc; ++c) {
      const int off = (int)nb.nodes.size();
      for (NdNode& nd : sub[c].nodes) {
        if (nd.left >= 0) { nd.left += off; nd.right += off; }
        nb.nodes.push_back(std::move(nd));
      }
      roots.push_back(off + sub_root[c]);
    }
  }
  // Fine dissection tree: parents, depths, sizes, and each subtree's
  // boundary size r (vertices outside the subtree adjacent to it, all of
  // them in ancestor separators).
  const int nf = (int)nb.nodes.size();
  std::vector<int> fpar(nf, -1), fdep(nf, 0), fw(nf), flo(nf), fhi(nf), fr(nf, 0);
  std::vector<int> fperm(n), fiperm(n);
  {
    int col = 0;
    for (int x = 0; x < nf; ++x) {  // nodes are in postorder
      const NdNode& nd = nb.nodes[x];
      fw[x] = (int)nd.verts.size();
      flo[x] = nd.left >= 0 ? flo[nd.left] : col;
      for (int v : nd.verts) { fperm[col] = v; fiperm[v] = col; ++col; }
      fhi[x] = col;
      if (nd.left >= 0) { fpar[nd.left] = x; fpar[nd.right] = x; }
    }
  }
  int dmax = 0;
  for (int x = nf - 1; x >= 0; --x) { fdep[x] = fpar[x] < 0 ? 0 : fdep[fpar[x]] + 1; dmax = std::max(dmax, fdep[x]); }
  {
    std::vector<int> mk(n, -1);
    for (int x = 0; x < nf; ++x)
      for (int c = flo[x]; c < fhi[x]; ++c) {
        const int v = fperm[c];
        for (int p = Ap[v]; p < Ap[v + 1]; ++p) {
          const int u = Ai[p];
          if (fiperm[u] >= fhi[x] && mk[u] != x) { mk[u] = x; fr[x]++; }
        }
      }
  }
  // W[x][b]: columns of x's subtree nodes with depth <= b.
  const int L = dmax + 1;
  std::vector<int64_t> Wb((size_t)nf * L, 0);
  for (int x = 0; x < nf; ++x) {
    const NdNode& nd = nb.nodes[x];
    for (int b = fdep[x]; b < L; ++b) {
      int64_t w = fw[x];
      if (nd.left >= 0 && b > fdep[x]) w += Wb[(size_t)nd.left * L + b] + Wb[(size_t)nd.right * L + b];
      Wb[(size_t)x * L + b] = w;
    }
  }
  // Bands by dynamic programming over levels.
  const int64_t front_cap = (int64_t)(150 * 1024) / (3 * std::max(4, in.value_bytes));
  const int maxb = std::max(1, in.max_band_levels);
  std::vector<std::vector<int>> at_depth(L);
  for (int x = 0; x < nf; ++x) at_depth[fdep[x]].push_back(x);
  // A band costs two launches (forward, backward) over its blocks; the root
  // band costs one launch over its dense w x w inverse.
  auto band_cost = [&](int a, int b) -> double {
    double entries = 0;
    for (int x : at_depth[a]) {
      const double W = (double)Wb[(size_t)x * L + b];
      if (W + fr[x] > front_cap) return 1e300;
      entries += a == 0 ? W * W : W * (W + 1) / 2 + (double)fr[x] * W;
    }
    return (a == 0 ? 1.0 : 2.0) * (in.band_launch_us * 1e-6 + entries * in.value_bytes / (in.band_gbps * 1e9));
  };
  std::vector<double> best(L + 1, 1e300);
  std::vector<int> from(L + 1, 0);
  best[0] = 0;
  for (int e = 1; e <= L; ++e)
    for (int a = std::max(0, e - maxb); a < e; ++a) {
      const double c = best[a] + band_cost(a, e - 1);
      if (c < best[e]) { best[e] = c; from[e] = a; }
    }
  if (!(best[L] < 1e299)) { S.status = VH_E_UNSUPPORTED; S.err = "a separator front exceeds the LDS budget"; return false; }
  std::vector<int> bands;
  for (int e = L; e > 0; e = from[e]) bands.push_back(from[e]);
  std::reverse(bands.begin(), bands.end());
  std::vector<int> band_of_depth(L, 0), band_end(bands.size());
  for (size_t k = 0; k < bands.size(); ++k) {
    band_end[k] = k + 1 < bands.size() ? bands[k + 1] - 1 : L - 1;
    for (int d = bands[k]; d <= band_end[k]; ++d) band_of_depth[d] = (int)k;
  }
  S.band_start = bands;
  S.nd_levels = L;

  // Supernodes: each node at a band's first level with its descendants inside
  // the band. Columns in postorder of supernodes; within one, the members'
  // columns in dissection postorder.
  S.perm.assign(n, 0); S.iperm.assign(n, 0);
  S.sn.clear();
  S.children.clear();
  std::vector<int> sub_lo;
  {
    int col = 0;
    std::vector<std::vector<int>> kids;  // children per supernode, filled as they complete
    std::function<int(int)> visit = [&](int root) -> int {
      const int bend = band_end[band_of_depth[fdep[root]]];
      std::vector<int> members, stack{root};
      while (!stack.empty()) {
        const int x = stack.back(); stack.pop_back();
        members.push_back(x);
        const NdNode& nd = nb.nodes[x];
        if (nd.left >= 0 && fdep[x] < bend) { stack.push_back(nd.left); stack.push_back(nd.right); }
      }
      std::sort(members.begin(), members.end());  // fine node ids are postorder
      std::vector<int> ch;
      const int lo = col;
      for (int x : members) {
        const NdNode& nd = nb.nodes[x];
        if (nd.left >= 0 && fdep[x] == bend) { ch.push_back(visit(nd.left)); ch.push_back(visit(nd.right)); }
      }
      Supernode sn;
      sn.col0 = col;
      for (int x : members)
        for (int v : nb.nodes[x].verts) { S.perm[col] = v; S.iperm[v] = col; ++col; }
      sn.w = col - sn.col0;
      sn.depth = band_of_depth[fdep[root]];
      const int id = (int)S.sn.size();
      S.sn.push_back(sn);
      sub_lo.push_back(lo);
      kids.push_back(ch);
      for (int c : ch) S.sn[c].parent = id;
      return id;
    };
    for (int r : roots) visit(r);
    for (size_t s = 0; s < S.sn.size(); ++s) {
      S.sn[s].child_off = (int)S.children.size();
      S.sn[s].nchild = (int)kids[s].size();
      S.children.insert(S.children.end(), kids[s].begin(), kids[s].end());
    }
  }
  const int ns = (int)S.sn.size();
  S.max_depth = (int)bands.size() - 1;

  // Permuted matrix (rows on several threads).
  S.Ap.assign(n + 1, 0);
  S.Ai.resize(Ai.size()); S.Ax.resize(Ax.size());
  for (int i = 0; i < n; ++i) S.Ap[i + 1] = S.Ap[i] + (Ap[S.perm[i] + 1] - Ap[S.perm[i]]);
  parallel_ranges(n, host_threads(n, 8192), [&](int64_t rb, int64_t re, int) {
    std::vector<std::pair<int32_t, double>> row;
    for (int i = (int)rb; i < (int)re; ++i) {
      const int o = S.perm[i];
      row.clear();
      for (int p = Ap[o]; p < Ap[o + 1]; ++p) row.push_back({S.iperm[Ai[p]], Ax[p]});
      std::sort(row.begin(), row.end());
      int64_t w = S.Ap[i];
      for (auto& e : row) { S.Ai[w] = e.first; S.Ax[w] = e.second; ++w; }
    }
  });

  // Row structure I_P = vertices outside P's subtree adjacent to it
  // (supernodes on several threads, each with its own marks).
  S.rowidx.clear();
  {
    std::vector<std::vector<int>> rows_of(ns);
    const int nth = host_threads(ns, 64);
    parallel_ranges(ns, nth, [&](int64_t sb, int64_t se, int) {
      std::vector<int> mark(n, -1);
      for (int s = (int)sb; s < (int)se; ++s) {
        const int lo = sub_lo[s], hi = S.sn[s].col0 + S.sn[s].w;
        std::vector<int>& rows = rows_of[s];
        for (int i = lo; i < hi; ++i)
          for (int p = S.Ap[i]; p < S.Ap[i + 1]; ++p) {
            const int j = S.Ai[p];
            if (j >= hi && mark[j] != s) { mark[j] = s; rows.push_back(j); }
          }
        std::sort(rows.begin(), rows.end());
      }
    });
    for (int s = 0; s < ns; ++s) {
      Supernode& x = S.sn[s];
      x.r = (int)rows_of[s].size();
      x.rows_off = (int)S.rowidx.size();
      S.rowidx.insert(S.rowidx.end(), rows_of[s].begin(), rows_of[s].end());
    }
  }
  // Offsets for M, update vectors and the extend-add maps.
  int64_t moff = 0;
  int uoff = 0;
  S.max_front = 0;
  for (int s = 0; s < ns; ++s) {
    Supernode& x = S.sn[s];
    x.m_off = moff;
    moff += (int64_t)x.w * (x.w + 1) / 2 + (int64_t)x.r * x.w;
    x.u_off = uoff;
    uoff += x.r;
    S.max_front = std::max(S.max_front, x.w + x.r);
  }
  S.m_total = moff;
  S.nnz_factor = moff;
  S.u_total = uoff;
  S.emap.assign((size_t)uoff, 0);
  std::vector<int> loc(n, -1);
  for (int s = 0; s < ns; ++s) {
    const Supernode& x = S.sn[s];
    if (x.parent < 0) continue;
    const Supernode& p = S.sn[x.parent];
    for (int q = 0; q < p.r; ++q) loc[S.rowidx[p.rows_off + q]] = p.w + q;
    for (int q = 0; q < x.r; ++q) {
      const int v = S.rowidx[x.rows_off + q];
      int pos;
      if (v >= p.col0 && v < p.col0 + p.w) pos = v - p.col0;
      else pos = loc[v];
      if (pos < 0) { S.status = VH_E_INTERNAL; S.err = "extend-add map: row not in parent front"; return false; }
      S.emap[x.u_off + q] = pos;
    }
    for (int q = 0; q < p.r; ++q) loc[S.rowidx[p.rows_off + q]] = -1;
    S.sn[s].map_off = x.u_off;
  }

  // Forward push slots (inverse of the extend-add maps) and backward push
  // lists.
  {
    S.uslot.assign((size_t)S.u_total, 0);
    std::vector<std::vector<int32_t>> lists;
    int64_t slot = 0;
    for (int s = 0; s < ns; ++s) {
      Supernode& x = S.sn[s];
      const int m = x.w + x.r;
      lists.assign(m, {});
      for (int c = 0; c < x.nchild; ++c) {
        const Supernode& y = S.sn[S.children[x.child_off + c]];
        for (int q = 0; q < y.r; ++q) lists[S.emap[y.u_off + q]].push_back(y.u_off + q);
      }
      int k = 0;
      for (int j = 0; j < m; ++j) k = std::max(k, (int)lists[j].size());
      // fm_off and the slot numbers below are 32-bit on the device.
      if (slot + (int64_t)m * k > INT32_MAX) { S.status = VH_E_UNSUPPORTED; S.err = "front slots exceed 2^31"; return false; }
      x.fm_off = (int)slot;
      x.fslot_k = k;
      for (int j = 0; j < m; ++j)
        for (size_t q = 0; q < lists[j].size(); ++q) S.uslot[lists[j][q]] = (int32_t)(slot + (int64_t)j * k + q);
      slot += (int64_t)m * k;
    }
    S.fslot_total = slot;
    std::vector<int32_t> cnt(n + 1, 0);
    for (int s = 0; s < ns; ++s)
      for (int q = 0; q < S.sn[s].r; ++q) cnt[S.rowidx[S.sn[s].rows_off + q] + 1]++;
    for (int i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
    S.dpush_ptr = cnt;
    S.dpush_idx.assign((size_t)cnt[n], 0);
    std::vector<int32_t> cur(cnt.begin(), cnt.end() - 1);
    for (int s = 0; s < ns; ++s)
      for (int q = 0; q < S.sn[s].r; ++q) S.dpush_idx[cur[S.rowidx[S.sn[s].rows_off + q]]++] = S.sn[s].u_off + q;
  }

  // Root inverses: every root (no off-diagonal rows) gets its dense w x w
  // inverse at Sroot offset s_off. The numeric factorization itself runs on
  // the device (vh_factor.hpp) from this symbolic structure.
  S.sroot_total = 0;
  for (int s = 0; s < ns; ++s) {
    Supernode& x = S.sn[s];
    if (x.parent >= 0) continue;
    if (x.r != 0) { S.status = VH_E_INTERNAL; S.err = "root supernode has off-diagonal rows"; return false; }
    x.s_off = S.sroot_total;
    S.sroot_total += (int64_t)x.w * x.w;
  }

  // Elements with permuted vertex labels, sorted by their smallest label for
  // locality of the position gathers.
  {
    std::vector<int> ord(t);
    std::iota(ord.begin(), ord.end(), 0);
    auto tmin = [&](int e) {
      return std::min(S.iperm[in.tri[3 * e]], std::min(S.iperm[in.tri[3 * e + 1]], S.iperm[in.tri[3 * e + 2]]));
    };
    stable_order_by_key(ord, n, tmin);
    S.tri_idx.resize(3 * (size_t)t); S.tri_dm.resize(4 * (size_t)t); S.tri_w.resize(t);
    S.tri_orig = ord;
    for (int k = 0; k < t; ++k) {
      const int e = ord[k];
      for (int c = 0; c < 3; ++c) S.tri_idx[3 * k + c] = S.iperm[in.tri[3 * e + c]];
      for (int c = 0; c < 4; ++c) S.tri_dm[4 * k + c] = dm[4 * (size_t)e + c];
      S.tri_w[k] = tw[e];
    }
    std::vector<int> hord(h);
    std::iota(hord.begin(), hord.end(), 0);
    auto hmin = [&](int e) {
      int m = S.iperm[in.hinge[4 * e]];
      for (int c = 1; c < 4; ++c) m = std::min(m, S.iperm[in.hinge[4 * e + c]]);
      return m;
    };
    stable_order_by_key(hord, n, hmin);
    S.hinge_idx.resize(4 * (size_t)h); S.hinge_rest.resize(h); S.hinge_w.resize(h);
    S.hinge_st.assign(ds ? 4 * (size_t)h : 0, 0.0);
    for (int k = 0; k < h; ++k) {
      const int e = hord[k];
      for (int c = 0; c < 4; ++c) S.hinge_idx[4 * k + c] = S.iperm[in.hinge[4 * e + c]];
      S.hinge_rest[k] = hrest[e];
      S.hinge_w[k] = hw[e];
      if (ds)
        for (int c = 0; c < 4; ++c) S.hinge_st[4 * k + c] = hst[4 * (size_t)e + c];
    }
  }
  S.mass.resize(n); S.pinned.resize(n);
  for (int i = 0; i < n; ++i) { S.mass[i] = mass[S.perm[i]]; S.pinned[i] = in.pinned[S.perm[i]] ? 1 : 0; }
  // RHS slots, one per (local-step chunk, vertex it touches): per chunk, its
  // corners sorted by (vertex, place in the block) (chunks on several
  // threads); then each vertex's slots numbered in chunk order.
  {
    constexpr int kC = 256;
    const int tb = (t + kC - 1) / kC, hb = (h + kC - 1) / kC, nch = tb + hb;
    S.lc_pos.assign(3 * (size_t)t + 4 * (size_t)h, 0);
    std::vector<std::vector<int32_t>> lvv(nch);
    std::vector<std::vector<uint16_t>> lve(nch);
    parallel_ranges(nch, host_threads(nch, 64), [&](int64_t cb, int64_t ce, int) {
      std::vector<int64_t> key;
      for (int64_t c = cb; c < ce; ++c) {
        const bool tri = c < tb;
        const int K = tri ? 3 : 4;
        const int64_t e0 = (tri ? c : c - tb) * kC;
        const int64_t ne = std::min<int64_t>(kC, (tri ? t : h) - e0);
        const int32_t* idx = tri ? S.tri_idx.data() : S.hinge_idx.data();
        key.clear();
        for (int64_t le = 0; le < ne; ++le)
          for (int k = 0; k < K; ++k) key.push_back(((int64_t)idx[(size_t)(e0 + le) * K + k] << 10) | (le * K + k));
        std::sort(key.begin(), key.end());
        const size_t base = tri ? 3 * (size_t)e0 : 3 * (size_t)t + 4 * (size_t)e0;
        std::vector<int32_t>& V = lvv[c];
        std::vector<uint16_t>& E = lve[c];
        for (size_t q = 0; q < key.size(); ++q) {
          const int32_t v = (int32_t)(key[q] >> 10);
          if (V.empty() || V.back() != v) {
            V.push_back(v);
            E.push_back(0);
          }
          S.lc_pos[base + q] = (uint16_t)(key[q] & 1023);
          E.back() = (uint16_t)(q + 1);
        }
      }
    });
    S.lv_ptr.assign((size_t)nch + 1, 0);
    for (int c = 0; c < nch; ++c) S.lv_ptr[c + 1] = S.lv_ptr[c] + (int32_t)lvv[c].size();
    std::vector<int32_t> cnt((size_t)n + 1, 0);
    for (int c = 0; c < nch; ++c)
      for (int32_t v : lvv[c]) cnt[(size_t)v + 1]++;
    for (int i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
    S.inc_ptr = cnt;
    std::vector<int32_t> cur(cnt.begin(), cnt.end() - 1);
    S.lv_slot.resize((size_t)S.lv_ptr[nch]);
    S.lv_end.resize((size_t)S.lv_ptr[nch]);
    for (int c = 0; c < nch; ++c) {
      const size_t b = (size_t)S.lv_ptr[c];
      for (size_t j = 0; j < lvv[c].size(); ++j) {
        S.lv_slot[b + j] = cur[lvv[c][j]]++;
        S.lv_end[b + j] = lve[c][j];
      }
    }
  }
  S.upper_cols.clear();
  S.band_front.assign(S.max_depth + 1, 0);
  for (int s = 0; s < ns; ++s) {
    const Supernode& x = S.sn[s];
    S.band_front[x.depth] = std::max(S.band_front[x.depth], x.w + x.r);
    if (x.depth < S.max_depth)
      for (int j = 0; j < x.w; ++j) S.upper_cols.push_back(x.col0 + j);
  }
  std::sort(S.upper_cols.begin(), S.upper_cols.end());
  if (S.m_total >= (int64_t)1 << 31) { S.status = VH_E_UNSUPPORTED; S.err = "solve blocks exceed 2^31 entries"; return false; }
  if (S.sroot_total >= (int64_t)1 << 31) {
    S.status = VH_E_UNSUPPORTED; S.err = "root inverses exceed 2^31 entries"; return false;
  }

  // Task lists. A task is a run of rows of one stored form; g lanes share a
  // row (g ~ row length / 8, 4..64, so a row is usually 8 loads per lane)
  // and each lane group takes up to pr = 2 rows, whose loads are issued
  // together before the task's front is assembled (the kernels'
  // pref_rows; groups of 64 lanes prefetch one longer row). A task holds
  // about task_entries block entries: at least one row per group, at most pr.
  const int kChunk = std::max(256, in.task_entries);
  const int pr = 2;
  auto pow2ceil = [](int v) { int p = 1; while (p < v) p <<= 1; return p; };
  auto lanes_for = [&](int len) { return std::min(64, std::max(4, pow2ceil((std::max(1, len) + 7) / 8))); };
  auto rows_for = [&](int len, int g) {
    const int ngrp = 256 / g;
    return std::max(ngrp, std::min(pr * ngrp, kChunk / std::max(1, len)));
  };
  S.fwd_tasks.clear(); S.bwd_tasks.clear();
  S.fwd_level_off.assign(S.max_depth + 2, 0);
  S.bwd_level_off.assign(S.max_depth + 2, 0);
  std::vector<int> nfc(ns, 0), nbc(ns, 0);
  for (int lvl = 0; lvl <= S.max_depth; ++lvl) {  // forward runs the deepest level first
    const int depth = S.max_depth - lvl;
    S.fwd_level_off[lvl] = (int)S.fwd_tasks.size();
    for (int s = 0; s < ns; ++s) {
      const Supernode& x = S.sn[s];
      if (x.depth != depth) continue;  // w == 0 nodes still pass child updates up
      const int g = lanes_for(x.w);
      const int m = x.w + x.r;
      const int rows_per = rows_for(x.w, g);
      const int kind = x.parent < 0 ? 2 : 0;
      for (int b = 0; b < m; b += rows_per) {
        S.fwd_tasks.push_back({s, b, std::min(m, b + rows_per), g, kind});
        nfc[s]++;
      }
    }
  }
  S.fwd_level_off[S.max_depth + 1] = (int)S.fwd_tasks.size();
  for (int lvl = 0; lvl <= S.max_depth; ++lvl) {  // backward runs the root level first
    S.bwd_level_off[lvl] = (int)S.bwd_tasks.size();
    for (int s = 0; s < ns; ++s) {
      const Supernode& x = S.sn[s];
      if (x.depth != lvl || x.parent < 0) continue;  // the root is finished by its root tasks
      const int m = x.w + x.r;
      const int g = lanes_for(m);
      const int rows_per = rows_for(m, g);
      for (int b = 0; b < x.w; b += rows_per) {
        S.bwd_tasks.push_back({s, b, std::min(x.w, b + rows_per), g, 1});
        nbc[s]++;
      }
      if (x.w == 0) {  // empty separator: one empty task keeps the children's wait chain intact
        S.bwd_tasks.push_back({s, 0, 0, g, 1});
        nbc[s]++;
      }
    }
  }
  S.bwd_level_off[S.max_depth + 1] = (int)S.bwd_tasks.size();
  S.ptasks = S.fwd_tasks;
  S.ptasks.insert(S.ptasks.end(), S.bwd_tasks.begin(), S.bwd_tasks.end());
  S.need_f.assign(ns, 0); S.need_b.assign(ns, 0); S.parent.assign(ns, -1);
  for (int s = 0; s < ns; ++s) {
    S.parent[s] = S.sn[s].parent;
    for (int c = 0; c < S.sn[s].nchild; ++c) S.need_f[s] += nfc[S.children[S.sn[s].child_off + c]];
    S.need_b[s] = S.sn[s].parent >= 0 ? nbc[S.sn[s].parent] : nfc[s];
  }
  const std::string why = device_index_overflow(S.m_total, S.sroot_total, S.fslot_total,
                                                3 * ((int64_t)3 * t + (int64_t)4 * h));
  if (!why.empty()) { S.status = VH_E_UNSUPPORTED; S.err = why; return false; }
  S.status = VH_OK;
  return true;
}

std::string device_index_overflow(int64_t m_total, int64_t sroot_entries, int64_t front_slots, int64_t rhs_values) {
  if (m_total > INT32_MAX) return "solve blocks exceed 2^31 entries (split the batch)";
  if (sroot_entries > INT32_MAX) return "root inverses exceed 2^31 entries (split the batch)";
  if (front_slots > INT32_MAX) return "front slots exceed 2^31";
  if (rhs_values > INT32_MAX) return "RHS slot values exceed 2^31 (split the batch)";
  return std::string();
}

}  // namespace vh
