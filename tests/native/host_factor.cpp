// Test-only CPU copy of the supernodal factorization (tests/ infrastructure).
// The product builds the solve blocks on the device (vistio_amd/csrc/vh_factor.hpp)
// from the symbolic set-up of vistio_amd/csrc/vh_setup.cpp. This file computes the
// same blocks on the host, sequentially in double, from that same symbolic
// set-up: a multifrontal dense LL^T per supernode (extend-add of the
// children's update matrices, dense Cholesky, L_IP = F_IP L_PP^{-T}, Schur
// update) and the partitioned-inverse blocks M_P = [L_PP^{-1}; L_IP L_PP^{-1}],
// M_P^T and the roots' dense inverses. The schedule replay
// (setup_emul.cpp) solves with them, and the GPU tests compare the device's
// blocks with them.
#include <cmath>
#include <cstdint>
#include <vector>

#include "vh_setup.hpp"

namespace {

// Dense lower Cholesky in place on the leading w x w block of F (ld = m).
bool dense_chol(double* F, int m, int w) {
  for (int j = 0; j < w; ++j) {
    double* Fj = F + (size_t)j * m;
    double d = Fj[j];
    for (int k = 0; k < j; ++k) d -= Fj[k] * Fj[k];
    if (!(d > 0.0)) return false;
    const double ljj = std::sqrt(d);
    Fj[j] = ljj;
    const double inv = 1.0 / ljj;
    for (int i = j + 1; i < w; ++i) {
      double* Fi = F + (size_t)i * m;
      double s = Fi[j];
      for (int k = 0; k < j; ++k) s -= Fi[k] * Fj[k];
      Fi[j] = s * inv;
    }
  }
  return true;
}

}  // namespace

// Fills M and Mt (m_total entries each) and Sroot (sroot_total entries).
// Returns false when a pivot is not positive (the matrix is not SPD) or the
// structure is inconsistent.
bool host_factor(const vh::Setup& S, std::vector<double>& M, std::vector<double>& Mt, std::vector<double>& Sroot) {
  const int n = S.n;
  const int ns = (int)S.sn.size();
  M.assign((size_t)S.m_total, 0.0);
  Mt.assign((size_t)S.m_total, 0.0);
  Sroot.assign((size_t)S.sroot_total, 0.0);
  std::vector<int> loc(n, -1);
  std::vector<std::vector<double>> upd(ns);
  std::vector<double> F, Linv;
  for (int s = 0; s < ns; ++s) {
    const vh::Supernode& x = S.sn[s];
    const int w = x.w, r = x.r, m = w + r;
    F.assign((size_t)m * m, 0.0);
    for (int q = 0; q < r; ++q) loc[S.rowidx[x.rows_off + q]] = w + q;
    for (int jj = 0; jj < w; ++jj) {
      const int J = x.col0 + jj;
      for (int p = S.Ap[J]; p < S.Ap[J + 1]; ++p) {
        const int i = S.Ai[p];
        if (i < J) continue;
        const int pos = i < x.col0 + w ? i - x.col0 : loc[i];
        if (pos < 0) return false;
        F[(size_t)pos * m + jj] += S.Ax[p];
      }
    }
    for (int q = 0; q < r; ++q) loc[S.rowidx[x.rows_off + q]] = -1;
    for (int c = 0; c < x.nchild; ++c) {
      const int ch = S.children[x.child_off + c];
      const vh::Supernode& y = S.sn[ch];
      const std::vector<double>& U = upd[ch];
      const int* mp = &S.emap[y.u_off];
      for (int a = 0; a < y.r; ++a)
        for (int b = 0; b <= a; ++b) F[(size_t)mp[a] * m + mp[b]] += U[(size_t)a * y.r + b];
      upd[ch].clear(); upd[ch].shrink_to_fit();
    }
    if (!dense_chol(F.data(), m, w)) {
      return false;
    }
    // L_IP = F_IP L_PP^{-T}
    for (int q = 0; q < r; ++q) {
      double* Fq = &F[(size_t)(w + q) * m];
      for (int j = 0; j < w; ++j) {
        const double* Lj = &F[(size_t)j * m];
        double v = Fq[j];
        for (int k = 0; k < j; ++k) v -= Fq[k] * Lj[k];
        Fq[j] = v / Lj[j];
      }
    }
    // Update matrix U = F_II - L_IP L_IP^T (lower)
    if (x.parent >= 0 && r > 0) {
      std::vector<double>& U = upd[s];
      U.assign((size_t)r * r, 0.0);
      for (int a = 0; a < r; ++a) {
        const double* La = &F[(size_t)(w + a) * m];
        for (int b = 0; b <= a; ++b) {
          const double* Lb = &F[(size_t)(w + b) * m];
          double v = F[(size_t)(w + a) * m + w + b];
          for (int k = 0; k < w; ++k) v -= La[k] * Lb[k];
          U[(size_t)a * r + b] = v;
        }
      }
    }
    // Linv = L_PP^{-1} (lower), column by column.
    Linv.assign((size_t)w * w, 0.0);
    for (int j = 0; j < w; ++j) {
      Linv[(size_t)j * w + j] = 1.0 / F[(size_t)j * m + j];
      for (int i = j + 1; i < w; ++i) {
        const double* Li = &F[(size_t)i * m];
        double v = 0;
        for (int k = j; k < i; ++k) v -= Li[k] * Linv[(size_t)k * w + j];
        Linv[(size_t)i * w + j] = v / Li[i];
      }
    }
    double* Mp = &M[(size_t)x.m_off];
    for (int i = 0; i < w; ++i)
      for (int j = 0; j <= i; ++j) Mp[(size_t)i * (i + 1) / 2 + j] = Linv[(size_t)i * w + j];
    double* Wp = Mp + (size_t)w * (w + 1) / 2;
    for (int q = 0; q < r; ++q) {
      const double* Lq = &F[(size_t)(w + q) * m];
      for (int j = 0; j < w; ++j) {
        double v = 0;
        for (int k = j; k < w; ++k) v += Lq[k] * Linv[(size_t)k * w + j];
        Wp[(size_t)q * w + j] = v;
      }
    }
    // Transposed copy: row j holds column j of M_P, i.e. Linv[j..w-1][j]
    // followed by W[0..r-1][j]; entry i multiplies front position j + i.
    double* Tp = &Mt[(size_t)x.m_off];
    for (int j = 0; j < w; ++j) {
      double* Tj = Tp + (size_t)j * m - (size_t)j * (j - 1) / 2;
      for (int k = j; k < w; ++k) Tj[k - j] = Mp[(size_t)k * (k + 1) / 2 + j];
      for (int q = 0; q < r; ++q) Tj[w - j + q] = Wp[(size_t)q * w + j];
    }
    if (x.parent < 0) {  // root: full inverse of L_RR L_RR^T = Linv^T Linv
      if (r != 0) return false;
      double* Sp = Sroot.data() + x.s_off;
      for (int k = 0; k < w; ++k) {
        const double* Lk = &Linv[(size_t)k * w];
        for (int i = 0; i <= k; ++i) {
          const double a = Lk[i];
          double* Si = Sp + (size_t)i * w;
          for (int j = i; j <= k; ++j) Si[j] += a * Lk[j];
        }
      }
      for (int i = 0; i < w; ++i)
        for (int j = 0; j < i; ++j) Sp[(size_t)i * w + j] = Sp[(size_t)j * w + i];
    }
  }

  return true;
}
