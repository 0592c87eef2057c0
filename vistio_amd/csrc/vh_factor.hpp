// Device factorization of the constant PD system matrix (FaerSolver::factorize,
// crates/vistio-math/src/faer_solver.rs:67-93), in FP64, from the host's
// symbolic set-up (vh_setup.hpp: supernodes, their row structures and the
// extend-add maps), straight into the blocks the solve kernels apply.
//
// Multifrontal, one dissection level (band) at a time, deepest first. The
// front of supernode P is a dense (w + r) x (w + r) matrix padded to tiles
// of kFT x kFT: wp = w rounded up (padded pivots are identity columns, which
// factor trivially and couple to nothing) and rp = r rounded up (zero rows).
// Next to it lies E, mp x wp, which starts as [I; 0].
//
//   assemble   F = A's columns of P + the children's update matrices
//              (extend-add, one child after another, no atomics); E = [I; 0]
//   panel k    diag: L_kk = chol(F_kk) and Linv_kk = L_kk^{-1} in LDS
//              trsm: F_ik = F_ik Linv_kk^T (i > k),  E_kc = Linv_kk E_kc (c <= k)
//              update: F_ij -= F_ik F_jk^T (k < j <= i),  E_ic -= F_ik E_kc (i > k)
//
// The panels run a right-looking Cholesky of F's pivot columns and, in the
// same loop, the forward substitution L_F X = [I; 0] with L_F = [L_PP 0;
// L_IP I]. So afterwards F_II holds the update matrix for P's parent, and
// E = [L_PP^{-1}; -L_IP L_PP^{-1}]: the partitioned-inverse block
// M_P = [L_PP^{-1}; L_IP L_PP^{-1}] the solve applies, with no triangular
// inverse left to form. For a root, S = E^T E = (L L^T)^{-1} is its dense
// inverse. Every tile product runs on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64): a 64-lane wave computes one 32 x 32 tile as four
// 16 x 16 blocks, eight 4-deep steps each.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace vhf {

constexpr int kFT = 32;  // tile edge

struct FrontDev {
  long long f;    // offset of F (mp x mp, row-major) in its level's front buffer (doubles)
  long long e;    // offset of E (mp x wp, row-major) in its level's E buffer
  long long m_off, s_off;  // offsets of M_P / M_P^T and (roots) of the dense inverse
  int mp, wp, w, r;
  int col0, rows_off;      // columns col0 .. col0 + w - 1; rows rowidx[rows_off .. + r)
  int child_off, nchild;   // children: S.children[child_off ..]
  int u_off;               // this node's rows in its parent's front: emap[u_off .. + r)
  int lpos;                // index within its level (its Linv_kk scratch tile)
};

// A tile task: (supernode, tile row i, tile column j, kind).
enum TileKind { kFTrsm = 0, kETrsm = 1, kFUpd = 2, kEUpd = 3, kSRoot = 4 };

typedef double dbl4 __attribute__((ext_vector_type(4)));

// Padded front position of unpadded position p (w pivots, then rows).
__device__ __forceinline__ int padded(int p, int w, int wp) { return p < w ? p : wp + (p - w); }

// Position of global index i among the sorted rows rowidx[b .. b + r).
__device__ __forceinline__ int row_pos(const int* __restrict__ rowidx, int b, int r, int i) {
  int lo = 0, hi = r;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (rowidx[b + mid] < i) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// One workgroup per supernode of the level: A's entries of P's columns, the
// identity of the padded pivots and of E, then the children's update
// matrices (the lower triangle of each child's F_II) one child at a time.
// The level's buffers are zero on entry.
__global__ void __launch_bounds__(256) k_fac_assemble(const FrontDev* __restrict__ fr, const int* __restrict__ lev,
                                                      const int* __restrict__ Ap, const int* __restrict__ Ai,
                                                      const double* __restrict__ Ax, const int* __restrict__ rowidx,
                                                      const int* __restrict__ children, const int* __restrict__ emap,
                                                      double* __restrict__ Fd, double* __restrict__ Ed,
                                                      const double* __restrict__ Fc) {
  const FrontDev P = fr[lev[blockIdx.x]];
  double* F = Fd + P.f;
  double* E = Ed + P.e;
  const int tid = threadIdx.x;
  for (int p = tid; p < P.wp; p += blockDim.x) {
    E[(size_t)p * P.wp + p] = 1.0;
    if (p >= P.w) F[(size_t)p * P.mp + p] = 1.0;
  }
  for (int jj = tid; jj < P.w; jj += blockDim.x) {  // one column per thread: no two threads share an entry
    const int J = P.col0 + jj;
    for (int q = Ap[J]; q < Ap[J + 1]; ++q) {
      const int i = Ai[q];
      if (i < J) continue;
      const int pos = i < P.col0 + P.w ? i - P.col0 : P.wp + row_pos(rowidx, P.rows_off, P.r, i);
      F[(size_t)pos * P.mp + jj] += Ax[q];
    }
  }
  __syncthreads();
  for (int c = 0; c < P.nchild; ++c) {
    const FrontDev C = fr[children[P.child_off + c]];
    const double* U = Fc + C.f;
    const int* mp_ = emap + C.u_off;
    const int rc = C.r;
    for (int idx = tid; idx < rc * rc; idx += blockDim.x) {
      const int a = idx / rc, b = idx - a * rc;
      if (b > a) continue;
      const int pa = padded(mp_[a], P.w, P.wp), pb = padded(mp_[b], P.w, P.wp);
      F[(size_t)pa * P.mp + pb] += U[(size_t)(C.wp + a) * C.mp + (C.wp + b)];
    }
    __syncthreads();  // the next child may add to the same entries
  }
}

// Panel k, diagonal tile of each listed supernode: Cholesky of F_kk's lower
// triangle in LDS, then the inverse of that triangle (one column per
// thread). L_kk goes back into F, Linv_kk into the scratch tile D[lpos].
// A pivot that is not positive sets *fail and is replaced by 1.
__global__ void __launch_bounds__(256) k_fac_diag(const FrontDev* __restrict__ fr, const int* __restrict__ list, int k,
                                                  double* __restrict__ Fd, double* __restrict__ Dd, int* fail) {
  const FrontDev P = fr[list[blockIdx.x]];
  double* T = Fd + P.f + (size_t)k * kFT * P.mp + (size_t)k * kFT;
  __shared__ double L[kFT][kFT + 1];
  __shared__ double X[kFT][kFT + 1];
  const int tid = threadIdx.x;
  for (int idx = tid; idx < kFT * kFT; idx += blockDim.x) {
    const int i = idx / kFT, j = idx % kFT;
    L[i][j] = j <= i ? T[(size_t)i * P.mp + j] : 0.0;
    X[i][j] = 0.0;
  }
  __syncthreads();
  for (int j = 0; j < kFT; ++j) {
    if (tid == 0) {
      double d = L[j][j];
      if (!(d > 0.0)) {
        atomicOr(fail, 1);
        d = 1.0;
      }
      L[j][j] = sqrt(d);
    }
    __syncthreads();
    if (tid > j && tid < kFT) L[tid][j] /= L[j][j];
    __syncthreads();
    for (int idx = tid; idx < kFT * kFT; idx += blockDim.x) {
      const int i = idx / kFT, l = idx % kFT;
      if (l > j && l <= i) L[i][l] -= L[i][j] * L[l][j];
    }
    __syncthreads();
  }
  if (tid < kFT) {  // column tid of L^{-1}: forward substitution against e_tid
    const int c = tid;
    X[c][c] = 1.0 / L[c][c];
    for (int i = c + 1; i < kFT; ++i) {
      double s = 0.0;
      for (int q = c; q < i; ++q) s += L[i][q] * X[q][c];
      X[i][c] = -s / L[i][i];
    }
  }
  __syncthreads();
  double* D = Dd + (size_t)P.lpos * kFT * kFT;
  for (int idx = tid; idx < kFT * kFT; idx += blockDim.x) {
    const int i = idx / kFT, j = idx % kFT;
    if (j <= i) T[(size_t)i * P.mp + j] = L[i][j];
    D[idx] = X[i][j];
  }
}

// Loads a 32 x 32 tile at src (leading dimension ld) into s, transposed when
// tr (s[a][b] = src[b][a]). 64 lanes, one row of 32 doubles per half-wave.
__device__ __forceinline__ void load_tile(double (*s)[kFT + 1], const double* __restrict__ src, int ld, bool tr) {
  const int l = threadIdx.x;
  for (int idx = l; idx < kFT * kFT; idx += 64) {
    const int i = idx / kFT, j = idx % kFT;
    const double v = src[(size_t)i * ld + j];
    if (tr) s[j][i] = v; else s[i][j] = v;
  }
}

// acc[bm][bn] += sA[16 bm .. +16][0 .. 32) sB[0 .. 32)[16 bn .. +16] on the
// matrix cores. Lane l feeds A[l % 16][4 s + l / 16] and B[4 s + l / 16][l % 16]
// and holds D[4 q + l / 16][l % 16] in acc[..][..][q].
__device__ __forceinline__ void mfma_tile(dbl4 acc[2][2], const double (*sA)[kFT + 1], const double (*sB)[kFT + 1]) {
  const int l = threadIdx.x;
  const int r16 = l & 15, k4 = l >> 4;
#pragma unroll
  for (int s = 0; s < kFT / 4; ++s) {
    const double a0 = sA[r16][4 * s + k4], a1 = sA[16 + r16][4 * s + k4];
    const double b0 = sB[4 * s + k4][r16], b1 = sB[4 * s + k4][16 + r16];
    acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
    acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
  }
}

// One 64-lane workgroup per tile task (s, i, j, kind) of panel k:
//   kFTrsm  F_ik = F_ik Linv_kk^T           (i > k, j = k)
//   kETrsm  E_kc = Linv_kk E_kc             (i = k, j = c <= k)
//   kFUpd   F_ij -= F_ik F_jk^T             (k < j <= i)
//   kEUpd   E_ic -= F_ik E_kc               (i > k, j = c <= k)
//   kSRoot  S_ij = sum_{K >= max(i, j)} E_Ki^T E_Kj  (roots, written over F)
// Every task reads its inputs whole before writing its own tile, and no task
// of a launch writes a tile another task of the same launch reads.
__global__ void __launch_bounds__(64) k_fac_tiles(const FrontDev* __restrict__ fr, const int4* __restrict__ tasks, int k,
                                                  double* __restrict__ Fd, double* __restrict__ Ed,
                                                  const double* __restrict__ Dd) {
  const int4 t = tasks[blockIdx.x];
  const FrontDev P = fr[t.x];
  const int i = t.y, j = t.z, kind = t.w;
  __shared__ double sA[kFT][kFT + 1];
  __shared__ double sB[kFT][kFT + 1];
  double* F = Fd + P.f;
  double* E = Ed + P.e;
  const double* D = Dd + (size_t)P.lpos * kFT * kFT;
  auto Ft = [&](int a, int b) { return F + (size_t)a * kFT * P.mp + (size_t)b * kFT; };
  auto Et = [&](int a, int b) { return E + (size_t)a * kFT * P.wp + (size_t)b * kFT; };
  dbl4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = dbl4{0.0, 0.0, 0.0, 0.0};
  double* C;
  int ldc;
  bool sub = false;
  if (kind == kSRoot) {
    const int T = P.wp / kFT;
    for (int K = i > j ? i : j; K < T; ++K) {
      load_tile(sA, Et(K, i), P.wp, true);
      load_tile(sB, Et(K, j), P.wp, false);
      __syncthreads();
      mfma_tile(acc, sA, sB);
      __syncthreads();
    }
    C = Ft(i, j); ldc = P.mp;
  } else {
    if (kind == kFTrsm) {
      load_tile(sA, Ft(i, k), P.mp, false);
      load_tile(sB, D, kFT, true);
      C = Ft(i, k); ldc = P.mp;
    } else if (kind == kETrsm) {
      load_tile(sA, D, kFT, false);
      load_tile(sB, Et(k, j), P.wp, false);
      C = Et(k, j); ldc = P.wp;
    } else if (kind == kFUpd) {
      load_tile(sA, Ft(i, k), P.mp, false);
      load_tile(sB, Ft(j, k), P.mp, true);
      C = Ft(i, j); ldc = P.mp; sub = true;
    } else {
      load_tile(sA, Ft(i, k), P.mp, false);
      load_tile(sB, Et(k, j), P.wp, false);
      C = Et(i, j); ldc = P.wp; sub = true;
    }
    __syncthreads();
    mfma_tile(acc, sA, sB);
  }
  const int l = threadIdx.x;
  const int col = l & 15, row0 = l >> 4;
#pragma unroll
  for (int bm = 0; bm < 2; ++bm)
#pragma unroll
    for (int bn = 0; bn < 2; ++bn)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        double* p = C + (size_t)(16 * bm + 4 * q + row0) * ldc + 16 * bn + col;
        *p = sub ? *p - acc[bm][bn][q] : acc[bm][bn][q];
      }
}

// The solve blocks of each supernode of the level, in the compute type T:
// M_P row-major (rows 0 .. w-1 of L_PP^{-1}, lower; then the r rows of
// L_IP L_PP^{-1} = -E_I) and M_P^T (column j of M_P contiguous). Block
// (x, y) writes supernode lev[x]'s rows and columns y, y + gridDim.y, ...,
// so a level of few large supernodes still spreads over the chip.
template <class T>
__global__ void __launch_bounds__(256) k_fac_blocks(const FrontDev* __restrict__ fr, const int* __restrict__ lev,
                                                    const double* __restrict__ Ed, T* __restrict__ M,
                                                    T* __restrict__ Mt) {
  const FrontDev P = fr[lev[blockIdx.x]];
  const double* E = Ed + P.e;
  const int w = P.w, r = P.r, m = w + r;
  T* Mp = M + P.m_off;
  for (int row = blockIdx.y; row < m; row += gridDim.y) {
    if (row < w) {
      T* dst = Mp + (size_t)row * (row + 1) / 2;
      for (int j = threadIdx.x; j <= row; j += blockDim.x) dst[j] = (T)E[(size_t)row * P.wp + j];
    } else {
      const int q = row - w;
      T* dst = Mp + (size_t)w * (w + 1) / 2 + (size_t)q * w;
      for (int j = threadIdx.x; j < w; j += blockDim.x) dst[j] = (T)(-E[(size_t)(P.wp + q) * P.wp + j]);
    }
  }
  for (int j = blockIdx.y; j < w; j += gridDim.y) {
    T* dst = Mt + P.m_off + (size_t)j * m - (size_t)j * (j - 1) / 2;
    for (int kk = j + threadIdx.x; kk < m; kk += blockDim.x) {
      const double v = kk < w ? E[(size_t)kk * P.wp + j] : -E[(size_t)(P.wp + kk - w) * P.wp + j];
      dst[kk - j] = (T)v;
    }
  }
}

// A root's dense inverse S (left in its F by the kSRoot tasks) into Sroot;
// block (x, y) copies root roots[x]'s rows y, y + gridDim.y, ...
template <class T>
__global__ void __launch_bounds__(256) k_fac_root(const FrontDev* __restrict__ fr, const int* __restrict__ roots,
                                                  const double* __restrict__ Fd, T* __restrict__ Sroot) {
  const FrontDev P = fr[roots[blockIdx.x]];
  const double* S = Fd + P.f;
  const int w = P.w;
  T* dst = Sroot + P.s_off;
  for (int i = blockIdx.y; i < w; i += gridDim.y)
    for (int j = threadIdx.x; j < w; j += blockDim.x) dst[(size_t)i * w + j] = (T)S[(size_t)i * P.mp + j];
}

}  // namespace vhf
