// Exclusive scan and stable LSD radix sort of 32-bit keys with 32-bit values,
// written for 64-lane wavefronts. The broad phase sorts the triangles by
// Morton code (the order BvhTree::build reaches by its median splits,
// crates/vistio-contact/src/bvh.rs:133-178, is replaced by an LBVH over this
// order) and scans the per-leaf pair counts and the per-vertex contact counts.
//
// Every kernel takes the contact step's skip flag (a device int; the kernel
// returns when it is set) and nothing here records a memset or calls a
// library, so a launch sequence that uses these can be captured into a HIP
// graph. All arithmetic is on integers: the results do not depend on the
// order in which blocks run.
//
// Scan: reduce (one sum per tile of 2048; the last block to finish, which it
// learns from an atomic ticket, scans the tile sums), then a down-sweep (each
// tile scans itself from its base). Sort: per pass of 8 bits, a digit
// histogram per tile, one block that turns them into each tile's first slot
// per digit, and a scatter in which each item's rank among the equal digits
// of its tile comes from wave ballots, in item order, so the sort is stable.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace vhs {

constexpr int kThreads = 256;      // 4 waves of 64
constexpr int kPer = 8;            // items per thread
constexpr int kTile = kThreads * kPer;
constexpr int kBins = 256;         // 8-bit digits

inline long long tiles(long long len) { return (len + kTile - 1) / kTile; }

// Exclusive prefix of v over the 256 threads of the block; *total gets the
// block's sum. wsum: 4 words of LDS. Every thread of the block must call it.
__device__ __forceinline__ unsigned int block_exclusive(unsigned int v, unsigned int* wsum, unsigned int* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  unsigned int inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned int u = __shfl_up(inc, o, 64);
    if (lane >= o) inc += u;
  }
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  unsigned int base = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const unsigned int s = wsum[k];
    if (k < w) base += s;
    tot += s;
  }
  __syncthreads();  // wsum may be written again by the next call
  *total = tot;
  return base + inc - v;
}

// bsum[b] = sum of in[b * kTile .. (b + 1) * kTile). The last block to
// finish then replaces bsum[0..nb) by its exclusive scan and re-arms the
// ticket (a device word that is 0 between launches). No block waits for
// another: each writes its sum, fences, and takes a ticket.
__global__ void __launch_bounds__(kThreads) k_scan_reduce(const unsigned int* __restrict__ in, long long len,
                                                          unsigned int* bsum, long long nb, unsigned int* ticket,
                                                          const int* skip) {
  if (skip && *skip) return;
  __shared__ unsigned int wsum[4];
  __shared__ int last;
  const long long b0 = (long long)blockIdx.x * kTile;
  unsigned int s = 0;
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const long long i = b0 + k * kThreads + threadIdx.x;
    if (i < len) s += in[i];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    bsum[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __threadfence();  // the sum is visible on every XCD before the ticket is taken
    last = atomicAdd(ticket, 1u) == (unsigned int)(nb - 1);
  }
  __syncthreads();
  if (!last) return;
  __threadfence();  // every block's sum was fenced before its ticket
  unsigned int carry = 0;
  for (long long c0 = 0; c0 < nb; c0 += kTile) {
    const long long i0 = c0 + (long long)threadIdx.x * kPer;
    unsigned int v[kPer], t = 0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      v[k] = i0 + k < nb ? bsum[i0 + k] : 0u;
      t += v[k];
    }
    unsigned int tot;
    unsigned int p = carry + block_exclusive(t, wsum, &tot);
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      if (i0 + k < nb) bsum[i0 + k] = p;
      p += v[k];
    }
    carry += tot;
  }
  if (threadIdx.x == 0) atomicExch(ticket, 0u);
}

// out[i] = base of i's tile (bsum[tile], or 0 without bsum) + the sum of the
// tile's items before i. out may alias in: each thread reads its items before
// it writes them.
__global__ void __launch_bounds__(kThreads) k_scan_down(const unsigned int* in, long long len,
                                                        const unsigned int* __restrict__ bsum, unsigned int* out,
                                                        const int* skip) {
  if (skip && *skip) return;
  __shared__ unsigned int wsum[4];
  const long long i0 = (long long)blockIdx.x * kTile + (long long)threadIdx.x * kPer;
  unsigned int v[kPer], s = 0;
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    v[k] = i0 + k < len ? in[i0 + k] : 0u;
    s += v[k];
  }
  unsigned int tot;
  unsigned int p = (bsum ? bsum[blockIdx.x] : 0u) + block_exclusive(s, wsum, &tot);
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    if (i0 + k < len) out[i0 + k] = p;
    p += v[k];
  }
}

// Words of scan scratch (bsum) that scan_exclusive needs for len items.
inline long long scan_scratch_words(long long len) { return tiles(len) > 1 ? tiles(len) : 1; }

// out[0..len) = exclusive scan of in[0..len) (out may alias in). bsum:
// scan_scratch_words(len) words of device memory; ticket: a device word that
// holds 0 (the reduce leaves it 0).
inline void scan_exclusive(hipStream_t s, const unsigned int* in, unsigned int* out, long long len,
                           unsigned int* bsum, unsigned int* ticket, const int* skip) {
  if (len <= 0) return;
  const long long nb = tiles(len);
  if (nb == 1) {
    hipLaunchKernelGGL(k_scan_down, dim3(1), dim3(kThreads), 0, s, in, len, (const unsigned int*)nullptr, out, skip);
    return;
  }
  hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned int)nb), dim3(kThreads), 0, s, in, len, bsum, nb, ticket, skip);
  hipLaunchKernelGGL(k_scan_down, dim3((unsigned int)nb), dim3(kThreads), 0, s, in, len, bsum, out, skip);
}

// hist[b * 256 + d] = items of tile b whose digit (key >> shift) & 255 is d
// (tile-major, so that this write and every later access are coalesced).
__global__ void __launch_bounds__(kThreads) k_rs_hist(const unsigned int* __restrict__ key, long long n, int shift,
                                                      unsigned int* __restrict__ hist, const int* skip) {
  if (skip && *skip) return;
  __shared__ unsigned int cnt[kBins];
  cnt[threadIdx.x] = 0u;
  const long long b0 = (long long)blockIdx.x * kTile;
  unsigned int k8[kPer];
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const long long i = b0 + k * kThreads + threadIdx.x;
    k8[k] = i < n ? (key[i] >> shift) & (kBins - 1) : kBins;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kPer; ++k)
    if (k8[k] < kBins) atomicAdd(&cnt[k8[k]], 1u);
  __syncthreads();
  hist[(long long)blockIdx.x * kBins + threadIdx.x] = cnt[threadIdx.x];
}

constexpr int kSpineThreads = 1024;  // 4 threads per digit
constexpr int kSpineBatch = 32;      // tiles a thread loads at once

// In place: hist[b * 256 + d] -> the first output slot of tile b's items of
// digit d (all items of smaller digits, then those of digit d in tiles before
// b). One block; thread (d, q) takes the q-th quarter of the tiles for digit
// d, so each wave's loads cover 64 consecutive words.
__global__ void __launch_bounds__(kSpineThreads) k_rs_spine(unsigned int* __restrict__ hist, long long nb,
                                                            const int* skip) {
  if (skip && *skip) return;
  __shared__ unsigned int part[4][kBins];
  __shared__ unsigned int wsum[4];
  const int d = threadIdx.x & (kBins - 1), q = threadIdx.x >> 8, lane = threadIdx.x & 63;
  const long long per = (nb + 3) / 4, lo = q * per, hi = lo + per < nb ? lo + per : nb;
  unsigned int s = 0;
  for (long long b0 = lo; b0 < hi; b0 += kSpineBatch) {
    unsigned int v[kSpineBatch];
#pragma unroll
    for (int i = 0; i < kSpineBatch; ++i) v[i] = b0 + i < hi ? hist[(b0 + i) * kBins + d] : 0u;
#pragma unroll
    for (int i = 0; i < kSpineBatch; ++i) s += v[i];
  }
  part[q][d] = s;
  __syncthreads();
  unsigned int pre = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const unsigned int u = part[k][d];
    if (k < q) pre += u;
    tot += u;
  }
  // exclusive scan of the digit totals: every wave scans its 64 digits (the
  // four quarters get the same values); the first quarter publishes the sums
  unsigned int inc = tot;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned int u = __shfl_up(inc, o, 64);
    if (lane >= o) inc += u;
  }
  if (q == 0 && lane == 63) wsum[d >> 6] = inc;
  __syncthreads();
  unsigned int run = inc - tot + pre;
  for (int k = 0; k < (d >> 6); ++k) run += wsum[k];
  for (long long b0 = lo; b0 < hi; b0 += kSpineBatch) {
    unsigned int v[kSpineBatch];
#pragma unroll
    for (int i = 0; i < kSpineBatch; ++i) v[i] = b0 + i < hi ? hist[(b0 + i) * kBins + d] : 0u;
#pragma unroll
    for (int i = 0; i < kSpineBatch; ++i) {
      if (b0 + i < hi) hist[(b0 + i) * kBins + d] = run;
      run += v[i];
    }
  }
}

// One stable pass: item i of tile b with digit d goes to off[b * 256 + d] +
// (items of tile b before i with digit d). The tile's 8 rounds of 256
// consecutive items are loaded first; within a round, a wave finds the lanes
// that share its digit with 8 ballots, and the waves' per-digit counts order
// the waves. The counts alternate between two buffers, so a round needs two
// barriers: one before the counts are read, one before they are folded into
// the digit bases and cleared.
__global__ void __launch_bounds__(kThreads) k_rs_scatter(const unsigned int* __restrict__ kin,
                                                         const int* __restrict__ vin, unsigned int* __restrict__ kout,
                                                         int* __restrict__ vout, long long n, int shift,
                                                         const unsigned int* __restrict__ off, const int* skip) {
  if (skip && *skip) return;
  __shared__ unsigned int base[kBins];        // next output slot of each digit from this tile
  __shared__ unsigned int wcnt[2][4][kBins];  // items of each digit per wave, by round parity
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const long long b0 = (long long)blockIdx.x * kTile;
  unsigned int key[kPer];
  int val[kPer];
#pragma unroll
  for (int r = 0; r < kPer; ++r) {
    const long long i = b0 + r * kThreads + tid;
    key[r] = i < n ? kin[i] : 0u;
    val[r] = i < n ? vin[i] : 0;
  }
  base[tid] = off[(long long)blockIdx.x * kBins + tid];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    wcnt[0][k][tid] = 0u;
    wcnt[1][k][tid] = 0u;
  }
  const unsigned long long lower = (1ull << lane) - 1ull;
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kPer; ++r) {
    unsigned int (*wc)[kBins] = wcnt[r & 1];
    const bool valid = b0 + r * kThreads + tid < n;
    const unsigned int d = (key[r] >> shift) & (kBins - 1);
    unsigned long long peers = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < 8; ++bit) {
      const bool on = (d >> bit) & 1u;
      const unsigned long long m = __ballot(on);
      peers &= on ? m : ~m;
    }
    const unsigned int rank = (unsigned int)__popcll(peers & lower);
    if (valid && rank == 0u) wc[w][d] = (unsigned int)__popcll(peers);
    __syncthreads();
    if (valid) {
      unsigned int pos = base[d] + rank;
      for (int k = 0; k < w; ++k) pos += wc[k][d];
      kout[pos] = key[r];
      vout[pos] = val[r];
    }
    __syncthreads();
    // only thread tid touches column tid here; the other buffer is in use
    // by the next round, whose writes follow the next barrier
    base[tid] += wc[0][tid] + wc[1][tid] + wc[2][tid] + wc[3][tid];
#pragma unroll
    for (int k = 0; k < 4; ++k) wc[k][tid] = 0u;
  }
}

// Words of the per-tile histograms that sort_pairs needs for n items.
inline long long sort_hist_words(long long n) { return (long long)kBins * (tiles(n) > 0 ? tiles(n) : 1); }

// Stable sort of (key, val) by the low `bits` bits of key, in passes of 8
// bits. The pairs start in (k0, v0); with an even pass count (bits 25..32)
// they end there, with an odd one in (k1, v1): the return value says which
// (0 or 1). hist: sort_hist_words(n) words.
inline int sort_pairs(hipStream_t s, unsigned int* k0, int* v0, unsigned int* k1, int* v1, long long n, int bits,
                      unsigned int* hist, const int* skip) {
  if (n <= 1) return 0;
  const long long nb = tiles(n);
  int cur = 0;
  for (int shift = 0; shift < bits; shift += 8) {
    unsigned int* ki = cur ? k1 : k0;
    int* vi = cur ? v1 : v0;
    unsigned int* ko = cur ? k0 : k1;
    int* vo = cur ? v0 : v1;
    hipLaunchKernelGGL(k_rs_hist, dim3((unsigned int)nb), dim3(kThreads), 0, s, ki, n, shift, hist, skip);
    hipLaunchKernelGGL(k_rs_spine, dim3(1), dim3(kSpineThreads), 0, s, hist, nb, skip);
    hipLaunchKernelGGL(k_rs_scatter, dim3((unsigned int)nb), dim3(kThreads), 0, s, ki, vi, ko, vo, n, shift,
                       (const unsigned int*)hist, skip);
    cur ^= 1;
  }
  return cur;
}

}  // namespace vhs
