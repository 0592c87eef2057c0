// Host-side set-up of the PD system for the HIP path: rest data, lumped
// masses, the constant system matrix, a nested-dissection supernodal LL^T
// factorization and the per-supernode blocks the GPU solve applies.
//
// Reference counterparts (file:line in the Vistio tree):
//   ElementData::from_mesh / compute_dm_inv   crates/vistio-solver/src/element.rs:54-91, 228-263
//   compute_lumped_masses                     crates/vistio-solver/src/pd_solver.rs:1219-1242
//   BendingData::from_topology                crates/vistio-solver/src/bending.rs:98-148
//   assemble_system_matrix                    crates/vistio-solver/src/assembly.rs:44-134
//   FaerSolver::factorize                     crates/vistio-math/src/faer_solver.rs:67-93
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace vh {

struct SetupInput {
  int n = 0, t = 0, h = 0;
  const float* px = nullptr;
  const float* py = nullptr;
  const float* pz = nullptr;
  const uint32_t* tri = nullptr;    // 3t vertex indices
  const uint32_t* hinge = nullptr;  // 4h: v0, v1, wing_a, wing_b
  const uint8_t* pinned = nullptr;  // n flags
  double k_membrane = 1e3;          // weight = k * area
  double k_bending = 1e-4;          // weight = k * edge length
  double density = 0.15;            // kg/m^2
  // Per-element materials (several garments in one system); when set they
  // replace the uniform values above: t, h and t entries.
  const double* k_membrane_tri = nullptr;
  const double* k_bending_hinge = nullptr;
  const double* density_tri = nullptr;
  // Bending kind: 1 = dihedral springs (stencil -1, -1, 1, 1), 2 = Discrete
  // Shells (per-hinge normalised cotangent stencil); for the latter a nonzero
  // bend_warp / bend_weft pair modulates each hinge's weight by its edge's
  // warp/weft alignment.
  int bending_kind = 1;
  double bend_warp = 0.0, bend_weft = 0.0;
  double dt_matrix = 1.0 / 60.0;
  int leaf_size = 16;               // nested-dissection leaf size
  // Level banding of the dissection tree. Each band of consecutive levels
  // becomes one level of supernodes (a node plus its descendants inside the
  // band form one dense block), trading block entries for fewer dependent
  // launches. Bands are chosen by dynamic programming on the cost
  //   2 * (band_launch_us + entries * value_bytes / band_gbps)
  // per band (forward and backward sweep).
  int value_bytes = 8;              // 8 = FP64 blocks, 4 = FP32 blocks
  double band_launch_us = 6.0;
  double band_gbps = 4000.0;
  int max_band_levels = 3;          // 1 = one band per dissection level
  int task_entries = 1024;          // block entries per solve task (rows per task follow)
};

// One supernode = one node of the nested-dissection tree: a contiguous run of
// w columns (in the permuted order) with r off-diagonal rows I_P (all in
// ancestors). Its solve block is the (w + r) x w lower trapezoid
//   M_P = [ L_PP^{-1} ; L_IP L_PP^{-1} ]
// stored row-major: rows 0..w-1 packed lower-triangular (row i has i+1
// entries), rows w..w+r-1 dense (w entries each).
struct Supernode {
  int col0 = 0, w = 0, r = 0;
  int64_t m_off = 0;   // offset of M_P in the value array
  int rows_off = 0;    // offset of I_P in rowidx
  int u_off = 0;       // offset (in rows) of this node's forward update vector
  int map_off = 0;     // offset of this node's r-entry map into its parent's front
  int parent = -1;
  int child_off = 0, nchild = 0;
  int depth = 0;
  int fm_off = 0;      // first forward slot of this node's front (fslot_k slots per position)
  int fslot_k = 0;     // slots per front position: most child update rows feeding one position
  int64_t s_off = -1;  // roots: offset of this root's dense inverse in Sroot
};

// A solve task: rows [b, e) of one of supernode sn's stored forms with g
// lanes per row. kind 0 = forward (rows of M_P), 1 = backward (rows of M_P^T,
// i.e. columns of M_P), 2 = root (rows of Sroot). Forward tasks cover every
// band but the root's; the root band's forward level holds the root tasks;
// the backward list has no root level.
struct Task {
  int sn, b, e, g, kind;
};

struct Setup {
  int n = 0, t = 0, h = 0;
  std::vector<int> perm, iperm;  // perm[new] = old vertex id
  // elements, vertex labels already permuted
  std::vector<int32_t> tri_idx;      // 3t
  std::vector<int32_t> tri_orig;     // t  (device triangle k is input triangle tri_orig[k])
  std::vector<double> tri_dm;        // 4t (dm_inv, column-major as the reference)
  std::vector<double> tri_w;         // t  (k * area)
  std::vector<int32_t> hinge_idx;    // 4h
  std::vector<double> hinge_rest;    // h
  std::vector<double> hinge_w;       // h
  std::vector<double> hinge_st;      // 4h stencils (Discrete Shells only; empty = -1, -1, 1, 1)
  std::vector<double> mass;          // n, permuted (pinned -> 1e8)
  std::vector<double> mass_orig;     // n, original vertex order (lumped_masses())
  std::vector<uint8_t> pinned;       // n, permuted
  std::vector<int32_t> inc_ptr;      // vertex i sums RHS slots [inc_ptr[i], inc_ptr[i + 1])
  // system matrix in the permuted order, full symmetric CSR
  std::vector<int32_t> Ap, Ai;
  std::vector<double> Ax;
  // supernodal structure
  std::vector<Supernode> sn;         // postorder: children before parents
  std::vector<int32_t> rowidx, children, emap;
  // Forward push slots. Every update-vector row (child c, row q) owns one
  // slot uslot[c.u_off + q] of the parent's front: position j of supernode P
  // owns slots P.fm_off + j*P.fslot_k + k, k < P.fslot_k, filled in child
  // order; unused slots stay zero. fslot_total = number of slots.
  std::vector<int32_t> uslot;
  int64_t fslot_total = 0;
  // Backward push slots: supernode c's off-diagonal rows read -d[I_c[q]]
  // from slot c.u_off + q; column k's value is pushed to
  // dpush_idx[dpush_ptr[k] .. dpush_ptr[k + 1]).
  std::vector<int32_t> dpush_ptr, dpush_idx;
  // RHS slots. The local step runs one block per chunk of 256 triangles or
  // 256 hinges (chunk c: the triangle chunks, then the hinge chunks); a block
  // sums its corners' contributions per vertex in LDS, in a fixed order, and
  // writes one slot per vertex it touches. Chunk c's local vertices are
  // [lv_ptr[c], lv_ptr[c + 1]); local vertex j writes slot lv_slot[j] with
  // the sum of the chunk's corners lc_pos[cb + lv_end[j - 1] .. cb + lv_end[j])
  // (lv_end counts from the chunk's first local vertex, which starts at 0;
  // cb = 3 * 256 * c for triangle chunk c, 3t + 4 * 256 * c' for hinge chunk
  // c'; lc_pos holds a corner's place in its block, element * K + k with
  // K = 3 for triangles and 4 for hinges). Each vertex's slots are ordered by
  // chunk.
  std::vector<int32_t> lv_ptr, lv_slot;
  std::vector<uint16_t> lv_end, lc_pos;
  std::vector<int32_t> upper_cols;   // columns outside the deepest band
  std::vector<int32_t> band_front;   // max (w + r) per band
  // The solve blocks (built on the device, vh_factor.hpp): M holds M_P as
  // row-major trapezoids at m_off (forward sweep), Mt holds M_P^T, column j
  // of M_P stored contiguously (backward sweep). A root supernode (no
  // off-diagonal rows) is solved in one pass: d_R = (L_RR L_RR^T)^{-1} t_R
  // with its dense w_R x w_R row-major inverse at Sroot[R.s_off]; a mesh of
  // several connected components (several garments) has one dissection tree,
  // one root, per component. m_total and sroot_total are the array sizes.
  int64_t sroot_total = 0;
  int64_t m_total = 0;
  int u_total = 0;
  int max_front = 0;                 // max (w + r)
  int max_depth = 0;                 // bands - 1
  int nd_levels = 0;                 // levels of the fine dissection tree
  std::vector<int> band_start;       // first dissection level of each band
  std::vector<Task> fwd_tasks, bwd_tasks;
  std::vector<int> fwd_level_off, bwd_level_off;  // tasks of level l: [off[l], off[l+1])
  // Dependency-driven schedule of both sweeps in one launch: all forward
  // tasks (deepest level first) then all backward tasks (root first), so
  // every task depends only on tasks listed before it.
  std::vector<Task> ptasks;
  std::vector<int32_t> need_f;       // forward chunks of P's children (wait count of P's forward)
  std::vector<int32_t> need_b;       // backward chunks of P's parent, or forward chunks of the root
  std::vector<int32_t> parent;
  int64_t nnz_factor = 0;            // sum of w(w+1)/2 + r*w
  std::string err;
  int status = 0;                    // 0 ok, else a vh_status code
};

// Builds everything. Returns false and sets err/status on failure.
bool build_setup(const SetupInput& in, Setup& out);

// The device kernels hold block offsets (task records), root-inverse
// offsets, front-slot numbers and RHS-slot value indices (3 per element
// corner) as 32-bit integers. Returns the reason when one of these totals
// does not fit, else an empty string; build_setup refuses such a system
// with VH_E_UNSUPPORTED.
std::string device_index_overflow(int64_t m_total, int64_t sroot_entries, int64_t front_slots, int64_t rhs_values);

}  // namespace vh
