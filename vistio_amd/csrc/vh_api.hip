// C ABI of the gfx950 Projective Dynamics step (declared in include/vistio_hip.h).
//
// One vh_ctx owns one HIP stream and all device buffers of one solver
// instance. A step runs entirely on the device:
//   predict -> { local -> forward bands (first one also forms r) ->
//                backward bands (last one also commits x) [-> refinement] }
//              x max_iterations -> residual -> finalize
// following ProjectiveDynamicsSolver::step (crates/vistio-solver/src/pd_solver.rs:259-476).
// The whole step is captured once into a HIP graph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "vh_kernels.hpp"
#include "vh_contact.hpp"
#include "vh_factor.hpp"
#include "vh_setup.hpp"
#include "vh_sort.hpp"
#include "vh_narrow.hpp"
#include "vistio_hip.h"

// Device buffers and settings of the contact step (vh_ipc_init).
struct IpcState {
  bool on = false;
  vh_ipc_params P{};
  vhc::Scene sc{};
  double *gself = nullptr, *grad = nullptr, *nrm = nullptr, *fnrm = nullptr, *lambda = nullptr, *ext = nullptr;
  double *partials = nullptr, *scratch_a = nullptr, *scratch_b = nullptr;
  uint8_t *inc = nullptr, *ever = nullptr;
  int *xptr = nullptr, *xidx = nullptr;
  vhc::Box* box = nullptr;
  vhc::Rec* node = nullptr;  // t - 1 internal nodes, then the t leaves (leaf_of)
  unsigned int *code = nullptr, *code_s = nullptr;
  int *id = nullptr, *id_s = nullptr;
  int *left = nullptr, *range_lo = nullptr, *range_hi = nullptr;
  vhc::Box* st = nullptr;  // sparse table of leaf-box unions, st_levels x t (levels 0..8)
  int st_levels = 0;
  vhc::Box* coarse = nullptr;  // sparse table over whole 256-leaf cells (k_st_coarse)
  int* h_done = nullptr;   // pinned host copy of IpcDev::al_done
  unsigned int* h_np = nullptr;  // pinned host copies of BvhDev::npairs and ::ncont (same allocation)
  unsigned int *pcount = nullptr, *poff = nullptr;  // candidate pairs per leaf, their offsets
  // the pairs k_bvh_walk1 keeps: kLeafStage per leaf (whole workgroup
  // stages) and wmore_leaf per leaf after them
  int2 *wkept = nullptr, *wmore = nullptr;
  int wmore_leaf = 0;
  float* bpart = nullptr;                             // per-block centroid bounds of k_boxes
  unsigned int* sort_hist = nullptr;  // per-tile digit histograms of the Morton sort (vh_sort.hpp)
  unsigned int* sbuf = nullptr;       // tile sums of every scan (the scans run one after another)
  unsigned int* ticket = nullptr;     // the scans' finishing-block counter (0 between launches)
  int* orig = nullptr;  // caller's index of each device triangle
  int* sorig = nullptr;  // caller's index of each sorted leaf
  int2* pairs = nullptr;
  unsigned int cap = 0;
  vhc::CcdTri* ccdtri = nullptr;  // per-triangle bounds of the CCD pair cull
  int2* ccd_surv = nullptr;        // pairs that survive the self-contact or CCD cull, one region of ccd_reg per block
  unsigned int ccd_reg = 0;
  unsigned int* ccd_cnt = nullptr;  // survivors per block
  vhc::BvhDev* bd = nullptr;
  vhc::IpcDev* ip = nullptr;
  double cfl_dh = 0;  // d_hat of the CFL velocity clamp (compliant-scaled when compliant_contact)
  // Self-contact gradients summed in a fixed order (vh_contact.hpp,
  // k_self_count .. k_vgather): per-survivor offsets (one region of ccd_reg
  // per block, as ccd_surv), per-block totals and bases, the contact entries
  // (4 per contact: vertex, gradient) for up to ecap contacts, and their
  // grouping by vertex (counts, fill, segment starts over n + 1).
  unsigned int *cof = nullptr, *btot = nullptr, *sbase = nullptr;
  int* ev = nullptr;
  double* eg = nullptr;
  unsigned int* vidx = nullptr;
  unsigned int *vcnt = nullptr, *vfill = nullptr, *vptr = nullptr;
  unsigned int ecap = 0;
  // The launch sequences of the contact step, each captured once into a
  // graph and replayed: [0] the head of a step (begin, first detection),
  // [1] an AL iteration other than the last, [2] the last one (an AL
  // iteration: the PD loop, then the detection at its result, the multiplier
  // update and the decision). Recaptured when dt or a buffer the launches
  // name changes (gen counts the regrowths).
  struct Seq {
    hipGraphExec_t exec = nullptr;
    double dt = 0;
    unsigned int gen = 0;
  };
  Seq seq[3];
  unsigned int gen = 0;
};

struct vh_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  bool initialized = false;
  vh_pd_params P{};
  vh::Setup S;
  double setup_seconds = 0;   // host set-up (symbolic)
  double factor_seconds = 0;  // device factorization
  int prec = VH_PREC_F64;
  size_t tsz = 8;  // bytes of the compute type
  // state, FP64 [n][3]
  double *x = nullptr, *v = nullptr, *prev = nullptr, *pred = nullptr, *r = nullptr, *acc = nullptr;
  double* mass = nullptr;
  uint8_t* pinned = nullptr;
  // elements
  int *tri_idx = nullptr, *hinge_idx = nullptr;
  int *lv_ptr = nullptr, *lv_slot = nullptr;              // the local step's RHS slot plan (vh::Setup)
  unsigned short *lv_end = nullptr, *lc_pos = nullptr;
  vhk::LocalMap local_map{};  // k_local's block -> element chunk map
  int local_blocks = 0;
  double* tri_dm = nullptr;
  void *tri_w = nullptr, *hinge_rest = nullptr, *hinge_w = nullptr;
  void* hinge_st = nullptr;  // Discrete Shells stencils (T [h][4]); null for dihedral springs
  vhk::MemModel mem{};       // membrane model and its normalised warp/weft factors
  void* cslot = nullptr;  // RHS contribution slots, T [3t + 4h][3]
  int* inc_ptr = nullptr;
  // solve
  vhk::TaskDev* tasks = nullptr;  // forward tasks then backward tasks
  // The root band below other bands (k_root_dense): root_rows rows of one
  // root per task, root_nl entries per lane loaded in the first round.
  vhk::TaskDev* root_tasks = nullptr;
  int nroot_tasks = 0, root_rows = 0, root_nl = 0;
  size_t root_lds = 0;
  // Forward bands below the first and above the root with their block rows
  // staged in LDS (k_fwd_band_st): tasks of up to 256 st_E entries, those
  // of forward level l at [st_off[l], st_off[l + 1]) (empty: that band runs
  // k_fwd_band).
  vhk::TaskDev* st_tasks = nullptr;
  std::vector<int> st_off;
  int st_E = 0;
  // The first forward launch from the RHS slots, staged (k_fwd_band_st1):
  // nst1 tasks with their slot ranges, st1_NS slot values per thread (0:
  // k_fwd_band<T, true>).
  vhk::TaskDev* st1_tasks = nullptr;
  int nst1 = 0, st1_NS = 0;
  // The committing last backward band, staged (k_bwd_band_stc): nstc tasks
  // (0: k_bwd_band<T, true>).
  vhk::TaskDev* stc_tasks = nullptr;
  int nstc = 0;
  int *uslot = nullptr, *dpush_ptr = nullptr, *dpush_idx = nullptr, *upper_cols = nullptr;
  void *M = nullptr, *Mt = nullptr, *Sroot = nullptr, *y = nullptr, *d = nullptr, *F = nullptr, *V = nullptr;
  int *Ap = nullptr, *Ai = nullptr;
  double* Ax = nullptr;
  double* partials = nullptr;
  int resid_blocks = 0;  // blocks writing residual partials
  vhk::StepDev* st = nullptr;
  const double* cur_ext = nullptr;  // RHS external term of the launches being enqueued
  IpcState ipc;
  std::vector<void*> allocs;
  std::vector<double> hbuf;  // host staging (3n)
  // Drop-in steps on host f32 state: pinned staging of the caller-order SoA
  // arrays (16 n floats: in pos, vel, inv_mass, lambda; out pos, vel, prev,
  // pred, lambda), its device copy and the permutation, made on first use.
  float* h_io = nullptr;
  float* d_io = nullptr;
  int* d_perm = nullptr;
  // GpuBackend staging (vh_backend_*): device copies of the operands, kept
  // across calls and grown on demand; independent of the solver's state.
  float *bk_a = nullptr, *bk_b = nullptr, *bk_o = nullptr;
  // vh_narrow_vertex_triangle's device buffers, one growable allocation; its
  // first word is the scan's ticket (0 between launches)
  void* nb_arena = nullptr;
  size_t nb_bytes = 0;
  double* bk_part = nullptr;
  size_t bk_cap = 0;
  hipGraphExec_t step_graph = nullptr;
  double graph_dt = 0;
  bool use_graph = true;
};

namespace {

vh_status fail(vh_ctx* c, vh_status s, const std::string& msg) {
  if (c) c->err = msg;
  return s;
}

#define VH_HIP(call)                                                                           \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess)                                                                      \
      return fail(ctx, VH_E_GPU, std::string(#call ": ") + hipGetErrorString(e_));             \
  } while (0)

template <class T>
vh_status dalloc(vh_ctx* ctx, T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  void* q = nullptr;
  VH_HIP(hipMalloc(&q, count * sizeof(T)));
  ctx->allocs.push_back(q);
  *p = static_cast<T*>(q);
  return VH_OK;
}

vh_status dalloc_bytes(vh_ctx* ctx, void** p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 8;
  VH_HIP(hipMalloc(p, bytes));
  ctx->allocs.push_back(*p);
  return VH_OK;
}

template <class T>
vh_status upload(vh_ctx* ctx, T** p, const std::vector<T>& h) {
  vh_status s = dalloc(ctx, p, h.size());
  if (s != VH_OK) return s;
  if (!h.empty()) VH_HIP(hipMemcpy(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return VH_OK;
}

// Uploads a double array as the compute type (double or float).
vh_status upload_t(vh_ctx* ctx, void** p, const std::vector<double>& h) {
  vh_status s = dalloc_bytes(ctx, p, h.size() * ctx->tsz);
  if (s != VH_OK) return s;
  if (h.empty()) return VH_OK;
  if (ctx->tsz == 8) {
    VH_HIP(hipMemcpy(*p, h.data(), h.size() * 8, hipMemcpyHostToDevice));
  } else {
    std::vector<float> f(h.size());
    for (size_t k = 0; k < h.size(); ++k) f[k] = (float)h[k];
    VH_HIP(hipMemcpy(*p, f.data(), f.size() * 4, hipMemcpyHostToDevice));
  }
  return VH_OK;
}

void drop_graph(vh_ctx* ctx) {
  if (ctx->step_graph) (void)hipGraphExecDestroy(ctx->step_graph);
  ctx->step_graph = nullptr;
  ctx->graph_dt = 0;
}

void drop_al_graph(vh_ctx* ctx) {
  for (auto& q : ctx->ipc.seq) {
    if (q.exec) (void)hipGraphExecDestroy(q.exec);
    q.exec = nullptr;
  }
}

void free_all(vh_ctx* ctx) {
  drop_graph(ctx);
  drop_al_graph(ctx);
  for (void* p : ctx->allocs) (void)hipFree(p);
  ctx->allocs.clear();
  ctx->initialized = false;
  if (ctx->ipc.h_done) (void)hipHostFree(ctx->ipc.h_done);
  ctx->ipc = IpcState{};
  if (ctx->h_io) (void)hipHostFree(ctx->h_io);
  ctx->h_io = nullptr;
  ctx->d_io = nullptr;
  ctx->d_perm = nullptr;
}

inline int nblk(int n) { return (n + 255) / 256; }

// ---- kernel sequences -----------------------------------------------------

template <class T>
vhk::RhsSrc<T> rhs_src(vh_ctx* c, double dt) {
  return vhk::RhsSrc<T>{c->inc_ptr, (const T*)c->cslot, c->mass, c->pred, c->x, 1.0 / (dt * dt),
                        1.0 / (c->P.dt_matrix * c->P.dt_matrix), c->cur_ext};
}

vhk::CommitArgs commit_args(vh_ctx* c, bool on, bool want_resid, bool add_acc) {
  return vhk::CommitArgs{on ? 1 : 0,   c->P.has_ground, c->P.ground_height, c->x, c->prev, c->pinned, c->partials,
                         want_resid ? 1 : 0, c->acc, add_acc ? 1 : 0};
}

// Task blocks of a band launch: one contiguous run of tasks per XCD
// (BandArgs::xper; see task_of) for bands of 64 to 4096 tasks, and for the
// first launch of a sweep (it also gathers the RHS slots of its columns)
// whatever its size; otherwise one block per task, dealt to the XCDs in turn.
// Measured: the runs take a single 256x256 sheet from 949 to 974 steps/s;
// on 64 sheets in one system the runs slowed the larger bands by 4-7 %
// (tasks of unequal cost gather on one XCD) but sped up the first launch.
constexpr int kXcdMinTasks = 64, kXcdMaxTasks = 4096;
int band_xper(int ntasks, bool first) {
  const bool runs = ntasks >= kXcdMinTasks && (first || ntasks <= kXcdMaxTasks);
  return runs ? (ntasks + vhk::kXcds - 1) / vhk::kXcds : 0;
}
int band_task_blocks(int ntasks, bool first) {
  const int x = band_xper(ntasks, first);
  return x ? vhk::kXcds * x : ntasks;
}

// Bands of more tasks than this run the kernels compiled for more waves
// per SIMD (k_fwd_band_w8 for FP32; k_fwd_band_d5, k_bwd_band_d5 for FP64,
// not the committing backward band). On 64
// sheets in one system the FP32 one took forward bands 1-5 from 377 to
// 333 us; on one 256x256 sheet (bands of at most 2745 tasks) the same
// kernel everywhere cost 3.6 % of the step.
constexpr int kWideBandTasks = 4096;

// The dense root kernel's shape for a widest root of w columns: rows per
// workgroup (one per wave; each workgroup gathers the root's front once)
// and entries per lane loaded in the first round (a multiple of 8; longer
// rows continue 8 per lane at a time). Measured at 256x256 (FP32): a
// 1535-wide root in 512-thread workgroups 13.2 -> 11.7 us per launch (the
// 256-thread band kernel before it), a 2295-wide one in 1024-thread
// workgroups 18.3 -> 16.8 us; 512-thread workgroups for the 2295-wide root
// (288 workgroups, all CUs busy, the front gathered twice as often) measured
// slower than 1024 (1150 -> 1118 steps/s). FP64 keeps k_root_band: its 24
// entries per lane spill in 1024-thread workgroups (15.0 -> 19.0 us at 1535
// wide).
int root_threads(int w, bool f64) {
  if (f64) return 0;
  return w > 2048 ? 1024 : 512;
}
int root_loads(int w, bool f64) {
  const int cap = f64 ? 24 : 40;
  return std::min(cap, std::max(8, ((w + 63) / 64 + 7) / 8 * 8));
}

#define VH_ROOT_NL(X) X(8) X(16) X(24) X(32) X(40)

// Staged forward bands (k_fwd_band_st): 8 entries per thread (tasks of up
// to 2048 entries), for bands whose supernodes are at most 128 columns
// wide. Measured at 256x256 (5 bands): the 87-wide band 16.9 -> 12.8 us
// (FP32) and 21.8 -> 15.9 us (FP64), the 95-wide one unchanged; the
// 475-wide band slower staged (12.7 -> 17.6 us, the rows' LDS dot products
// run in sequence), 16 or 32 entries per thread slower everywhere (fewer
// workgroups). Staged backward bands were measured too and gained nothing.
constexpr int kStagedLoads = 8;
constexpr int kStagedMaxW = 128;

#define VH_ST_E(X) X(8)

template <class T>
const void* fwd_st_kernel(int E) {
  switch (E) {
#define VH_CASE(N) \
  case N: return (const void*)vhk::k_fwd_band_st<N, T>;
    VH_ST_E(VH_CASE)
#undef VH_CASE
    default: return nullptr;
  }
}

template <class T>
void launch_fwd_st(vh_ctx* c, const vhk::BandArgs<T>& a, int grid, size_t lds, int fmax, const vhk::StepDev* st) {
  switch (c->st_E) {
#define VH_CASE(N)                                                                                             \
  case N:                                                                                                      \
    hipLaunchKernelGGL((vhk::k_fwd_band_st<N, T>), dim3(grid), dim3(256), lds, c->stream, a, (const T*)c->M, \
                       (T*)c->y, (T*)c->F, fmax, st);                                                          \
    break;
    VH_ST_E(VH_CASE)
#undef VH_CASE
    default: break;
  }
}

size_t staged1_lds(const vh_ctx* c, int fmax) {
  return 3 * 256 * sizeof(double) + (3 * (size_t)fmax + 256 * (size_t)c->st_E + 256 * (size_t)c->st1_NS) * c->tsz +
         512 * sizeof(int);
}

#define VH_ST1_NS(X) X(8) X(16) X(32)

template <class T>
const void* fwd_st1_kernel(int ns) {
  switch (ns) {
#define VH_CASE(N) \
  case N: return (const void*)vhk::k_fwd_band_st1<kStagedLoads, N, T>;
    VH_ST1_NS(VH_CASE)
#undef VH_CASE
    default: return nullptr;
  }
}

template <class T>
void launch_fwd_st1(vh_ctx* c, const vhk::BandArgs<T>& a, int grid, size_t lds, int fmax, const vhk::StepDev* st) {
  switch (c->st1_NS) {
#define VH_CASE(N)                                                                                            \
  case N:                                                                                                     \
    hipLaunchKernelGGL((vhk::k_fwd_band_st1<kStagedLoads, N, T>), dim3(grid), dim3(256), lds, c->stream, a, \
                       (const T*)c->M, (T*)c->y, (T*)c->F, fmax, st);                                         \
    break;
    VH_ST1_NS(VH_CASE)
#undef VH_CASE
    default: break;
  }
}

size_t stagedc_lds(const vh_ctx* c, int fmax) {
  return 9 * 256 * sizeof(double) + (3 * (size_t)fmax + 256 * (size_t)c->st_E) * c->tsz + 3 * 256 * sizeof(int);
}

size_t staged_lds(const vh_ctx* c, int fmax) {
  return (3 * (size_t)fmax + 256 * (size_t)c->st_E) * c->tsz + 512 * sizeof(int);
}


const void* root_dense_kernel(int nl) {
  switch (nl) {
#define VH_CASE(N) \
  case N: return (const void*)vhk::k_root_dense<N, float>;
    VH_ROOT_NL(VH_CASE)
#undef VH_CASE
    default: return nullptr;
  }
}

void launch_root_dense(vh_ctx* c, const vhk::BandArgs<float>& a, const vhk::StepDev* st) {
  const dim3 grid(c->nroot_tasks), block(64 * c->root_rows);
  switch (c->root_nl) {
#define VH_CASE(N)                                                                                         \
  case N:                                                                                                  \
    hipLaunchKernelGGL((vhk::k_root_dense<N, float>), grid, block, c->root_lds, c->stream, c->root_tasks, a, \
                       (const float*)c->Sroot, (const float*)c->F, (float*)c->V, (float*)c->d, st);        \
    break;
    VH_ROOT_NL(VH_CASE)
#undef VH_CASE
    default: break;
  }
}

// Records an event after each launch when profiling (ev == nullptr: off).
struct Marks {
  hipEvent_t* ev = nullptr;
  size_t k = 0;
  void mark(hipStream_t s) {
    if (ev) (void)hipEventRecord(ev[k++], s);
  }
};

// One solve of A d = r: forward bands deepest first, then the root band
// (forward and backward in one launch through its dense inverse), then the
// backward bands below the root. from_slots: the first launch forms r from
// the RHS slots. commit: the last launch adds d (plus acc when add_acc) to x.
// Launch count: 2 * bands - 1.
template <class T>
void enqueue_solve(vh_ctx* c, double dt, bool from_slots, bool commit, bool add_acc, bool want_resid,
                   const vhk::StepDev* st, Marks& mk) {
  const vh::Setup& S = c->S;
  const int nb = S.max_depth + 1;
  const int nfwd = (int)S.fwd_tasks.size();
  const int nupper = (int)S.upper_cols.size();
  vhk::BandArgs<T> a{};
  a.uslot = c->uslot; a.dpush_ptr = c->dpush_ptr; a.dpush_idx = c->dpush_idx;
  a.src = rhs_src<T>(c, dt); a.r = c->r; a.upper_cols = c->upper_cols; a.nupper = nupper;
  for (int l = 0; l + 1 < nb; ++l) {
    if (l > 0 && c->st_E > 0 && c->st_off[l + 1] > c->st_off[l]) {
      const int b = c->st_off[l], e = c->st_off[l + 1];
      a.tasks = c->st_tasks + b; a.ntasks = e - b; a.xper = band_xper(e - b, false);
      const int fmax = S.band_front[S.max_depth - l];
      launch_fwd_st<T>(c, a, band_task_blocks(e - b, false), staged_lds(c, fmax), fmax, st);
      mk.mark(c->stream);
      continue;
    }
    const bool first = l == 0 && from_slots;
    if (first && c->nst1 > 0) {
      a.tasks = c->st1_tasks; a.ntasks = c->nst1; a.xper = band_xper(c->nst1, true);
      const int fmax = S.band_front[S.max_depth];
      launch_fwd_st1<T>(c, a, band_task_blocks(c->nst1, true) + nblk(4 * nupper), staged1_lds(c, fmax), fmax, st);
      mk.mark(c->stream);
      continue;
    }
    const int b = S.fwd_level_off[l], e = S.fwd_level_off[l + 1];
    a.tasks = c->tasks + b; a.ntasks = e - b; a.xper = band_xper(e - b, l == 0);
    const int grid = band_task_blocks(e - b, l == 0) + (first ? nblk(4 * nupper) : 0);
    const size_t lds = 3 * (size_t)S.band_front[S.max_depth - l] * c->tsz;
    if (grid > 0) {
      if (first)
        hipLaunchKernelGGL((vhk::k_fwd_band<T, true>), dim3(grid), dim3(256), lds, c->stream, a, (const T*)c->M,
                           (T*)c->y, (T*)c->F, st);
      else {
        bool wide = false;
        wide = e - b > kWideBandTasks;
        if constexpr (std::is_same<T, float>::value) {
          if (wide)
            hipLaunchKernelGGL(vhk::k_fwd_band_w8, dim3(grid), dim3(256), lds, c->stream, a, (const T*)c->M,
                               (T*)c->y, (T*)c->F, st);
        } else {
          if (wide)
            hipLaunchKernelGGL(vhk::k_fwd_band_d5, dim3(grid), dim3(256), lds, c->stream, a, (const T*)c->M,
                               (T*)c->y, (T*)c->F, st);
        }
        if (!wide)
          hipLaunchKernelGGL((vhk::k_fwd_band<T, false>), dim3(grid), dim3(256), lds, c->stream, a, (const T*)c->M,
                             (T*)c->y, (T*)c->F, st);
      }
    }
    mk.mark(c->stream);
  }
  {
    // The root band: upper_cols is empty when it is the only band.
    const int b = S.fwd_level_off[nb - 1], e = S.fwd_level_off[nb];
    const bool only = nb == 1;
    const bool slots = only && from_slots, cm = only && commit;
    a.tasks = c->tasks + b; a.ntasks = e - b; a.xper = band_xper(e - b, false);
    const vhk::CommitArgs cc = commit_args(c, cm, want_resid, add_acc);
    const size_t lds = 3 * (size_t)S.band_front[0] * c->tsz;
    const T* Sr = (const T*)c->Sroot;
    const int rg = band_task_blocks(e - b, false);
    bool dense = false;
    if constexpr (std::is_same<T, float>::value) {
      dense = !only && c->nroot_tasks > 0;
      if (dense) launch_root_dense(c, a, st);
    }
    if (!dense && e > b) {
      if (slots && cm)
        hipLaunchKernelGGL((vhk::k_root_band<T, true, true>), dim3(rg), dim3(256), lds, c->stream, a, cc, Sr,
                           (T*)c->F, (T*)c->V, (T*)c->d, st);
      else if (slots)
        hipLaunchKernelGGL((vhk::k_root_band<T, true, false>), dim3(rg), dim3(256), lds, c->stream, a, cc, Sr,
                           (T*)c->F, (T*)c->V, (T*)c->d, st);
      else if (cm)
        hipLaunchKernelGGL((vhk::k_root_band<T, false, true>), dim3(rg), dim3(256), lds, c->stream, a, cc, Sr,
                           (T*)c->F, (T*)c->V, (T*)c->d, st);
      else
        hipLaunchKernelGGL((vhk::k_root_band<T, false, false>), dim3(rg), dim3(256), lds, c->stream, a, cc, Sr,
                           (T*)c->F, (T*)c->V, (T*)c->d, st);
    }
    mk.mark(c->stream);
  }
  for (int l = 1; l < nb; ++l) {
    const bool cm = l == nb - 1 && commit;
    if constexpr (std::is_same<T, float>::value) {
      if (cm && c->nstc > 0) {
        a.tasks = c->stc_tasks; a.ntasks = c->nstc; a.xper = band_xper(c->nstc, false);
        const vhk::CommitArgs cc = commit_args(c, true, want_resid, add_acc);
        const int fmax = S.band_front[l];
        hipLaunchKernelGGL((vhk::k_bwd_band_stc<kStagedLoads, float>),
                           dim3(band_task_blocks(c->nstc, false) + nblk(nupper)), dim3(256), stagedc_lds(c, fmax),
                           c->stream, a, cc, (const float*)c->Mt, (const float*)c->y, (float*)c->V, (float*)c->d,
                           fmax, st);
        mk.mark(c->stream);
        continue;
      }
    }
    const int b = S.bwd_level_off[l], e = S.bwd_level_off[l + 1];
    a.tasks = c->tasks + nfwd + b; a.ntasks = e - b; a.xper = band_xper(e - b, false);
    const vhk::CommitArgs cc = commit_args(c, cm, want_resid, add_acc);
    const int grid = band_task_blocks(e - b, false) + (cm ? nblk(nupper) : 0);
    const size_t lds = 3 * (size_t)S.band_front[l] * c->tsz;
    if (grid > 0) {
      if (cm)
        hipLaunchKernelGGL((vhk::k_bwd_band<T, true>), dim3(grid), dim3(256), lds, c->stream, a, cc, (const T*)c->Mt,
                           (const T*)c->y, (T*)c->V, (T*)c->d, st);
      else {
        bool wide = false;
        if constexpr (std::is_same<T, double>::value) {
          wide = e - b > kWideBandTasks;
          if (wide)
            hipLaunchKernelGGL(vhk::k_bwd_band_d5, dim3(grid), dim3(256), lds, c->stream, a, cc, (const T*)c->Mt,
                               (const T*)c->y, (T*)c->V, (T*)c->d, st);
        }
        if (!wide)
          hipLaunchKernelGGL((vhk::k_bwd_band<T, false>), dim3(grid), dim3(256), lds, c->stream, a, cc,
                             (const T*)c->Mt, (const T*)c->y, (T*)c->V, (T*)c->d, st);
      }
    }
    mk.mark(c->stream);
  }
}

template <class T>
vhk::ElemArgs<T> elem_args(vh_ctx* c) {
  return vhk::ElemArgs<T>{c->mem, c->x, c->tri_idx, c->tri_dm, (const T*)c->tri_w, c->hinge_idx,
                          (const T*)c->hinge_rest, (const T*)c->hinge_w, (const T*)c->hinge_st};
}

template <class T>
void enqueue_local(vh_ctx* c, const vhk::StepDev* st) {
  const vh::Setup& S = c->S;
  if (c->local_blocks > 0)
    hipLaunchKernelGGL(vhk::k_local<T>, dim3(c->local_blocks), dim3(256), 0, c->stream, S.t, S.h, c->local_map,
                       elem_args<T>(c), vhk::LocalPlan{c->lv_ptr, c->lv_slot, c->lv_end, c->lc_pos}, (T*)c->cslot,
                       st);
}

// The local step of one PD iteration (k_local into the slots). Returns true:
// the solve then starts from the slots.
bool enqueue_local_step(vh_ctx* c, const vhk::StepDev* st) {
  if (c->prec == VH_PREC_F64) enqueue_local<double>(c, st); else enqueue_local<float>(c, st);
  return true;
}

// One correction of the iterate: x += A^{-1} (b - A x). FP64: one solve whose
// last launch commits. FP32: `sweeps` solves; between them the correction is
// accumulated in acc and r updated with the FP64 residual, and the last
// solve's last launch commits acc + d. Without commit (a standalone solve)
// every sweep accumulates into acc.
void enqueue_correction(vh_ctx* c, double dt, int sweeps, bool from_slots, bool commit, bool want_resid,
                        const vhk::StepDev* st, Marks& mk) {
  if (c->prec == VH_PREC_F64) {
    enqueue_solve<double>(c, dt, from_slots, commit, false, want_resid, st, mk);
    return;
  }
  for (int s = 0; s < sweeps; ++s) {
    const bool last = s + 1 == sweeps;
    enqueue_solve<float>(c, dt, from_slots && s == 0, last && commit, s > 0, last && want_resid, st, mk);
    if (!last || !commit) {
      hipLaunchKernelGGL(vhk::k_ir_update, dim3(nblk(c->S.n)), dim3(256), 0, c->stream, c->S.n, c->Ap, c->Ai, c->Ax,
                         (const float*)c->d, c->r, c->acc, s == 0 ? 1 : 0, last ? 0 : 1, st);
      mk.mark(c->stream);
    }
  }
}

void enqueue_step(vh_ctx* c, double dt) {
  const vh::Setup& S = c->S;
  const vh_pd_params& P = c->P;
  hipLaunchKernelGGL(vhk::k_predict, dim3(nblk(S.n)), dim3(256), 0, c->stream, S.n, c->x, c->v, c->prev, c->pred,
                     c->pinned, dt, P.gravity[0], P.gravity[1], P.gravity[2], P.has_ground, P.ground_height, c->st);
  // With tolerance 0 the loop never exits early: only the last iteration's
  // residual (StepResult.final_residual) is formed, and the kernels do not
  // read the early-exit flag.
  const vhk::StepDev* stk = P.tolerance > 0 ? c->st : nullptr;
  for (int it = 0; it < P.max_iterations; ++it) {
    const bool last = it + 1 == P.max_iterations;
    const bool want_resid = last || P.tolerance > 0;
    const bool from_slots = enqueue_local_step(c, stk);
    Marks none;
    enqueue_correction(c, dt, last ? P.ir_final_sweeps : P.ir_sweeps, from_slots, true, want_resid, stk, none);
    if (want_resid && !last)
      hipLaunchKernelGGL(vhk::k_resid_reduce, dim3(1), dim3(256), 0, c->stream, c->partials, c->resid_blocks,
                         P.tolerance, it, c->st);
  }
  // the last iteration's residual is formed by one extra block of k_finalize
  const double frayl = P.rayleigh > 0 ? 1.0 / (1.0 + P.rayleigh * dt) : 0.0;
  const vhk::ResidArgs rs{c->partials, c->resid_blocks, P.tolerance, P.max_iterations - 1,
                          P.max_iterations > 0 ? c->st : nullptr};
  hipLaunchKernelGGL(vhk::k_finalize, dim3(nblk(S.n) + (rs.st ? 1 : 0)), dim3(256), 0, c->stream, S.n, c->x, c->prev,
                     c->v, c->pinned, 1.0 / dt, P.has_ground, P.ground_height, 1.0 - P.damping, frayl, rs);
}

// Enqueues one step: replays the captured graph (captured on first use for
// this dt), or launches the kernels directly when graphs are disabled.
vh_status launch_step(vh_ctx* ctx, double dt) {
  if (!ctx->use_graph) {
    enqueue_step(ctx, dt);
    VH_HIP(hipGetLastError());
    return VH_OK;
  }
  if (!ctx->step_graph || ctx->graph_dt != dt) {
    drop_graph(ctx);
    hipGraph_t g = nullptr;
    VH_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    enqueue_step(ctx, dt);
    const hipError_t le = hipGetLastError();
    const hipError_t ce = hipStreamEndCapture(ctx->stream, &g);
    if (le != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return fail(ctx, VH_E_GPU, std::string("capture: ") + hipGetErrorString(le));
    }
    if (ce != hipSuccess) return fail(ctx, VH_E_GPU, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
    const hipError_t ie = hipGraphInstantiate(&ctx->step_graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) {
      ctx->step_graph = nullptr;
      return fail(ctx, VH_E_GPU, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
    }
    ctx->graph_dt = dt;
  }
  VH_HIP(hipGraphLaunch(ctx->step_graph, ctx->stream));
  return VH_OK;
}

// ---- GpuBackend primitives ----------------------------------------------

constexpr int kBackendBlocks = 1024;

// out = a + scale * b, rounded as the reference's separate multiply and add
// (no contraction into an FMA).
__global__ void __launch_bounds__(256) k_backend_axpy(size_t n, const float* __restrict__ a,
                                                      const float* __restrict__ b, float scale,
                                                      float* __restrict__ out) {
#pragma clang fp contract(off)
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float p = scale * b[i];
    out[i] = a[i] + p;
  }
}
__global__ void __launch_bounds__(256) k_backend_fill(size_t n, float* __restrict__ p, float value) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = value;
}
// Per-block sums of a[i] * b[i] in f64 (each product is exact in f64).
__global__ void __launch_bounds__(256) k_backend_dot(size_t n, const float* __restrict__ a,
                                                     const float* __restrict__ b, double* __restrict__ part) {
  __shared__ double sh[256];
  double s = 0.0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    s += (double)a[i] * (double)b[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}
__global__ void __launch_bounds__(256) k_backend_dot_sum(int nb, double* __restrict__ part) {
  __shared__ double sh[256];
  double s = 0.0;
  for (int k = threadIdx.x; k < nb; k += blockDim.x) s += part[k];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[kBackendBlocks] = sh[0];
}

vh_status backend_reserve(vh_ctx* ctx, size_t n) {
  if (!ctx->bk_part) VH_HIP(hipMalloc((void**)&ctx->bk_part, (kBackendBlocks + 1) * sizeof(double)));
  if (n <= ctx->bk_cap) return VH_OK;
  for (float** p : {&ctx->bk_a, &ctx->bk_b, &ctx->bk_o})
    if (*p) { (void)hipFree(*p); *p = nullptr; }
  ctx->bk_cap = 0;
  for (float** p : {&ctx->bk_a, &ctx->bk_b, &ctx->bk_o}) VH_HIP(hipMalloc((void**)p, n * sizeof(float)));
  ctx->bk_cap = n;
  return VH_OK;
}

unsigned int backend_blocks(size_t n) {
  return (unsigned int)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, kBackendBlocks));
}

// ---- allocation-free dot product ---------------------------------------------
// vh_backend_dot_noalloc keeps everything its kernels touch beyond the
// caller's inputs and result in one file-scope device array per GPU: the
// partial sums (kBackendBlocks + 1 doubles, padded to 256 bytes), then room
// for the two operands (kNoAllocMaxElems floats each). Host code never
// passes the array's name to a kernel: vh_create looks up its address on the
// device (hipGetSymbolAddress) once per device and keeps it in
// g_noalloc_scratch; every use goes through that address.
constexpr uint64_t kNoAllocMaxElems = 1ull << 21;  // 2 Mi floats per operand
constexpr size_t kNoAllocPartBytes = ((kBackendBlocks + 1) * sizeof(double) + 255) / 256 * 256;
constexpr size_t kNoAllocBytes = kNoAllocPartBytes + 2 * kNoAllocMaxElems * sizeof(float);
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];
constexpr int kMaxDevices = 64;
void* g_noalloc_scratch[kMaxDevices] = {};

// Looks up this device's copy of `scratch` (the device is current). Called
// by vh_create; later calls for the same device return at once.
hipError_t noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  static std::mutex mu;
  std::lock_guard<std::mutex> lock(mu);
  if (g_noalloc_scratch[device]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (!p) return hipErrorInvalidSymbol;
  g_noalloc_scratch[device] = p;
  return hipSuccess;
}

// ---- contact step -------------------------------------------------------

// Blocks of the grid-stride kernels over candidate pairs: enough to fill the
// chip, independent of the pair capacity (a launch whose work is skipped
// stays cheap).
constexpr unsigned int kMaxPairBlocks = 4096;
unsigned int pair_blocks(unsigned int cap) {
  const unsigned int b = std::max(1u, std::min((cap + 255) / 256, kMaxPairBlocks));
  return (b + 7) / 8 * 8;  // a multiple of 8: pair_range gives each XCD an equal share of blocks
}

// The most pairs one block of a pair_blocks(cap) grid visits (pair_range
// over at most cap pairs): the survivor region of a k_ccd_cull block.
unsigned int ccd_region(unsigned int cap) {
  const unsigned int span = pair_blocks(cap) / 8 * 256, per = (cap + 7) / 8;
  return std::max(1u, (per + span - 1) / span) * 256;
}

const int* ipc_done_flag(vh_ctx* c) {
  return reinterpret_cast<const int*>(reinterpret_cast<const char*>(c->ipc.ip) + offsetof(vhc::IpcDev, al_done));
}

__global__ void k_ipc_clear(vhc::IpcDev* ip) {
  if (threadIdx.x != 0) return;
  ip->al_done = 0; ip->pd_done = 0; ip->overflow = 0; ip->toi = vhc::kNoToi;
  ip->viol = 0; ip->active = 0; ip->self_contacts = 0;
}

__global__ void k_ipc_clear_overflow(vhc::IpcDev* ip) {
  if (threadIdx.x == 0) ip->overflow = 0;
}

// Replaces a device buffer tracked in ctx->allocs by one of `bytes` bytes
// (contents not kept). The caller has synchronized the stream.
hipError_t realloc_tracked(vh_ctx* c, void** p, size_t bytes) {
  auto it = std::find(c->allocs.begin(), c->allocs.end(), *p);
  if (it != c->allocs.end()) c->allocs.erase(it);
  hipError_t e = *p ? hipFree(*p) : hipSuccess;
  *p = nullptr;
  if (e != hipSuccess) return e;
  e = hipMalloc(p, std::max<size_t>(bytes, 8));
  if (e != hipSuccess) {
    *p = nullptr;
    return e;
  }
  c->allocs.push_back(*p);
  return hipSuccess;
}

// Replaces the self-contact entry buffers by ones with room for `need`
// contacts and half again, and clears the overflow flag (as grow_pairs).
hipError_t grow_contacts(vh_ctx* c, unsigned int need) {
  IpcState& I = c->ipc;
  const uint64_t cap = std::min<uint64_t>((uint64_t)need + need / 2 + 256, 1ull << 29);
  if (need > cap) return hipErrorOutOfMemory;
  I.ecap = 0;  // with no room, k_self_emit writes nothing
  hipError_t e = realloc_tracked(c, (void**)&I.ev, 4 * cap * sizeof(int));
  if (e == hipSuccess) e = realloc_tracked(c, (void**)&I.eg, 12 * cap * sizeof(double));
  if (e == hipSuccess) e = realloc_tracked(c, (void**)&I.vidx, 4 * cap * sizeof(unsigned int));
  if (e != hipSuccess) return e;
  I.ecap = (unsigned int)cap;
  ++I.gen;
  hipLaunchKernelGGL(k_ipc_clear_overflow, dim3(1), dim3(64), 0, c->stream, I.ip);
  return hipGetLastError();
}

// Replaces the candidate-pair buffer by one with room for `need` pairs and
// half again, and clears the overflow flag. The caller has synchronized the
// stream (the old buffer is freed). The reference collects its pairs in a
// growable Vec (bvh.rs:207-209), so no pair count is ever refused.
hipError_t grow_pairs(vh_ctx* c, unsigned int need) {
  IpcState& I = c->ipc;
  const uint64_t cap = std::min<uint64_t>((uint64_t)need + need / 2 + 256, 1ull << 31);
  if (need > cap) return hipErrorOutOfMemory;
  auto it = std::find(c->allocs.begin(), c->allocs.end(), static_cast<void*>(I.pairs));
  if (it != c->allocs.end()) c->allocs.erase(it);
  hipError_t e = hipFree(I.pairs);
  I.pairs = nullptr;
  I.cap = 0;
  if (e != hipSuccess) return e;
  void* q = nullptr;
  e = hipMalloc(&q, cap * sizeof(int2));
  if (e != hipSuccess) return e;
  c->allocs.push_back(q);
  I.pairs = static_cast<int2*>(q);
  I.cap = (unsigned int)cap;
  it = std::find(c->allocs.begin(), c->allocs.end(), static_cast<void*>(I.ccd_surv));
  if (it != c->allocs.end()) c->allocs.erase(it);
  e = hipFree(I.ccd_surv);
  I.ccd_surv = nullptr;
  const unsigned int reg = ccd_region(I.cap);
  q = nullptr;
  if (e == hipSuccess) e = hipMalloc(&q, (size_t)pair_blocks(I.cap) * reg * sizeof(int2));
  if (e != hipSuccess) {
    I.cap = 0;  // with no survivor regions, no pair kernel may visit a pair
    return e;
  }
  c->allocs.push_back(q);
  I.ccd_surv = static_cast<int2*>(q);
  I.ccd_reg = reg;
  // the self-contact offsets share the survivor regions' layout
  e = realloc_tracked(c, (void**)&I.cof, (size_t)pair_blocks(I.cap) * reg * sizeof(unsigned int));
  if (e != hipSuccess) {
    I.cap = 0;
    return e;
  }
  ++I.gen;
  hipLaunchKernelGGL(k_ipc_clear_overflow, dim3(1), dim3(64), 0, c->stream, I.ip);
  return hipGetLastError();
}

// The self part of the CCD over the candidate pairs of the last detection,
// from x to x + dl: the cull, then the tests of its survivors.
void enqueue_ccd_pairs(vh_ctx* c, const double* x, const double* dl) {
  IpcState& I = c->ipc;
  const int t = (int)c->S.t;
  hipLaunchKernelGGL(vhc::k_ccd_tri_bounds, dim3(std::max(1, nblk(t))), dim3(256), 0, c->stream, t, c->tri_idx, x, dl,
                     I.ccdtri, I.ip);
  const int ee = I.P.ccd_edge_edge ? 1 : 0;
  hipLaunchKernelGGL(vhc::k_ccd_cull, dim3(pair_blocks(I.cap)), dim3(256), 0, c->stream, I.pairs, &I.bd->npairs,
                     I.cap, I.ccdtri, I.ccd_surv, I.ccd_reg, I.ccd_cnt, I.ip, ee);
  if (ee)
    hipLaunchKernelGGL(vhc::k_ccd_pairs<true>, dim3(pair_blocks(I.cap)), dim3(256), 0, c->stream, I.ccd_surv,
                       I.ccd_reg, I.ccd_cnt, c->tri_idx, x, dl, I.ip);
  else
    hipLaunchKernelGGL(vhc::k_ccd_pairs<false>, dim3(pair_blocks(I.cap)), dim3(256), 0, c->stream, I.ccd_surv,
                       I.ccd_reg, I.ccd_cnt, c->tri_idx, x, dl, I.ip);
}

// After a candidate-pair query has been enqueued: reads the pair count and
// the self-contact count (one 8-byte read and a synchronization) and, when
// one exceeded its buffer, grows it and enqueues `redo`, which repeats the
// query and what consumed it. The counts are functions of the positions, so
// a regrowth per buffer suffices. contacts: the enqueued work includes a
// self-contact detection. `flag`, when given, is read into h_done with the
// counts.
template <class F>
hipError_t fit_pairs(vh_ctx* c, F&& redo, bool contacts, const int* flag = nullptr) {
  IpcState& I = c->ipc;
  static_assert(offsetof(vhc::BvhDev, ncont) == offsetof(vhc::BvhDev, npairs) + sizeof(unsigned int),
                "npairs and ncont are read together");
  for (int pass = 0; pass < 4; ++pass) {
    hipError_t e =
        hipMemcpyAsync(I.h_np, &I.bd->npairs, 2 * sizeof(unsigned int), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && flag) e = hipMemcpyAsync(I.h_done, flag, sizeof(int), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return e;
    const bool pairs_over = I.h_np[0] > I.cap;
    const bool cont_over = contacts && !pairs_over && I.h_np[1] > I.ecap;
    if (!pairs_over && !cont_over) return hipSuccess;
    e = pairs_over ? grow_pairs(c, I.h_np[0]) : grow_contacts(c, I.h_np[1]);
    if (e != hipSuccess) return e;
    redo();
  }
  return hipErrorUnknown;
}

// Levels of the sparse table over nc whole cells: 0 .. floor(log2 nc).
int coarse_levels(int nc) {
  int m = 0;
  while (nc > 0 && (1 << (m + 1)) <= nc) ++m;
  return nc > 0 ? m + 1 : 0;
}

// The leaf records, after the t - 1 internal nodes in the same allocation.
vhc::Rec* leaf_of(vh_ctx* c) { return c->ipc.node + (c->S.t > 0 ? c->S.t - 1 : 0); }

// Candidate triangle pairs of positions x (boxes inflated by margin).
void enqueue_pairs(vh_ctx* c, const double* x, float margin, const int* skip) {
  IpcState& I = c->ipc;
  const int t = c->S.t;
  hipLaunchKernelGGL(vhc::k_bvh_reset, dim3(1), dim3(64), 0, c->stream, I.bd, skip);
  if (t == 0) return;
  hipLaunchKernelGGL(vhc::k_boxes, dim3(nblk(t)), dim3(256), 0, c->stream, t, c->tri_idx, x, margin, I.box,
                     I.bpart, skip);
  hipLaunchKernelGGL(vhc::k_scene_bounds, dim3(1), dim3(256), 0, c->stream, nblk(t), I.bpart, I.bd, skip);
  hipLaunchKernelGGL(vhc::k_morton, dim3(nblk(t)), dim3(256), 0, c->stream, t, I.box, I.bd, I.code, I.id, skip);
  // stable, so equal codes keep the triangle order
  const int in_s = vhs::sort_pairs(c->stream, I.code, I.id, I.code_s, I.id_s, t, 30, I.sort_hist, skip);
  const unsigned int* code_s = in_s ? I.code_s : I.code;
  const int* id_s = in_s ? I.id_s : I.id;
  // sparse-table levels 0 .. low - 1 in one launch
  const int low = t < 2 ? 1 : std::min(I.st_levels, vhc::kStLowLevels + 1);
  hipLaunchKernelGGL(vhc::k_st_low, dim3(nblk(t)), dim3(256), 0, c->stream, t, low, I.box, id_s, I.st, skip);
  if (t < 2) return;
  hipLaunchKernelGGL(vhc::k_karras, dim3(nblk(t - 1)), dim3(256), 0, c->stream, t, code_s, I.left, I.range_lo,
                     I.range_hi, skip);
  const int nc = t >> 8;  // whole 256-leaf cells; ranges of 512 leaves or more use them
  if (nc >= 2)
    hipLaunchKernelGGL(vhc::k_st_coarse, dim3(1), dim3(vhc::kCoarseThreads), 0, c->stream, t, nc, coarse_levels(nc),
                       I.st + (size_t)vhc::kStLowLevels * t, I.coarse, skip);
  hipLaunchKernelGGL(vhc::k_node_bounds, dim3(nblk(t)), dim3(256), 0, c->stream, t, I.st, I.coarse, nc, id_s, I.left,
                     I.range_lo, I.range_hi, I.node, leaf_of(c), I.orig, I.sorig, skip);
  const int walk_blocks = (nblk(t) + 7) / 8 * 8;  // see leaf_chunk
  hipLaunchKernelGGL(vhc::k_bvh_walk1, dim3(walk_blocks), dim3(256), 0, c->stream, t, I.node, leaf_of(c), I.sorig,
                     I.pcount, I.wkept, I.wmore, I.wmore_leaf, I.bd, skip);
  vhs::scan_exclusive(c->stream, I.pcount, I.poff, t, I.sbuf, I.ticket, skip);
  // the kept pairs copied to the list; a leaf with more is walked again
  hipLaunchKernelGGL(vhc::k_bvh_compact, dim3(nblk(t)), dim3(256), 0, c->stream, t, I.pcount, I.poff, I.wkept,
                     I.wmore, I.wmore_leaf, I.pairs, I.cap, I.bd, skip);
  hipLaunchKernelGGL(vhc::k_bvh_more, dim3(walk_blocks), dim3(256), 0, c->stream, t, I.node, leaf_of(c), I.sorig,
                     I.pcount, I.poff, I.wmore_leaf, I.pairs, I.cap, I.bd, skip);
}

// CollisionPipeline::detect_ipc_contacts at positions x with scene sc.
void enqueue_detect(vh_ctx* c, const double* x, const vhc::Scene& sc, bool reset_lambda, const int* skip,
                    double* hess = nullptr) {
  IpcState& I = c->ipc;
  const int n = c->S.n;
  hipLaunchKernelGGL(vhc::k_ipc_detect_begin, dim3(1), dim3(64), 0, c->stream, I.ip);
  if (sc.self_collision) {
    enqueue_pairs(c, x, (float)sc.thickness, skip);
    hipLaunchKernelGGL(vhc::k_tri_spheres, dim3(std::max(1, nblk(c->S.t))), dim3(256), 0, c->stream, c->S.t,
                       c->tri_idx, x, I.ccdtri, skip);
    // the survivor regions are shared with the CCD, which runs after the detection
    const unsigned int pb = pair_blocks(I.cap);
    hipLaunchKernelGGL(vhc::k_self_cull, dim3(pb), dim3(256), 0, c->stream, I.pairs, &I.bd->npairs, I.cap, sc,
                       I.ccdtri, I.ccd_surv, I.ccd_reg, I.ccd_cnt, skip);
    // the gradients in a fixed order: count, offsets, entries, by vertex
    hipLaunchKernelGGL(vhc::k_self_count, dim3(pb), dim3(256), 0, c->stream, &I.bd->npairs, I.cap, I.ccd_surv,
                       I.ccd_reg, I.ccd_cnt, c->tri_idx, x, I.xptr, I.xidx, sc, I.cof, I.btot, I.ip, skip);
    hipLaunchKernelGGL(vhc::k_self_base, dim3(1), dim3(256), 0, c->stream, (int)pb, I.btot, I.sbase, I.ecap, I.bd,
                       I.ip, skip);
    hipLaunchKernelGGL(vhc::k_self_emit, dim3(pb), dim3(256), 0, c->stream, I.ccd_surv, I.ccd_reg, I.ccd_cnt, I.cof,
                       I.sbase, c->tri_idx, x, I.xptr, I.xidx, sc, I.ecap, I.ev, I.eg, skip);
    hipLaunchKernelGGL(vhc::k_fill_u, dim3(nblk(n + 1)), dim3(256), 0, c->stream, n + 1, I.vcnt, skip);
    hipLaunchKernelGGL(vhc::k_fill_u, dim3(nblk(n + 1)), dim3(256), 0, c->stream, n + 1, I.vfill, skip);
    hipLaunchKernelGGL(vhc::k_vcount, dim3(1024), dim3(256), 0, c->stream, I.bd, I.ecap, I.ev, I.vcnt, skip);
    vhs::scan_exclusive(c->stream, I.vcnt, I.vptr, (long long)n + 1, I.sbuf, I.ticket, skip);
    hipLaunchKernelGGL(vhc::k_vscatter, dim3(1024), dim3(256), 0, c->stream, I.bd, I.ecap, I.ev, I.vptr, I.vfill,
                       I.vidx, skip);
    hipLaunchKernelGGL(vhc::k_vgather, dim3(nblk(n)), dim3(256), 0, c->stream, n, I.vptr, I.vidx, I.eg, I.gself,
                       skip);
  }
  hipLaunchKernelGGL(vhc::k_vertex_contacts, dim3(nblk(n)), dim3(256), 0, c->stream, n, x, sc, I.gself, I.grad, I.nrm,
                     I.inc, I.ever, I.fnrm, I.lambda, reset_lambda ? 1 : 0, I.ip, hess, skip);
}

// The PD loop of one AL iteration (pd_solver.rs:660-1000): the frozen
// contact term, then max_iterations x (local step, solve, step fraction
// from the CCD, update); every launch skips once the device flags say the
// loop has converged, so the sequence is fixed and is replayed as a graph.
void enqueue_al_pd(vh_ctx* c, double dt) {
  IpcState& I = c->ipc;
  const vh_pd_params& P = c->P;
  const int n = c->S.n;
  const double* dl = c->prec == VH_PREC_F64 ? (const double*)c->d : c->acc;
  hipLaunchKernelGGL(vhc::k_ipc_ext, dim3(nblk(3 * n)), dim3(256), 0, c->stream, n, I.grad, I.lambda, I.ext, I.ip);
  c->cur_ext = I.ext;
  for (int it = 0; it < P.max_iterations; ++it) {
    const bool from_slots = enqueue_local_step(c, c->st);
    Marks none;
    const int sweeps = it + 1 == P.max_iterations ? P.ir_final_sweeps : P.ir_sweeps;
    enqueue_correction(c, dt, sweeps, from_slots, false, false, c->st, none);
    hipLaunchKernelGGL(vhc::k_ipc_sums, dim3(nblk(n)), dim3(256), 0, c->stream, n, c->x, dl, I.grad, c->pinned,
                       I.partials, I.ip);
    if (I.sc.self_collision) enqueue_ccd_pairs(c, c->x, dl);
    if (I.sc.has_ground || I.sc.has_sphere)
      hipLaunchKernelGGL(vhc::k_ccd_vertices, dim3(nblk(n)), dim3(256), 0, c->stream, n, c->x, dl, I.sc, I.ip);
    hipLaunchKernelGGL(vhc::k_ipc_alpha, dim3(1), dim3(256), 0, c->stream, I.partials, nblk(n), P.tolerance, I.ip);
    hipLaunchKernelGGL(vhc::k_ipc_update, dim3(nblk(n)), dim3(256), 0, c->stream, n, c->x, dl, c->pinned, I.ip);
    hipLaunchKernelGGL(vhc::k_ipc_pd_latch, dim3(1), dim3(64), 0, c->stream, I.ip, c->st);
  }
  c->cur_ext = nullptr;
}

// Enqueues the launches of `enqueue` (kernels only: no memset or library
// call) by replaying q's graph, captured on first use and again after dt or
// a pair buffer changed; or launches them directly when graphs are disabled.
template <class F>
hipError_t launch_seq(vh_ctx* c, IpcState::Seq& q, double dt, F&& enqueue) {
  IpcState& I = c->ipc;
  if (!c->use_graph) {
    enqueue();
    return hipGetLastError();
  }
  if (!q.exec || q.dt != dt || q.gen != I.gen) {
    if (q.exec) (void)hipGraphExecDestroy(q.exec);
    q.exec = nullptr;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    enqueue();
    const hipError_t le = hipGetLastError();
    const hipError_t ce = hipStreamEndCapture(c->stream, &g);
    if (le != hipSuccess || ce != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return le != hipSuccess ? le : ce;
    }
    e = hipGraphInstantiate(&q.exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      q.exec = nullptr;
      return e;
    }
    q.dt = dt;
    q.gen = I.gen;
  }
  return hipGraphLaunch(q.exec, c->stream);
}

// step_with_ipc (pd_solver.rs:549-1157) as three replayed graphs: the head
// of the step, then one per AL iteration. The PD loop has a fixed trip count
// and stops through device flags; after each AL iteration the host reads the
// AL flag (and, with self-collision, the pair and contact counts) and
// launches no further AL iteration once it is set (the kernels of one would
// all be skipped).
hipError_t enqueue_ipc_step(vh_ctx* c, double dt) {
  IpcState& I = c->ipc;
  const vh::Setup& S = c->S;
  const vh_pd_params& P = c->P;
  const int n = S.n;
  const vh_ipc_params& Q = I.P;
  const int* skip = ipc_done_flag(c);
  const double max_vel = 0.5 * std::sqrt(I.cfl_dh) / dt;
  hipError_t e = launch_seq(c, I.seq[0], dt, [&] {
    hipLaunchKernelGGL(vhc::k_ipc_begin, dim3(nblk(n)), dim3(256), 0, c->stream, n, c->x, c->v, c->prev, c->pred,
                       c->pinned, I.ever, dt, P.gravity[0], P.gravity[1], P.gravity[2], max_vel, Q.has_ground,
                       Q.ground_height, Q.al_mu_initial, I.ip, c->st);
    enqueue_detect(c, c->x, I.sc, true, skip);
  });
  if (e != hipSuccess) return e;
  if (I.sc.self_collision) {
    e = fit_pairs(c, [&] { enqueue_detect(c, c->x, I.sc, true, skip); }, true);
    if (e != hipSuccess) return e;
  }
  for (int al = 0; al < Q.al_max_iterations; ++al) {
    const bool last = al + 1 == Q.al_max_iterations;
    // The AL update: detection at the PD result, multiplier update, decision.
    // After a pair overflow the last two do nothing and all three are redone
    // with a larger buffer; the pair count is read with the AL flag.
    auto al_close = [&] {
      enqueue_detect(c, c->x, I.sc, false, skip);
      hipLaunchKernelGGL(vhc::k_ipc_al_end, dim3(nblk(n)), dim3(256), 0, c->stream, n, I.grad, I.inc, I.lambda,
                         Q.al_tolerance, last ? 0 : 1, I.ip);
      hipLaunchKernelGGL(vhc::k_ipc_al_commit, dim3(1), dim3(64), 0, c->stream, Q.al_tolerance, Q.al_mu_growth,
                         last ? 1 : 0, I.ip, c->st);
    };
    e = launch_seq(c, I.seq[last ? 2 : 1], dt, [&] {
      enqueue_al_pd(c, dt);
      al_close();
    });
    if (e != hipSuccess) return e;
    if (I.sc.self_collision) {
      e = fit_pairs(c, al_close, true, skip);
      if (e != hipSuccess) return e;
      if (*I.h_done) break;
    } else if (!last) {
      e = hipMemcpyAsync(I.h_done, skip, sizeof(int), hipMemcpyDeviceToHost, c->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
      if (e != hipSuccess) return e;
      if (*I.h_done) break;
    }
  }
  vhc::FinalArgs f{};
  f.inv_dt = 1.0 / dt; f.mu_f = Q.friction_coefficient; f.cdamp = Q.contact_damping;
  f.cdamp_max = Q.contact_damping_max; f.v_thresh = Q.contact_velocity_threshold;
  f.one_minus_damp = 1.0 - P.damping; f.frayl = P.rayleigh > 0 ? 1.0 / (1.0 + P.rayleigh * dt) : 0.0;
  f.ground = Q.ground_height; f.adaptive = Q.adaptive_contact_damping; f.has_ground = Q.has_ground;
  hipLaunchKernelGGL(vhc::k_ipc_finalize, dim3(nblk(n)), dim3(256), 0, c->stream, n, c->x, c->prev, c->v, c->pinned,
                     I.ever, I.fnrm, f, I.ip);
  return hipGetLastError();
}

// Reads the step result and the contact diagnostics (synchronizes).
vh_status read_ipc_result(vh_ctx* ctx, vh_step_result* out, vh_ipc_stats* st) {
  vhc::IpcDev h{};
  unsigned int np = 0;
  VH_HIP(hipMemcpyAsync(&h, ctx->ipc.ip, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  VH_HIP(hipMemcpyAsync(&np, &ctx->ipc.bd->npairs, sizeof(np), hipMemcpyDeviceToHost, ctx->stream));
  VH_HIP(hipStreamSynchronize(ctx->stream));
  if (out) {
    out->iterations = (uint32_t)h.iterations;
    out->final_residual = h.resid;
    out->converged = (h.converged || h.resid < ctx->P.tolerance) ? 1 : 0;
  }
  if (st) {
    st->al_iterations = h.al_iters; st->al_converged = h.converged; st->mu = h.mu;
    st->max_violation = __builtin_bit_cast(double, h.viol);
    st->active = (int64_t)h.active; st->self_contacts = (int64_t)h.self_contacts;
    st->pairs = ctx->ipc.sc.self_collision ? (int64_t)np : 0; st->last_alpha = h.last_alpha;
  }
  if (h.overflow) return fail(ctx, VH_E_INTERNAL, "candidate pairs or self contacts exceeded their buffers");
  if (!(h.resid == h.resid)) return fail(ctx, VH_E_NONFINITE, "non-finite positions in the IPC step");
  return VH_OK;
}

// Permuted AoS [n][3] <-> caller SoA (x[n], y[n], z[n]).
void to_dev_order(const vh::Setup& S, const double* x, const double* y, const double* z, std::vector<double>& out) {
  out.resize(3 * (size_t)S.n);
  for (int k = 0; k < S.n; ++k) {
    const int o = S.perm[k];
    out[3 * k] = x[o]; out[3 * k + 1] = y[o]; out[3 * k + 2] = z[o];
  }
}
void from_dev_order(const vh::Setup& S, const std::vector<double>& in, double* x, double* y, double* z) {
  for (int k = 0; k < S.n; ++k) {
    const int o = S.perm[k];
    x[o] = in[3 * k]; y[o] = in[3 * k + 1]; z[o] = in[3 * k + 2];
  }
}

double* state_ptr(vh_ctx* c, int which) {
  switch (which) {
    case 0: return c->x;
    case 1: return c->v;
    case 2: return c->prev;
    case 3: return c->pred;
    default: return nullptr;
  }
}

vh_status read_step_result(vh_ctx* ctx, vh_step_result* out) {
  vhk::StepDev h{};
  VH_HIP(hipMemcpyAsync(&h, ctx->st, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  VH_HIP(hipStreamSynchronize(ctx->stream));
  if (out) {
    out->iterations = (uint32_t)h.iterations;
    out->final_residual = h.residual;
    out->converged = h.residual < ctx->P.tolerance ? 1 : 0;
  }
  if (!(h.residual == h.residual)) return fail(ctx, VH_E_NONFINITE, "non-finite positions in the PD step");
  return VH_OK;
}

// FaerSolver::factorize (faer_solver.rs:67-93) on the device: the level
// loop of vh_factor.hpp, deepest level first, writing M, Mt and Sroot in the
// compute type. A level's fronts live until its parent level has assembled
// them. Synchronizes; VH_E_NOT_SPD when a pivot is not positive.
struct FactorTemps {
  std::vector<void*> p;
  ~FactorTemps() {
    for (void* q : p) (void)hipFree(q);
  }
  void* take(void* q) {
    p.push_back(q);
    return q;
  }
  void drop(void* q) {
    auto it = std::find(p.begin(), p.end(), q);
    if (it != p.end()) p.erase(it);
    (void)hipFree(q);
  }
};

vh_status device_factor(vh_ctx* ctx) {
  const vh::Setup& S = ctx->S;
  const int ns = (int)S.sn.size();
  const int nb = S.max_depth + 1;
  constexpr int FT = vhf::kFT;
  auto up32 = [](int v) { return (v + FT - 1) / FT * FT; };
  std::vector<vhf::FrontDev> fr(ns);
  std::vector<std::vector<int>> lev(nb);
  std::vector<long long> fsz(nb, 0), esz(nb, 0);
  for (int s = 0; s < ns; ++s) {
    const vh::Supernode& x = S.sn[s];
    vhf::FrontDev& f = fr[s];
    f.w = x.w; f.r = x.r; f.wp = up32(x.w); f.mp = f.wp + up32(x.r);
    f.f = fsz[x.depth]; fsz[x.depth] += (long long)f.mp * f.mp;
    f.e = esz[x.depth]; esz[x.depth] += (long long)f.mp * f.wp;
    f.m_off = x.m_off; f.s_off = x.parent < 0 ? x.s_off : 0;
    f.col0 = x.col0; f.rows_off = x.rows_off; f.child_off = x.child_off; f.nchild = x.nchild; f.u_off = x.u_off;
    f.lpos = (int)lev[x.depth].size();
    lev[x.depth].push_back(s);
  }
  // Launch lists: per level, per panel k, the supernodes with a panel k and
  // the tile tasks of the two tile launches; the roots' inverse tasks last.
  std::vector<int> ints;  // level lists and diag lists
  std::vector<int4> tasks;
  struct L3 { size_t diag_off, trsm_off, upd_off; int ndiag, ntrsm, nupd; };
  std::vector<std::vector<L3>> plan(nb);
  std::vector<size_t> lev_off(nb);
  size_t root_off = 0, root_tasks_off = 0;
  int nroot_tasks = 0;
  for (int d = 0; d < nb; ++d) {
    lev_off[d] = ints.size();
    ints.insert(ints.end(), lev[d].begin(), lev[d].end());
    int kmax = 0;
    for (int s : lev[d]) kmax = std::max(kmax, fr[s].wp / FT);
    for (int k = 0; k < kmax; ++k) {
      L3 L{};
      L.diag_off = ints.size();
      for (int s : lev[d])
        if (fr[s].wp / FT > k) ints.push_back(s);
      L.ndiag = (int)(ints.size() - L.diag_off);
      L.trsm_off = tasks.size();
      for (int s : lev[d]) {
        const int T = fr[s].mp / FT;
        if (fr[s].wp / FT <= k) continue;
        for (int i = k + 1; i < T; ++i) tasks.push_back(make_int4(s, i, k, vhf::kFTrsm));
        for (int c = 0; c <= k; ++c) tasks.push_back(make_int4(s, k, c, vhf::kETrsm));
      }
      L.ntrsm = (int)(tasks.size() - L.trsm_off);
      L.upd_off = tasks.size();
      for (int s : lev[d]) {
        const int T = fr[s].mp / FT;
        if (fr[s].wp / FT <= k) continue;
        for (int i = k + 1; i < T; ++i) {
          for (int j = k + 1; j <= i; ++j) tasks.push_back(make_int4(s, i, j, vhf::kFUpd));
          for (int c = 0; c <= k; ++c) tasks.push_back(make_int4(s, i, c, vhf::kEUpd));
        }
      }
      L.nupd = (int)(tasks.size() - L.upd_off);
      plan[d].push_back(L);
    }
  }
  root_off = ints.size();
  for (int s : lev[0])
    if (S.sn[s].parent < 0) ints.push_back(s);
  const int nroots = (int)(ints.size() - root_off);
  root_tasks_off = tasks.size();
  for (int s : lev[0]) {
    if (S.sn[s].parent >= 0) continue;
    const int T = fr[s].wp / FT;
    for (int i = 0; i < T; ++i)
      for (int j = 0; j < T; ++j) tasks.push_back(make_int4(s, i, j, vhf::kSRoot));
  }
  nroot_tasks = (int)(tasks.size() - root_tasks_off);

  FactorTemps tmp;
  auto dmalloc = [&](size_t bytes, void** q) -> hipError_t {
    *q = nullptr;
    const hipError_t e = hipMalloc(q, std::max<size_t>(bytes, 8));
    if (e == hipSuccess) tmp.take(*q);
    return e;
  };
  void *d_fr, *d_ints, *d_tasks, *d_rowidx, *d_children, *d_emap, *d_fail;
  VH_HIP(dmalloc(fr.size() * sizeof(vhf::FrontDev), &d_fr));
  VH_HIP(dmalloc(ints.size() * sizeof(int), &d_ints));
  VH_HIP(dmalloc(tasks.size() * sizeof(int4), &d_tasks));
  VH_HIP(dmalloc(S.rowidx.size() * sizeof(int), &d_rowidx));
  VH_HIP(dmalloc(S.children.size() * sizeof(int), &d_children));
  VH_HIP(dmalloc(S.emap.size() * sizeof(int), &d_emap));
  VH_HIP(dmalloc(sizeof(int), &d_fail));
  VH_HIP(hipMemcpy(d_fr, fr.data(), fr.size() * sizeof(vhf::FrontDev), hipMemcpyHostToDevice));
  if (!ints.empty()) VH_HIP(hipMemcpy(d_ints, ints.data(), ints.size() * sizeof(int), hipMemcpyHostToDevice));
  if (!tasks.empty()) VH_HIP(hipMemcpy(d_tasks, tasks.data(), tasks.size() * sizeof(int4), hipMemcpyHostToDevice));
  if (!S.rowidx.empty())
    VH_HIP(hipMemcpy(d_rowidx, S.rowidx.data(), S.rowidx.size() * sizeof(int), hipMemcpyHostToDevice));
  if (!S.children.empty())
    VH_HIP(hipMemcpy(d_children, S.children.data(), S.children.size() * sizeof(int), hipMemcpyHostToDevice));
  if (!S.emap.empty()) VH_HIP(hipMemcpy(d_emap, S.emap.data(), S.emap.size() * sizeof(int), hipMemcpyHostToDevice));
  VH_HIP(hipMemset(d_fail, 0, sizeof(int)));
  const vhf::FrontDev* frp = static_cast<const vhf::FrontDev*>(d_fr);
  const int* ip = static_cast<const int*>(d_ints);
  const int4* tp = static_cast<const int4*>(d_tasks);
  hipStream_t st = ctx->stream;
  void *Fc = nullptr, *Ec = nullptr, *Dc = nullptr;  // the child level's buffers
  for (int d = nb - 1; d >= 0; --d) {
    const int nl = (int)lev[d].size();
    void *Fd, *Ed, *Dd;
    VH_HIP(dmalloc(fsz[d] * sizeof(double), &Fd));
    VH_HIP(dmalloc(esz[d] * sizeof(double), &Ed));
    VH_HIP(dmalloc((size_t)std::max(nl, 1) * FT * FT * sizeof(double), &Dd));
    VH_HIP(hipMemsetAsync(Fd, 0, fsz[d] * sizeof(double), st));
    VH_HIP(hipMemsetAsync(Ed, 0, esz[d] * sizeof(double), st));
    if (nl > 0)
      hipLaunchKernelGGL(vhf::k_fac_assemble, dim3(nl), dim3(256), 0, st, frp, ip + lev_off[d], ctx->Ap, ctx->Ai,
                         ctx->Ax, (const int*)d_rowidx, (const int*)d_children, (const int*)d_emap, (double*)Fd,
                         (double*)Ed, (const double*)Fc);
    VH_HIP(hipGetLastError());
    if (Fc) {  // the children are assembled into this level: their fronts can go
      VH_HIP(hipStreamSynchronize(st));
      tmp.drop(Fc); tmp.drop(Ec); tmp.drop(Dc);
    }
    for (size_t k = 0; k < plan[d].size(); ++k) {
      const L3& L = plan[d][k];
      hipLaunchKernelGGL(vhf::k_fac_diag, dim3(L.ndiag), dim3(256), 0, st, frp, ip + L.diag_off, (int)k, (double*)Fd,
                         (double*)Dd, (int*)d_fail);
      if (L.ntrsm > 0)
        hipLaunchKernelGGL(vhf::k_fac_tiles, dim3(L.ntrsm), dim3(64), 0, st, frp, tp + L.trsm_off, (int)k,
                           (double*)Fd, (double*)Ed, (const double*)Dd);
      if (L.nupd > 0)
        hipLaunchKernelGGL(vhf::k_fac_tiles, dim3(L.nupd), dim3(64), 0, st, frp, tp + L.upd_off, (int)k,
                           (double*)Fd, (double*)Ed, (const double*)Dd);
    }
    if (nl > 0) {
      // at least ~2048 workgroups per level: a level of few large
      // supernodes splits each one's rows over several
      const dim3 grid(nl, std::max(1, std::min(256, 2048 / nl)));
      if (ctx->prec == VH_PREC_F64)
        hipLaunchKernelGGL(vhf::k_fac_blocks<double>, grid, dim3(256), 0, st, frp, ip + lev_off[d],
                           (const double*)Ed, (double*)ctx->M, (double*)ctx->Mt);
      else
        hipLaunchKernelGGL(vhf::k_fac_blocks<float>, grid, dim3(256), 0, st, frp, ip + lev_off[d],
                           (const double*)Ed, (float*)ctx->M, (float*)ctx->Mt);
    }
    if (d == 0 && nroots > 0) {
      if (nroot_tasks > 0)
        hipLaunchKernelGGL(vhf::k_fac_tiles, dim3(nroot_tasks), dim3(64), 0, st, frp, tp + root_tasks_off, 0,
                           (double*)Fd, (double*)Ed, (const double*)Dd);
      const dim3 rgrid(nroots, std::max(1, std::min(256, 2048 / nroots)));
      if (ctx->prec == VH_PREC_F64)
        hipLaunchKernelGGL(vhf::k_fac_root<double>, rgrid, dim3(256), 0, st, frp, ip + root_off,
                           (const double*)Fd, (double*)ctx->Sroot);
      else
        hipLaunchKernelGGL(vhf::k_fac_root<float>, rgrid, dim3(256), 0, st, frp, ip + root_off,
                           (const double*)Fd, (float*)ctx->Sroot);
    }
    VH_HIP(hipGetLastError());
    Fc = Fd; Ec = Ed; Dc = Dd;
  }
  int h_fail = 0;
  VH_HIP(hipMemcpyAsync(&h_fail, d_fail, sizeof(int), hipMemcpyDeviceToHost, st));
  VH_HIP(hipStreamSynchronize(st));
  if (h_fail) return fail(ctx, VH_E_NOT_SPD, "Cholesky factorization failed: matrix is not positive definite");
  return VH_OK;
}

}  // namespace

extern "C" {

vh_status vh_create(int32_t device, vh_ctx** out) {
  if (!out) return VH_E_INVALID_ARG;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return VH_E_GPU;
  if (device < 0 || device >= count) return VH_E_INVALID_ARG;
  vh_ctx* ctx = new (std::nothrow) vh_ctx();
  if (!ctx) return VH_E_INTERNAL;
  ctx->device = device;
  if (const char* g = getenv("VISTIO_HIP_GRAPH")) ctx->use_graph = atoi(g) != 0;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return VH_E_GPU;
  }
  if (noalloc_setup(device) != hipSuccess) {
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return VH_E_GPU;
  }
  *out = ctx;
  return VH_OK;
}

vh_status vh_destroy(vh_ctx* ctx) {
  if (!ctx) return VH_E_INVALID_ARG;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  free_all(ctx);
  for (void* p : {(void*)ctx->bk_a, (void*)ctx->bk_b, (void*)ctx->bk_o, (void*)ctx->bk_part, ctx->nb_arena})
    if (p) (void)hipFree(p);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return VH_OK;
}

const char* vh_last_error(const vh_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

vh_status vh_pd_init(vh_ctx* ctx, uint32_t n, const float* px, const float* py, const float* pz, uint32_t t,
                     const uint32_t* tri, uint32_t h, const uint32_t* hinge, const uint8_t* pinned,
                     const vh_pd_params* params, float* lumped_mass_out) {
  return vh_pd_init_ex(ctx, n, px, py, pz, t, tri, h, hinge, pinned, params, nullptr, lumped_mass_out);
}

vh_status vh_pd_init_ex(vh_ctx* ctx, uint32_t n, const float* px, const float* py, const float* pz, uint32_t t,
                        const uint32_t* tri, uint32_t h, const uint32_t* hinge, const uint8_t* pinned,
                        const vh_pd_params* params, const vh_material_arrays* mat, float* lumped_mass_out) {
  if (!ctx) return VH_E_INVALID_ARG;
  if (!params) return fail(ctx, VH_E_INVALID_ARG, "params is null");
  VH_HIP(hipSetDevice(ctx->device));
  VH_HIP(hipStreamSynchronize(ctx->stream));
  free_all(ctx);
  const vh_pd_params& P = *params;
  if (P.model < VH_MODEL_ARAP || P.model > VH_MODEL_ANISOTROPIC_COROTATIONAL)
    return fail(ctx, VH_E_UNSUPPORTED, "unsupported constitutive model");
  if (P.bending != VH_BENDING_NONE && P.bending != VH_BENDING_DIHEDRAL && P.bending != VH_BENDING_DISCRETE_SHELLS)
    return fail(ctx, VH_E_UNSUPPORTED, "unsupported bending model");
  if (!(P.warp_stiffness >= 0) || !(P.weft_stiffness >= 0) || !(P.bend_warp >= 0) || !(P.bend_weft >= 0))
    return fail(ctx, VH_E_INVALID_CONFIG, "warp/weft stiffness must be >= 0");
  if (P.precision != VH_PREC_F64 && P.precision != VH_PREC_F32_IR)
    return fail(ctx, VH_E_UNSUPPORTED, "unsupported precision mode");
  if (P.max_iterations < 0) return fail(ctx, VH_E_INVALID_CONFIG, "max_iterations must be >= 0");
  if (P.precision == VH_PREC_F32_IR &&
      (P.ir_sweeps < 1 || P.ir_sweeps > 3 || P.ir_final_sweeps < 1 || P.ir_final_sweeps > 3))
    return fail(ctx, VH_E_INVALID_CONFIG, "ir_sweeps and ir_final_sweeps must be in 1..3");
  if (n == 0) return fail(ctx, VH_E_INVALID_MESH, "empty mesh");
  ctx->P = P;
  ctx->mem.kind = P.model;
  ctx->mem.kw = ctx->mem.kf = 0.0;
  if (P.model == VH_MODEL_ORTHOTROPIC_LINEAR) {  // orthotropic.rs:64-66
    const double mk = std::max(std::max(P.warp_stiffness, P.weft_stiffness), 1e-6);
    ctx->mem.kw = P.warp_stiffness / mk;
    ctx->mem.kf = P.weft_stiffness / mk;
  } else if (P.model == VH_MODEL_ANISOTROPIC_COROTATIONAL) {  // anisotropic.rs:88-89
    ctx->mem.kw = std::min(std::max(P.warp_stiffness, 0.0), 1.0);
    ctx->mem.kf = std::min(std::max(P.weft_stiffness, 0.0), 1.0);
  }
  ctx->prec = P.precision;
  ctx->tsz = P.precision == VH_PREC_F64 ? 8 : 4;
  const auto t0 = std::chrono::steady_clock::now();
  vh::SetupInput in;
  in.n = (int)n; in.t = (int)t; in.h = P.bending == VH_BENDING_NONE ? 0 : (int)h;
  in.bending_kind = P.bending == VH_BENDING_DISCRETE_SHELLS ? 2 : 1;
  in.bend_warp = P.bend_warp; in.bend_weft = P.bend_weft;
  in.px = px; in.py = py; in.pz = pz; in.tri = tri; in.hinge = hinge; in.pinned = pinned;
  in.k_membrane = P.k_membrane; in.k_bending = P.k_bending; in.density = P.density; in.dt_matrix = P.dt_matrix;
  if (mat) {
    in.k_membrane_tri = mat->k_membrane;
    in.k_bending_hinge = mat->k_bending;
    in.density_tri = mat->density;
  }
  in.leaf_size = P.leaf_size > 0 ? P.leaf_size : 16;
  in.value_bytes = (int)ctx->tsz;
  // Default band depth: 4 levels in FP32-IR, whose staged narrow bands
  // make deeper merging pay (256x256: 5 -> 4 bands, 1070 -> 1075 steps/s;
  // 128x128: 4 -> 3 bands, 2411 -> 2626), 3 in FP64 (4 levels: 727 -> 655
  // at 256x256).
  in.max_band_levels = P.max_band_levels > 0 ? P.max_band_levels : (ctx->prec == VH_PREC_F64 ? 3 : 4);
  if (P.band_launch_us > 0) in.band_launch_us = P.band_launch_us;
  if (P.band_gbps > 0) in.band_gbps = P.band_gbps;
  if (P.task_entries > 0) in.task_entries = P.task_entries;
  ctx->S = vh::Setup();
  if (!vh::build_setup(in, ctx->S)) return fail(ctx, (vh_status)ctx->S.status, ctx->S.err);
  vh::Setup& S = ctx->S;
  ctx->setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (lumped_mass_out)
    for (int i = 0; i < S.n; ++i) lumped_mass_out[i] = (float)S.mass_orig[i];

  vh_status s = VH_OK;
#define UP(expr)              \
  do {                        \
    s = (expr);               \
    if (s != VH_OK) return s; \
  } while (0)
  const size_t n3 = 3 * (size_t)S.n;
  UP(dalloc(ctx, &ctx->x, n3)); UP(dalloc(ctx, &ctx->v, n3));
  UP(dalloc(ctx, &ctx->prev, n3)); UP(dalloc(ctx, &ctx->pred, n3));
  UP(dalloc(ctx, &ctx->r, n3)); UP(dalloc(ctx, &ctx->acc, n3));
  UP(upload(ctx, &ctx->mass, S.mass));
  UP(upload(ctx, &ctx->pinned, S.pinned));
  UP(upload(ctx, &ctx->tri_idx, S.tri_idx));
  UP(upload(ctx, &ctx->lv_ptr, S.lv_ptr)); UP(upload(ctx, &ctx->lv_slot, S.lv_slot));
  UP(upload(ctx, &ctx->lv_end, S.lv_end)); UP(upload(ctx, &ctx->lc_pos, S.lc_pos));
  {
    // k_local's block order (vhk::local_chunk): per XCD, one eighth of the
    // triangle chunks and one eighth of the hinge chunks.
    const int tb = nblk(S.t), hb = nblk(S.h);
    ctx->local_map.tper = (tb + vhk::kLocalXcds - 1) / vhk::kLocalXcds;
    ctx->local_map.hper = (hb + vhk::kLocalXcds - 1) / vhk::kLocalXcds;
    ctx->local_blocks = tb + hb > 0 ? vhk::kLocalXcds * (ctx->local_map.tper + ctx->local_map.hper) : 0;
  }
  UP(upload(ctx, &ctx->tri_dm, S.tri_dm));
  UP(upload_t(ctx, &ctx->tri_w, S.tri_w));
  UP(upload(ctx, &ctx->hinge_idx, S.hinge_idx));
  UP(upload_t(ctx, &ctx->hinge_rest, S.hinge_rest));
  UP(upload_t(ctx, &ctx->hinge_w, S.hinge_w));
  ctx->hinge_st = nullptr;
  if (!S.hinge_st.empty()) UP(upload_t(ctx, &ctx->hinge_st, S.hinge_st));
  UP(dalloc_bytes(ctx, &ctx->cslot, 3 * (size_t)std::max(1, S.inc_ptr[S.n]) * ctx->tsz));
  UP(upload(ctx, &ctx->inc_ptr, S.inc_ptr));
  // Task records hold block and root-inverse offsets as 32-bit integers.
  if (S.m_total > INT32_MAX || S.sroot_total > INT32_MAX || S.fslot_total > INT32_MAX)
    return fail(ctx, VH_E_UNSUPPORTED, "solve blocks exceed 2^31 entries (split the batch)");
  std::vector<vhk::TaskDev> td(S.ptasks.size());
  for (size_t k = 0; k < td.size(); ++k) {
    const vh::Task& T = S.ptasks[k];
    const vh::Supernode& q = S.sn[T.sn];
    td[k].q0 = make_int4(q.col0, q.w, q.r, (int)q.m_off);
    td[k].q1 = make_int4(T.b, T.e, T.g, q.u_off);
    td[k].q2 = make_int4(q.fm_off, q.fslot_k, q.parent < 0 ? (int)q.s_off : 0, 0);
  }
  UP(upload(ctx, &ctx->tasks, td));
  // Dense root tasks (k_root_dense) when the root band is not the only one.
  ctx->root_tasks = nullptr;
  ctx->nroot_tasks = 0; ctx->root_rows = 0; ctx->root_nl = 0; ctx->root_lds = 0;
  if (S.max_depth > 0) {
    int wmax = 0;
    for (const vh::Supernode& q : S.sn)
      if (q.parent < 0) wmax = std::max(wmax, q.w);
    ctx->root_rows = root_threads(wmax, ctx->prec == VH_PREC_F64) / 64;
    if (ctx->root_rows == 0) wmax = 0;
    ctx->root_nl = root_loads(wmax, ctx->prec == VH_PREC_F64);
    ctx->root_lds = 3 * (size_t)wmax * ctx->tsz;
    std::vector<vhk::TaskDev> rt;
    for (const vh::Supernode& q : S.sn) {
      if (q.parent >= 0 || ctx->root_rows == 0) continue;
      for (int b = 0; b < q.w; b += ctx->root_rows) {
        vhk::TaskDev x{};
        x.q0 = make_int4(q.col0, q.w, 0, 0);
        x.q1 = make_int4(b, std::min(q.w, b + ctx->root_rows), 64, 0);
        x.q2 = make_int4(q.fm_off, q.fslot_k, (int)q.s_off, 0);
        rt.push_back(x);
      }
    }
    ctx->nroot_tasks = (int)rt.size();
    if (!rt.empty()) UP(upload(ctx, &ctx->root_tasks, rt));
  }
  // Staged forward tasks: each supernode's rows in runs of at most 256 rows
  // and 256 st_E entries; a band with a row longer than that keeps
  // k_fwd_band.
  ctx->st_tasks = nullptr;
  ctx->st_E = kStagedLoads;
  ctx->st_off.assign(S.max_depth + 2, 0);
  if (ctx->st_E > 0 && S.max_depth > 1) {
    const int cap = 256 * ctx->st_E;
    std::vector<vhk::TaskDev> sv;
    for (int l = 0; l <= S.max_depth; ++l) {
      ctx->st_off[l] = (int)sv.size();
      const int depth = S.max_depth - l;
      if (l == 0 || depth == 0) continue;
      const size_t lvl0 = sv.size();
      bool fits = true;
      const int wlim = std::min(cap, kStagedMaxW);
      for (const vh::Supernode& q : S.sn) {
        if (q.depth != depth) continue;
        if (q.w > wlim) { fits = false; break; }
        const int m = q.w + q.r;
        int g = 4;
        while (g < 64 && g * 8 < q.w) g *= 2;
        for (int b = 0; b < m;) {
          int e = b;
          int64_t ent = 0;
          while (e < m && e - b < 256) {
            const int len = e < q.w ? e + 1 : q.w;
            if (e > b && ent + len > cap) break;
            ent += len;
            ++e;
          }
          vhk::TaskDev x{};
          x.q0 = make_int4(q.col0, q.w, q.r, (int)q.m_off);
          x.q1 = make_int4(b, e, g, q.u_off);
          x.q2 = make_int4(q.fm_off, q.fslot_k, 0, 0);
          sv.push_back(x);
          b = e;
        }
      }
      if (!fits) sv.resize(lvl0);
    }
    ctx->st_off[S.max_depth + 1] = (int)sv.size();
    if (!sv.empty()) UP(upload(ctx, &ctx->st_tasks, sv));
  }
  // Staged first forward launch: the deepest band's supernodes (no
  // children) cut into tasks of at most 255 columns, 256 rows, 256 st_E
  // entries and 8192 RHS slot values; the slot run of the task's columns
  // goes into the record.
  ctx->st1_tasks = nullptr;
  ctx->nst1 = 0; ctx->st1_NS = 0;
  if (ctx->st_E > 0 && S.max_depth > 0) {
    const int cap = 256 * ctx->st_E, scap = 256 * 32;
    std::vector<vhk::TaskDev> fv;
    bool fits = true;
    int maxsv = 0;
    for (const vh::Supernode& q : S.sn) {
      if (q.depth != S.max_depth || !fits) continue;
      if (q.fslot_k != 0 || q.w > std::min(cap, 255) ||
          3 * (int64_t)(S.inc_ptr[q.col0 + q.w] - S.inc_ptr[q.col0]) > scap) { fits = false; break; }
      const int m = q.w + q.r;
      int g = 4;
      while (g < 64 && g * 8 < q.w) g *= 2;
      for (int b = 0; b < m;) {
        int e = b;
        int64_t ent = 0;
        while (e < m && e - b < 256) {
          const int len = e < q.w ? e + 1 : q.w;
          if (e > b && ent + len > cap) break;
          ent += len;
          ++e;
        }
        const int tl = std::min(e, q.w);
        vhk::TaskDev x{};
        x.q0 = make_int4(q.col0, q.w, q.r, (int)q.m_off);
        x.q1 = make_int4(b, e, g, q.u_off);
        x.q2 = make_int4(q.fm_off, 0, S.inc_ptr[q.col0], S.inc_ptr[q.col0 + tl]);
        maxsv = std::max(maxsv, 3 * (S.inc_ptr[q.col0 + tl] - S.inc_ptr[q.col0]));
        fv.push_back(x);
        b = e;
      }
    }
    if (fits && !fv.empty()) {
      ctx->st1_NS = maxsv <= 256 * 8 ? 8 : (maxsv <= 256 * 16 ? 16 : 32);
      ctx->nst1 = (int)fv.size();
      UP(upload(ctx, &ctx->st1_tasks, fv));
      const size_t l1 = staged1_lds(ctx, S.band_front[S.max_depth]);
      if (l1 > 160 * 1024) return fail(ctx, VH_E_UNSUPPORTED, "staged solve task exceeds the 160 KB LDS");
      if (l1 > 64 * 1024) {
        const void* k = ctx->prec == VH_PREC_F64 ? fwd_st1_kernel<double>(ctx->st1_NS) : fwd_st1_kernel<float>(ctx->st1_NS);
        if (k) VH_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)l1));
      }
    }
  }
  // Staged committing band (FP32): the deepest band's columns (rows of Mt)
  // in tasks of at most 256 rows and 256 st_E entries. Measured at 256x256:
  // FP32-IR 1108 -> 1124 steps/s (4 bands); FP64 753 -> 748, so FP64 keeps
  // k_bwd_band<double, true>.
  ctx->stc_tasks = nullptr;
  ctx->nstc = 0;
  if (ctx->st_E > 0 && S.max_depth > 0 && ctx->prec != VH_PREC_F64) {
    const int cap = 256 * ctx->st_E;
    std::vector<vhk::TaskDev> cv;
    bool fits = true;
    for (const vh::Supernode& q : S.sn) {
      if (q.depth != S.max_depth || q.parent < 0) continue;
      const int m = q.w + q.r;
      if (m > cap) { fits = false; break; }
      int g = 4;
      while (g < 64 && g * 8 < m) g *= 2;
      for (int b = 0; b < q.w;) {
        int e = b;
        int64_t ent = 0;
        while (e < q.w && e - b < 256) {
          const int len = m - e;
          if (e > b && ent + len > cap) break;
          ent += len;
          ++e;
        }
        vhk::TaskDev x{};
        x.q0 = make_int4(q.col0, q.w, q.r, (int)q.m_off);
        x.q1 = make_int4(b, e, g, q.u_off);
        x.q2 = make_int4(q.fm_off, q.fslot_k, 0, 0);
        cv.push_back(x);
        b = e;
      }
    }
    if (fits && !cv.empty()) {
      ctx->nstc = (int)cv.size();
      UP(upload(ctx, &ctx->stc_tasks, cv));
      const size_t lc = stagedc_lds(ctx, S.band_front[S.max_depth]);
      if (lc > 160 * 1024) return fail(ctx, VH_E_UNSUPPORTED, "staged solve task exceeds the 160 KB LDS");
      if (lc > 64 * 1024)
        VH_HIP(hipFuncSetAttribute((const void*)vhk::k_bwd_band_stc<kStagedLoads, float>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lc));
    }
  }
  {
    size_t lds = 0;
    for (int f : S.band_front) lds = std::max(lds, staged_lds(ctx, f));
    if (lds > 160 * 1024) return fail(ctx, VH_E_UNSUPPORTED, "staged solve task exceeds the 160 KB LDS");
    if (lds > 64 * 1024) {
      const void* k = ctx->prec == VH_PREC_F64 ? fwd_st_kernel<double>(ctx->st_E) : fwd_st_kernel<float>(ctx->st_E);
      if (k) VH_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    }
  }
  UP(upload(ctx, &ctx->uslot, S.uslot));
  UP(upload(ctx, &ctx->dpush_ptr, S.dpush_ptr));
  UP(upload(ctx, &ctx->dpush_idx, S.dpush_idx));
  UP(upload(ctx, &ctx->upper_cols, S.upper_cols));
  UP(dalloc_bytes(ctx, &ctx->M, (size_t)S.m_total * ctx->tsz));
  UP(dalloc_bytes(ctx, &ctx->Mt, (size_t)S.m_total * ctx->tsz));
  UP(dalloc_bytes(ctx, &ctx->Sroot, (size_t)S.sroot_total * ctx->tsz));
  UP(dalloc_bytes(ctx, &ctx->y, n3 * ctx->tsz));
  UP(dalloc_bytes(ctx, &ctx->d, n3 * ctx->tsz));
  UP(dalloc_bytes(ctx, &ctx->F, 3 * (size_t)std::max<int64_t>(1, S.fslot_total) * ctx->tsz));
  UP(dalloc_bytes(ctx, &ctx->V, 3 * (size_t)std::max(1, S.u_total) * ctx->tsz));
  // A in FP64: the device factorization assembles its fronts from it, and
  // the FP32-IR step forms its residual with it.
  UP(upload(ctx, &ctx->Ap, S.Ap));
  UP(upload(ctx, &ctx->Ai, S.Ai));
  UP(upload(ctx, &ctx->Ax, S.Ax));
  {
    const auto tf = std::chrono::steady_clock::now();
    UP(device_factor(ctx));
    ctx->factor_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - tf).count();
  }
  // Residual partials come from the committing launch: the last backward
  // band, or the root band when it is the only band.
  const int nb = S.max_depth + 1;
  ctx->resid_blocks = nb == 1 ? band_task_blocks(S.fwd_level_off[1] - S.fwd_level_off[0], false)
                              : band_task_blocks(ctx->nstc > 0 ? ctx->nstc : S.bwd_level_off[nb] - S.bwd_level_off[nb - 1],
                                                 false) +
                                    nblk((int)S.upper_cols.size());
  UP(dalloc(ctx, &ctx->partials, 2 * (size_t)std::max(1, ctx->resid_blocks)));
  UP(dalloc(ctx, &ctx->st, 1));
#undef UP
  VH_HIP(hipMemset(ctx->st, 0, sizeof(vhk::StepDev)));
  VH_HIP(hipMemset(ctx->F, 0, 3 * (size_t)std::max<int64_t>(1, S.fslot_total) * ctx->tsz));
  VH_HIP(hipMemset(ctx->V, 0, 3 * (size_t)std::max(1, S.u_total) * ctx->tsz));
  size_t lds = 0;
  for (int f : S.band_front) lds = std::max(lds, 3 * (size_t)f * ctx->tsz);
  if (lds > 160 * 1024) return fail(ctx, VH_E_UNSUPPORTED, "supernode front exceeds the 160 KB LDS of one workgroup");
  if (lds > 64 * 1024) {
    const bool f64 = ctx->prec == VH_PREC_F64;
    const void* ks[] = {
        f64 ? (const void*)vhk::k_fwd_band<double, false> : (const void*)vhk::k_fwd_band<float, false>,
        f64 ? (const void*)vhk::k_fwd_band<double, true> : (const void*)vhk::k_fwd_band<float, true>,
        f64 ? (const void*)vhk::k_bwd_band<double, false> : (const void*)vhk::k_bwd_band<float, false>,
        f64 ? (const void*)vhk::k_bwd_band<double, true> : (const void*)vhk::k_bwd_band<float, true>,
        f64 ? (const void*)vhk::k_root_band<double, false, false> : (const void*)vhk::k_root_band<float, false, false>,
        f64 ? (const void*)vhk::k_root_band<double, false, true> : (const void*)vhk::k_root_band<float, false, true>,
        f64 ? (const void*)vhk::k_root_band<double, true, false> : (const void*)vhk::k_root_band<float, true, false>,
        f64 ? (const void*)vhk::k_root_band<double, true, true> : (const void*)vhk::k_root_band<float, true, true>,
        f64 ? (const void*)vhk::k_fwd_band_d5 : (const void*)vhk::k_fwd_band_w8,
        f64 ? (const void*)vhk::k_bwd_band_d5 : (const void*)vhk::k_bwd_band<float, false>};
    for (const void* k : ks) VH_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  }
  if (ctx->root_lds > 64 * 1024) {
    VH_HIP(hipFuncSetAttribute(root_dense_kernel(ctx->root_nl), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)ctx->root_lds));
  }
  // Initial state (state.rs:119-139): pos = prev = mesh positions, vel = 0.
  std::vector<double> h3(n3);
  for (int k = 0; k < S.n; ++k) {
    const int o = S.perm[k];
    h3[3 * k] = px[o]; h3[3 * k + 1] = py[o]; h3[3 * k + 2] = pz[o];
  }
  VH_HIP(hipMemcpy(ctx->x, h3.data(), n3 * 8, hipMemcpyHostToDevice));
  VH_HIP(hipMemcpy(ctx->prev, h3.data(), n3 * 8, hipMemcpyHostToDevice));
  VH_HIP(hipMemset(ctx->v, 0, n3 * 8));
  VH_HIP(hipMemset(ctx->pred, 0, n3 * 8));
  ctx->initialized = true;
  return VH_OK;
}

vh_status vh_state_set_f64(vh_ctx* ctx, int32_t which, const double* x, const double* y, const double* z) {
  if (!ctx) return VH_E_INVALID_ARG;
  if (!ctx->initialized) return fail(ctx, VH_E_NOT_INITIALIZED, "Solver not initialized. Call init() first.");
  double* dst = state_ptr(ctx, which);
  if (!dst || !x || !y || !z) return fail(ctx, VH_E_INVALID_ARG, "bad state selector or null array");
  VH_HIP(hipSetDevice(ctx->device));
  to_dev_order(ctx->S, x, y, z, ctx->hbuf);
  VH_HIP(hipMemcpyAsync(dst, ctx->hbuf.data(), ctx->hbuf.size() * 8, hipMemcpyHostToDevice, ctx->stream));
  VH_HIP(hipStreamSynchronize(ctx->stream));
  return VH_OK;
}

vh_status vh_state_get_f64(vh_ctx* ctx, int32_t which, double* x, double* y, double* z) {
  if (!ctx) return VH_E_INVALID_ARG;
  if (!ctx->initialized) return fail(ctx, VH_E_NOT_INITIALIZED, "Solver not initialized. Call init() first.");
  double* src = state_ptr(ctx, which);
  if (!src || !x || !y || !z) return fail(ctx, VH_E_INVALID_ARG, "bad state selector or null array");
  VH_HIP(hipSetDevice(ctx->device));
  ctx->hbuf.resize(3 * (size_t)ctx->S.n);
  VH_HIP(hipMemcpyAsync(ctx->hbuf.data(), src, ctx->hbuf.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
  VH_HIP(hipStreamSynchronize(ctx->stream));
  from_dev_order(ctx->S, ctx->hbuf, x, y, z);
  return VH_OK;
}

vh_status vh_state_set_inv_mass(vh_ctx* ctx, const float* inv_mass) {
  if (!ctx) return VH_E_INVALID_ARG;
  if (!ctx->initialized) return fail(ctx, VH_E_NOT_INITIALIZED, "Solver not initialized. Call init() first.");
  if (!inv_mass) return fail(ctx, VH_E_INVALID_ARG, "inv_mass is null");
  VH_HIP(hipSetDevice(ctx->device));
  std::vector<uint8_t> pin(ctx->S.n);
  for (int k = 0; k < ctx->S.n; ++k) pin[k] = inv_mass[ctx->S.perm[k]] == 0.0f ? 1 : 0;
  VH_HIP(hipMemcpyAsync(ctx->pinned, pin.data(), pin.size(), hipMemcpyHostToDevice, ctx->stream));
  VH_HIP(hipStreamSynchronize(ctx->stream));
  return VH_OK;
}

vh_status vh_pd_set_ground(vh_ctx* ctx, int32_t has_ground, double ground_height) {
  if (!ctx) return VH_E_INVALID_ARG;
  if (!ctx->initialized) return fail(ctx, VH_E_NOT_INITIALIZED, "Solver not initialized. Call init() first.");
  if (has_ground && !std::isfinite(ground_height)) return fail(ctx, VH_E_INVALID_ARG, "ground height must be finite");
  const int32_t on = has_ground ? 1 : 0;
  const double h = on ? ground_height : 0.0;
  if (on != ctx->P.has_ground || h != ctx->P.ground_height) {
    ctx->P.has_ground = on;
    ctx->P.ground_height = h;
    drop_graph(ctx);  // the captured sequences hold the old value in their kernel arguments
    drop_al_graph(ctx);
  }
  return VH_OK;
}

vh_status vh_pd_step_resident(vh_ctx* ctx, double dt, uint32_t nsteps, vh_step_result* out) {
  if (!ctx) return VH_E_INVALID_ARG;
  if (!ctx->initialized) return fail(ctx, VH_E_NOT_INITIALIZED, "Solver not initialized. Call init() first.");
  if (!(dt > 0)) return fail(ctx, VH_E_INVALID_CONFIG, "dt must be > 0");
  VH_HIP(hipSetDevice(ctx->device));
  const auto t0 = std::chrono::steady_clock::now();
  for (uint32_t s = 0; s < nsteps; ++s) {
    const vh_status ls = launch_step(ctx, dt);
    if (ls != VH_OK) return ls;
  }
  vh_status st = read_step_result(ctx, out);
  if (out) out->wall_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return st;
}

// ---- host f32 state in and out of the drop-in steps ------------------------
// The caller's SoA arrays are packed into pinned memory and copied in one
// transfer; a kernel permutes them into the device order and widens them to
// FP64. On the way out a kernel narrows (the same round-to-nearest as a host
// cast) and permutes into caller order, and one transfer brings them back.
// Slots of h_io / d_io, n floats each: in 0-2 pos, 3-5 vel, 6 inv_mass,
// 7-9 lambda; out 0-2 pos, 3-5 vel, 6-8 prev, 9-11 pred, 12-14 lambda.
constexpr int kIoSlots = 16;

__global__ void __launch_bounds__(256) k_io_in(int n, const int* __restrict__ perm, const float* __restrict__ in,
                                               int has_inv, int has_lambda, double* __restrict__ x,
                                               double* __restrict__ v, uint8_t* __restrict__ pinned,
                                               double* __restrict__ lambda) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const size_t o = (size_t)perm[k], nn = (size_t)n;
  for (int c = 0; c < 3; ++c) {
    x[3 * (size_t)k + c] = (double)in[c * nn + o];
    v[3 * (size_t)k + c] = (double)in[(3 + c) * nn + o];
    if (has_lambda) lambda[3 * (size_t)k + c] = (double)in[(7 + c) * nn + o];
  }
  if (has_inv) pinned[k] = in[6 * nn + o] == 0.0f ? 1 : 0;
}

__global__ void __launch_bounds__(256) k_io_out(int n, const int* __restrict__ perm, const double* __restrict__ x,
                                                const double* __restrict__ v, const double* __restrict__ prev,
                                                const double* __restrict__ pred, const double* __restrict__ lambda,
                                                float* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const size_t o = (size_t)perm[k], nn = (size_t)n;
  for (int c = 0; c < 3; ++c) {
    out[c * nn + o] = (float)x[3 * (size_t)k + c];
    out[(3 + c) * nn + o] = (float)v[3 * (size_t)k + c];
    out[(6 + c) * nn + o] = (float)prev[3 * (size_t)k + c];
    out[(9 + c) * nn + o] = (float)pred[3 * (size_t)k + c];
    if (lambda) out[(12 + c) * nn + o] = (float)lambda[3 * (size_t)k + c];
  }
}

vh_status io_reserve(vh_ctx* ctx) {
  if (ctx->h_io) return VH_OK;
  const size_t n = (size_t)ctx->S.n;
  VH_HIP(hipHostMalloc((void**)&ctx->h_io, kIoSlots * std::max<size_t>(n, 1) * sizeof(float), hipHostMallocDefault));
  vh_status st = dalloc(ctx, &ctx->d_io, kIoSlots * std::max<size_t>(n, 1));
  if (st == VH_OK) st = upload(ctx, &ctx->d_perm, ctx->S.perm);
  if (st != VH_OK) {
    (void)hipHostFree(ctx->h_io);
    ctx->h_io = nullptr;
  }
  return st;
}

// Packs and uploads pos, vel, inv_mass (optional) and lambda (optional), then
// scatters them into the device state.
vh_status io_upload(vh_ctx* ctx, const vh_state_f32* s, const float* lx, const float* ly, const float* lz) {
  const int n = ctx->S.n;
  const size_t nb = (size_t)n * sizeof(float);
  float* h = ctx->h_io;
  const float* src[10] = {s->pos_x, s->pos_y, s->pos_z, s->vel_x, s->vel_y, s->vel_z, s->inv_mass, lx, ly, lz};
  int last = 0;
  for (int q = 0; q < 10; ++q)
    if (src[q]) { std::memcpy(h + (size_t)q * n, src[q], nb); last = q; }
  VH_HIP(hipMemcpyAsync(ctx->d_io, h, (size_t)(last + 1) * nb, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_io_in, dim3(std::max(1, nblk(n))), dim3(256), 0, ctx->stream, n, ctx->d_perm, ctx->d_io,
                     s->inv_mass ? 1 : 0, lx ? 1 : 0, ctx->x, ctx->v, ctx->pinned, ctx->ipc.lambda);
  VH_HIP(hipGetLastError());
  return VH_OK;
}

// Gathers pos, vel, prev, pred (and lambda) into caller order, downloads them
// and synchronizes; io_scatter then copies them into the caller's arrays.
vh_status io_download(vh_ctx* ctx, bool with_lambda) {
  const int n = ctx->S.n;
  const size_t nb = (size_t)n * sizeof(float);
  hipLaunchKernelGGL(k_io_out, dim3(std::max(1, nblk(n))), dim3(256), 0, ctx->stream, n, ctx->d_perm, ctx->x, ctx->v,
                     ctx->prev, ctx->pred, with_lambda ? ctx->ipc.lambda : nullptr, ctx->d_io);
  VH_HIP(hipGetLastError());
  VH_HIP(hipMemcpyAsync(ctx->h_io, ctx->d_io, (size_t)(with_lambda ? 15 : 12) * nb, hipMemcpyDeviceToHost,
                        ctx->stream));
  return VH_OK;
}

void io_scatter(vh_ctx* ctx, vh_state_f32* s, float* lx, float* ly, float* lz) {
  const size_t n = (size_t)ctx->S.n, nb = n * sizeof(float);
  float* dst[15] = {s->pos_x, s->pos_y, s->pos_z, s->vel_x, s->vel_y, s->vel_z, s->prev_x, s->prev_y, s->prev_z,
                    s->pred_x, s->pred_y, s->pred_z, lx, ly, lz};
  for (int q = 0; q < 15; ++q)
    if (dst[q]) std::memcpy(dst[q], ctx->h_io + q * n, nb);
}

vh_status vh_pd_step(vh_ctx* ctx, double dt, vh_state_f32* s, vh_step_result* out) {
  if (!ctx) return VH_E_INVALID_ARG;
  if (!ctx->initialized) return fail(ctx, VH_E_NOT_INITIALIZED, "Solver not initialized. Call init() first.");
  if (!(dt > 0)) return fail(ctx, VH_E_INVALID_CONFIG, "dt must be > 0");
  if (!s || !s->pos_x || !s->pos_y || !s->pos_z || !s->vel_x || !s->vel_y || !s->vel_z || !s->prev_x ||
      !s->prev_y || !s->prev_z || !s->pred_x || !s->pred_y || !s->pred_z)
    return fail(ctx, VH_E_INVALID_ARG, "null state array");
  const auto t0 = std::chrono::steady_clock::now();
  VH_HIP(hipSetDevice(ctx->device));
  {
    vh_status e = io_reserve(ctx);
    if (e == VH_OK) e = io_upload(ctx, s, nullptr, nullptr, nullptr);
    if (e != VH_OK) return e;
  }
  {
    const vh_status ls = launch_step(ctx, dt);
    if (ls != VH_OK) return ls;
  }
  {
    const vh_status e = io_download(ctx, false);
    if (e != VH_OK) return e;
  }
  vh_status st = read_step_result(ctx, out);  // synchronizes
  io_scatter(ctx, s, nullptr, nullptr, nullptr);
  if (out) out->wall_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return st;
}

vh_status vh_pd_residual(vh_ctx* ctx, double dt, double* out) {
  if (!ctx) return VH_E_INVALID_ARG;
  if (!ctx->initialized) return fail(ctx, VH_E_NOT_INITIALIZED, "Solver not initialized. Call init() first.");
  if (!out) return fail(ctx, VH_E_INVALID_ARG, "out is null");
  if (!(dt > 0)) return fail(ctx, VH_E_INVALID_CONFIG, "dt must be > 0");
  VH_HIP(hipSetDevice(ctx->device));
  const int n = ctx->S.n;
  if (enqueue_local_step(ctx, nullptr)) {
    if (ctx->prec == VH_PREC_F64)
      hipLaunchKernelGGL(vhk::k_gather<double>, dim3(nblk(n)), dim3(256), 0, ctx->stream, n,
                         rhs_src<double>(ctx, dt), ctx->r);
    else
      hipLaunchKernelGGL(vhk::k_gather<float>, dim3(nblk(n)), dim3(256), 0, ctx->stream, n, rhs_src<float>(ctx, dt),
                         ctx->r);
  }
  VH_HIP(hipGetLastError());
  const size_t n3 = 3 * (size_t)n;
  ctx->hbuf.resize(n3);
  VH_HIP(hipMemcpyAsync(ctx->hbuf.data(), ctx->r, n3 * 8, hipMemcpyDeviceToHost, ctx->stream));
  VH_HIP(hipStreamSynchronize(ctx->stream));
  from_dev_order(ctx->S, ctx->hbuf, out, out + n, out + 2 * (size_t)n);
  return VH_OK;
}

vh_status vh_pd_solve(vh_ctx* ctx, const double* b, double* x, int32_t sweeps) {
  if (!ctx) return VH_E_INVALID_ARG;
  if (!ctx->initialized) return fail(ctx, VH_E_NOT_INITIALIZED, "Solver not initialized. Call init() first.");
  if (!b || !x) return fail(ctx, VH_E_INVALID_ARG, "null array");
  if (ctx->prec == VH_PREC_F32_IR && (sweeps < 1 || sweeps > 8)) return fail(ctx, VH_E_INVALID_ARG, "sweeps must be in 1..8");
  VH_HIP(hipSetDevice(ctx->device));
  const int n = ctx->S.n;
  const size_t n3 = 3 * (size_t)n;
  to_dev_order(ctx->S, b, b + n, b + 2 * (size_t)n, ctx->hbuf);
  VH_HIP(hipMemcpyAsync(ctx->r, ctx->hbuf.data(), n3 * 8, hipMemcpyHostToDevice, ctx->stream));
  Marks none;
  enqueue_correction(ctx, ctx->P.dt_matrix, sweeps, false, false, false, nullptr, none);
  VH_HIP(hipGetLastError());
  if (ctx->prec == VH_PREC_F64) {
    VH_HIP(hipMemcpyAsync(ctx->hbuf.data(), ctx->d, n3 * 8, hipMemcpyDeviceToHost, ctx->stream));
  } else {
    VH_HIP(hipMemcpyAsync(ctx->hbuf.data(), ctx->acc, n3 * 8, hipMemcpyDeviceToHost, ctx->stream));
  }
  VH_HIP(hipStreamSynchronize(ctx->stream));
  from_dev_order(ctx->S, ctx->hbuf, x, x + n, x + 2 * (size_t)n);
  return VH_OK;
}

vh_status vh_pd_info_get(vh_ctx* ctx, vh_pd_info* out) {
  if (!ctx || !out) return VH_E_INVALID_ARG;
  if (!ctx->initialized) return fail(ctx, VH_E_NOT_INITIALIZED, "Solver not initialized. Call init() first.");
  const vh::Setup& S = ctx->S;
  out->n = (uint32_t)S.n; out->t = (uint32_t)S.t; out->h = (uint32_t)S.h;
  out->supernodes = (uint32_t)S.sn.size();
  out->levels = (uint32_t)(S.max_depth + 1);
  out->nnz_a = (int64_t)S.Ai.size();
  out->nnz_factor = S.nnz_factor;
  out->max_front = S.max_front;
  out->setup_seconds = ctx->setup_seconds;
  out->factor_seconds = ctx->factor_seconds;
  return VH_OK;
}

vh_status vh_pd_blocks_get(vh_ctx* ctx, int32_t which, double* out, uint64_t cap, uint64_t* count) {
  if (!ctx || !count) return VH_E_INVALID_ARG;
  if (!ctx->initialized) return fail(ctx, VH_E_NOT_INITIALIZED, "Solver not initialized. Call init() first.");
  if (which < 0 || which > 2) return fail(ctx, VH_E_INVALID_ARG, "which must be 0 (M), 1 (M^T) or 2 (root inverses)");
  const vh::Setup& S = ctx->S;
  const uint64_t total = (uint64_t)(which == 2 ? S.sroot_total : S.m_total);
  *count = total;
  const uint64_t k = std::min(cap, total);
  if (k == 0) return VH_OK;
  if (!out) return fail(ctx, VH_E_INVALID_ARG, "null output");
  VH_HIP(hipSetDevice(ctx->device));
  const void* src = which == 0 ? ctx->M : which == 1 ? ctx->Mt : ctx->Sroot;
  VH_HIP(hipStreamSynchronize(ctx->stream));
  if (ctx->tsz == 8) {
    VH_HIP(hipMemcpy(out, src, k * 8, hipMemcpyDeviceToHost));
  } else {
    std::vector<float> f(k);
    VH_HIP(hipMemcpy(f.data(), src, k * 4, hipMemcpyDeviceToHost));
    for (uint64_t q = 0; q < k; ++q) out[q] = (double)f[q];
  }
  return VH_OK;
}

vh_status vh_pd_profile(vh_ctx* ctx, double dt, uint32_t iterations, vh_profile* out) {
  if (!ctx || !out) return VH_E_INVALID_ARG;
  if (!ctx->initialized) return fail(ctx, VH_E_NOT_INITIALIZED, "Solver not initialized. Call init() first.");
  if (!(dt > 0) || iterations == 0) return fail(ctx, VH_E_INVALID_ARG, "dt > 0 and iterations > 0 required");
  const vh::Setup& S = ctx->S;
  const int nb = S.max_depth + 1;
  if (nb > VH_MAX_BANDS) return fail(ctx, VH_E_UNSUPPORTED, "more bands than vh_profile holds");
  VH_HIP(hipSetDevice(ctx->device));
  std::memset(out, 0, sizeof(*out));
  const bool f64 = ctx->prec == VH_PREC_F64;
  const int sweeps = f64 ? 1 : std::max(1, ctx->P.ir_sweeps);
  const int solve_launches = 2 * nb - 1;
  const int per_iter = 1 + sweeps * solve_launches + (sweeps - 1);
  out->bands = (uint32_t)nb;
  out->iterations = iterations;
  out->sweeps = (uint32_t)sweeps;
  out->launches = (uint32_t)per_iter;
  const double ts = (double)ctx->tsz;
  // Algorithmic bytes per launch.
  const double n = S.n, t = S.t, h = S.h;
  // k_local: positions once, element data, the slot plan (6 bytes per slot
  // and 2 per corner), one slot written per (block, vertex)
  const double lslots = (double)S.inc_ptr[S.n];
  out->local_bytes = n * 24 + t * (12 + 32 + ts + 6) + h * (16 + 2 * ts + 8) + lslots * (6 + 3 * ts);
  std::vector<double> band_entries(nb, 0), band_rows(nb, 0), band_cols(nb, 0);
  for (const vh::Supernode& q : S.sn) {
    band_entries[q.depth] += q.parent < 0 ? (double)q.w * q.w : (double)q.w * (q.w + 1) / 2 + (double)q.r * q.w;
    band_rows[q.depth] += q.w + q.r;
    band_cols[q.depth] += q.w;
  }
  const double slots_in = (double)S.inc_ptr[S.n] * 3 * ts + n * 56;  // RHS slots, mass, pred, x -> r
  std::vector<double> fb(nb, 0), bb(nb, 0);
  for (int l = 0; l + 1 < nb; ++l) {
    const int fd = S.max_depth - l;  // forward launch l runs band fd
    // block entries + front inputs (r or RHS slots, child slots) + update rows out
    fb[l] = band_entries[fd] * ts + band_rows[fd] * 3 * ts + band_cols[fd] * 24 + (l == 0 ? slots_in : 0.0) +
            band_rows[fd] * 3 * ts;
  }
  // root: dense inverse + front inputs + d and its pushes (+ commit when alone)
  fb[nb - 1] = band_entries[0] * ts + band_rows[0] * 3 * ts + band_cols[0] * 24 + band_cols[0] * 3 * ts * 2 +
               (nb == 1 ? slots_in + n * 48 : 0.0);
  for (int l = 1; l < nb; ++l)
    bb[l] = band_entries[l] * ts + band_rows[l] * 3 * ts + band_cols[l] * 3 * ts * 2 + (l == nb - 1 ? n * 48 : 0.0);
  const double ir_bytes = f64 ? 0.0 : (double)S.Ai.size() * 12 + n * (12 + 48 + 48);
  std::vector<hipEvent_t> ev((size_t)iterations * per_iter + 1);
  for (auto& e : ev) VH_HIP(hipEventCreate(&e));
  Marks mk;
  mk.ev = ev.data();
  mk.mark(ctx->stream);
  for (uint32_t it = 0; it < iterations; ++it) {
    const bool from_slots = enqueue_local_step(ctx, nullptr);
    mk.mark(ctx->stream);
    enqueue_correction(ctx, dt, sweeps, from_slots, true, false, nullptr, mk);
  }
  VH_HIP(hipGetLastError());
  VH_HIP(hipStreamSynchronize(ctx->stream));
  if (mk.k != ev.size()) {
    for (auto& e : ev) (void)hipEventDestroy(e);
    return fail(ctx, VH_E_INTERNAL, "profile: launch count mismatch");
  }
  auto span = [&](size_t a, size_t b) -> double {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ev[a], ev[b]);
    return 1000.0 * ms;
  };
  for (uint32_t it = 0; it < iterations; ++it) {
    const size_t base = (size_t)it * per_iter;
    size_t q = base;
    out->local_us += span(q, q + 1); ++q;
    for (int s = 0; s < sweeps; ++s) {
      for (int l = 0; l < nb; ++l) { out->fwd_us[l] += span(q, q + 1); ++q; }
      for (int l = 1; l < nb; ++l) { out->bwd_us[l] += span(q, q + 1); ++q; }
      if (s + 1 < sweeps) { out->ir_us += span(q, q + 1); ++q; }
    }
    out->iteration_us += span(base, base + per_iter);
  }
  const double inv = 1.0 / iterations;
  out->local_us *= inv; out->ir_us *= inv; out->iteration_us *= inv;
  for (int l = 0; l < nb; ++l) {
    out->fwd_us[l] *= inv; out->bwd_us[l] *= inv;
    out->fwd_bytes[l] = fb[l] * sweeps; out->bwd_bytes[l] = bb[l] * sweeps;
  }
  out->ir_bytes = ir_bytes * (sweeps - 1);
  for (auto& e : ev) (void)hipEventDestroy(e);
  return VH_OK;
}

vh_status vh_pd_profile_graph(vh_ctx* ctx, double dt, uint32_t reps, vh_graph_profile* out) {
  constexpr int kLocalPerGraph = 10;
  if (!ctx || !out) return VH_E_INVALID_ARG;
  if (!ctx->initialized) return fail(ctx, VH_E_NOT_INITIALIZED, "Solver not initialized. Call init() first.");
  if (!(dt > 0) || reps == 0) return fail(ctx, VH_E_INVALID_ARG, "dt > 0 and reps > 0 required");
  VH_HIP(hipSetDevice(ctx->device));
  // algorithmic bytes of one sweep and of the local step, as vh_pd_profile counts them
  vh_profile p{};
  {
    const vh::Setup& S = ctx->S;
    if (S.max_depth + 1 > VH_MAX_BANDS) return fail(ctx, VH_E_UNSUPPORTED, "more bands than vh_profile holds");
  }
  const int saved = ctx->P.ir_sweeps;
  ctx->P.ir_sweeps = 1;
  const vh_status ps = vh_pd_profile(ctx, dt, 1, &p);
  ctx->P.ir_sweeps = saved;
  if (ps != VH_OK) return ps;
  std::memset(out, 0, sizeof(*out));
  out->reps = reps;
  out->solve_launches = 2 * p.bands - 1;
  for (uint32_t l = 0; l < p.bands; ++l) out->solve_bytes += p.fwd_bytes[l] + p.bwd_bytes[l];
  out->local_bytes = p.local_bytes;
  hipGraphExec_t ex[2] = {nullptr, nullptr};
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto cleanup = [&] {
    for (auto& x : ex)
      if (x) (void)hipGraphExecDestroy(x);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  };
  for (int which = 0; which < 2; ++which) {
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      if (which == 0) {
        // several local steps in one graph: gaps between launches as in a step
        for (int k = 0; k < kLocalPerGraph; ++k) (void)enqueue_local_step(ctx, nullptr);
      } else {
        Marks none;
        if (ctx->prec == VH_PREC_F64) enqueue_solve<double>(ctx, dt, true, true, false, false, nullptr, none);
        else enqueue_solve<float>(ctx, dt, true, true, false, false, nullptr, none);
      }
      const hipError_t le = hipGetLastError();
      e = hipStreamEndCapture(ctx->stream, &g);
      if (e == hipSuccess) e = le;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&ex[which], g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      cleanup();
      return fail(ctx, VH_E_GPU, std::string("profile graph: ") + hipGetErrorString(e));
    }
  }
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  double us[2] = {0, 0};
  for (int which = 0; which < 2 && e == hipSuccess; ++which) {
    for (int w = 0; w < 2 && e == hipSuccess; ++w) e = hipGraphLaunch(ex[which], ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(e0, ctx->stream);
    for (uint32_t k = 0; k < reps && e == hipSuccess; ++k) e = hipGraphLaunch(ex[which], ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(e1, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    us[which] = 1000.0 * ms / reps;
  }
  cleanup();
  if (e != hipSuccess) return fail(ctx, VH_E_GPU, std::string("profile graph: ") + hipGetErrorString(e));
  out->local_us = us[0] / kLocalPerGraph;
  out->solve_us = us[1];
  return VH_OK;
}

vh_status vh_pd_lumped_masses(vh_ctx* ctx, float* out) {
  if (!ctx || !out) return VH_E_INVALID_ARG;
  if (!ctx->initialized) return fail(ctx, VH_E_NOT_INITIALIZED, "Solver not initialized. Call init() first.");
  for (int i = 0; i < ctx->S.n; ++i) out[i] = (float)ctx->S.mass_orig[i];
  return VH_OK;
}


// ---- IPC contact ------------------------------------------------------------

vh_status vh_ipc_init(vh_ctx* ctx, const vh_ipc_params* params) {
  if (!ctx) return VH_E_INVALID_ARG;
  if (!ctx->initialized) return fail(ctx, VH_E_NOT_INITIALIZED, "Solver not initialized. Call init() first.");
  if (!params) return fail(ctx, VH_E_INVALID_ARG, "params is null");
  const vh_ipc_params& Q = *params;
  if (!(Q.d_hat > 0) || !(Q.kappa >= 0) || !(Q.thickness >= 0) || Q.al_max_iterations < 1 || Q.exclusion_depth < 0 ||
      (Q.has_sphere && !(Q.sphere_radius > 0)) || (Q.compliant_contact && !(Q.compliant_d_hat_scale > 0)))
    return fail(ctx, VH_E_INVALID_CONFIG, "invalid contact parameters");
  if (ctx->ipc.on) return fail(ctx, VH_E_INVALID_CONFIG, "contact already initialized on this context");
  VH_HIP(hipSetDevice(ctx->device));
  const vh::Setup& S = ctx->S;
  const int n = S.n, t = S.t;
  IpcState& I = ctx->ipc;
  I.P = Q;
  I.sc.has_ground = Q.has_ground; I.sc.ground = Q.ground_height;
  I.sc.has_sphere = Q.has_sphere;
  I.sc.cx = Q.sphere_center[0]; I.sc.cy = Q.sphere_center[1]; I.sc.cz = Q.sphere_center[2];
  I.sc.radius = Q.sphere_radius;
  I.sc.self_collision = Q.self_collision;
  I.sc.thickness = Q.thickness;
  // d_hat_effective (pd_solver.rs:612-618) sets the CFL clamp; detection
  // uses it too unless the handler ignores set_d_hat (detect_keeps_d_hat).
  I.cfl_dh = Q.compliant_contact ? Q.d_hat * Q.compliant_d_hat_scale : Q.d_hat;
  I.sc.dh = Q.detect_keeps_d_hat ? Q.d_hat : I.cfl_dh;
  I.sc.kappa = Q.kappa;
  // Exclusion sets (exclusion.rs:22-41): v plus `depth` one-ring expansions,
  // in the device's vertex labels.
  std::vector<std::vector<int>> ring(n);
  for (int e = 0; e < t; ++e)
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        if (a != b) ring[S.tri_idx[3 * e + a]].push_back(S.tri_idx[3 * e + b]);
  for (auto& r : ring) { std::sort(r.begin(), r.end()); r.erase(std::unique(r.begin(), r.end()), r.end()); }
  std::vector<int> xptr(n + 1, 0), xidx, cur, next;
  for (int v = 0; v < n; ++v) {
    cur.assign(1, v);
    for (int d = 0; d < Q.exclusion_depth; ++d) {
      next = cur;
      for (int u : cur) next.insert(next.end(), ring[u].begin(), ring[u].end());
      std::sort(next.begin(), next.end());
      next.erase(std::unique(next.begin(), next.end()), next.end());
      cur.swap(next);
    }
    xidx.insert(xidx.end(), cur.begin(), cur.end());
    xptr[v + 1] = (int)xidx.size();
  }
  const uint64_t cap64 = Q.max_pairs ? (uint64_t)Q.max_pairs : 32ull * (uint64_t)std::max(1, t);
  if (cap64 > (1ull << 31)) return fail(ctx, VH_E_UNSUPPORTED, "max_pairs too large");
  I.cap = (unsigned int)cap64;
  const size_t n3 = 3 * (size_t)n;
  const int tt = std::max(1, t);
#define UP(call)                          \
  do {                                    \
    vh_status s_ = (call);                \
    if (s_ != VH_OK) return s_;           \
  } while (0)
  UP(dalloc(ctx, &I.gself, n3)); UP(dalloc(ctx, &I.grad, n3)); UP(dalloc(ctx, &I.nrm, n3));
  UP(dalloc(ctx, &I.fnrm, n3)); UP(dalloc(ctx, &I.lambda, n3)); UP(dalloc(ctx, &I.ext, n3));
  UP(dalloc(ctx, &I.scratch_a, n3)); UP(dalloc(ctx, &I.scratch_b, n3));
  UP(dalloc(ctx, &I.partials, 3 * (size_t)nblk(n)));
  UP(dalloc(ctx, &I.inc, (size_t)n)); UP(dalloc(ctx, &I.ever, (size_t)n));
  UP(upload(ctx, &I.xptr, xptr)); UP(upload(ctx, &I.xidx, xidx));
  UP(dalloc(ctx, &I.box, (size_t)tt)); UP(dalloc(ctx, &I.node, 2 * (size_t)tt));
  UP(dalloc(ctx, &I.code, (size_t)tt)); UP(dalloc(ctx, &I.code_s, (size_t)tt));
  UP(dalloc(ctx, &I.id, (size_t)tt)); UP(dalloc(ctx, &I.id_s, (size_t)tt));
  UP(dalloc(ctx, &I.left, (size_t)tt));
  UP(dalloc(ctx, &I.range_lo, (size_t)tt)); UP(dalloc(ctx, &I.range_hi, (size_t)tt));
  I.st_levels = 1;
  while ((1ll << I.st_levels) <= (long long)tt && I.st_levels <= vhc::kStLowLevels) ++I.st_levels;  // levels 0 .. 8
  UP(dalloc(ctx, &I.st, (size_t)I.st_levels * (size_t)tt));
  UP(dalloc(ctx, &I.coarse, (size_t)std::max(1, coarse_levels(tt >> 8)) * (size_t)std::max(1, tt >> 8)));
  UP(upload(ctx, &I.orig, S.tri_orig));
  UP(dalloc(ctx, &I.sorig, (size_t)tt));
  UP(dalloc(ctx, &I.pairs, (size_t)I.cap));
  UP(dalloc(ctx, &I.ccdtri, (size_t)tt));
  I.ccd_reg = ccd_region(I.cap);
  UP(dalloc(ctx, &I.ccd_surv, (size_t)pair_blocks(I.cap) * I.ccd_reg));
  UP(dalloc(ctx, &I.ccd_cnt, (size_t)kMaxPairBlocks));
  UP(dalloc(ctx, &I.cof, (size_t)pair_blocks(I.cap) * I.ccd_reg));
  UP(dalloc(ctx, &I.btot, (size_t)kMaxPairBlocks)); UP(dalloc(ctx, &I.sbase, (size_t)kMaxPairBlocks));
  I.ecap = (unsigned int)std::max(8192, t / 4);
  UP(dalloc(ctx, &I.ev, 4 * (size_t)I.ecap)); UP(dalloc(ctx, &I.eg, 12 * (size_t)I.ecap));
  UP(dalloc(ctx, &I.vidx, 4 * (size_t)I.ecap));
  UP(dalloc(ctx, &I.vcnt, (size_t)n + 1)); UP(dalloc(ctx, &I.vfill, (size_t)n + 1));
  UP(dalloc(ctx, &I.vptr, (size_t)n + 1));
  UP(dalloc(ctx, &I.bd, 1)); UP(dalloc(ctx, &I.ip, 1));
  // pinned: [0] the AL flag, [1] the pair count, [2] the self-contact count
  VH_HIP(hipHostMalloc((void**)&I.h_done, 4 * sizeof(int), hipHostMallocDefault));
  *I.h_done = 0;
  I.h_np = reinterpret_cast<unsigned int*>(I.h_done + 1);
  I.h_np[0] = I.h_np[1] = 0;
  UP(dalloc(ctx, &I.pcount, (size_t)tt)); UP(dalloc(ctx, &I.poff, (size_t)tt));
  // k_bvh_walk1 keeps 16 + 48 pairs per leaf (a flat sheet has about 25,
  // the drape's sheet where it wraps the sphere up to about 60)
  I.wmore_leaf = 48;
  UP(dalloc(ctx, &I.wkept, (size_t)nblk(tt) * vhc::kPairStage));
  UP(dalloc(ctx, &I.wmore, (size_t)tt * I.wmore_leaf));
  UP(dalloc(ctx, &I.bpart, 6 * (size_t)nblk(tt)));
  UP(dalloc(ctx, &I.sort_hist, (size_t)vhs::sort_hist_words(tt)));
  UP(dalloc(ctx, &I.sbuf, (size_t)std::max(vhs::scan_scratch_words(tt), vhs::scan_scratch_words((long long)n + 1))));
  UP(dalloc(ctx, &I.ticket, 1));
#undef UP
  VH_HIP(hipMemset(I.gself, 0, n3 * 8));
  VH_HIP(hipMemset(I.lambda, 0, n3 * 8));
  VH_HIP(hipMemset(I.fnrm, 0, n3 * 8));
  VH_HIP(hipMemset(I.ever, 0, (size_t)n));
  VH_HIP(hipMemset(I.ticket, 0, sizeof(unsigned int)));
  VH_HIP(hipMemset(I.bd, 0, sizeof(vhc::BvhDev)));
  VH_HIP(hipMemset(I.ip, 0, sizeof(vhc::IpcDev)));
  I.on = true;
  return VH_OK;
}

vh_status vh_pd_step_ipc_resident(vh_ctx* ctx, double dt, uint32_t nsteps, vh_step_result* out,
                                  vh_ipc_stats* stats) {
  if (!ctx) return VH_E_INVALID_ARG;
  if (!ctx->initialized) return fail(ctx, VH_E_NOT_INITIALIZED, "Solver not initialized. Call init() first.");
  if (!ctx->ipc.on) return fail(ctx, VH_E_NOT_INITIALIZED, "contact not initialized: call vh_ipc_init");
  if (!(dt > 0)) return fail(ctx, VH_E_INVALID_CONFIG, "dt must be > 0");
  VH_HIP(hipSetDevice(ctx->device));
  const auto t0 = std::chrono::steady_clock::now();
  for (uint32_t k = 0; k < nsteps; ++k) {
    VH_HIP(enqueue_ipc_step(ctx, dt));
  }
  vh_status st = read_ipc_result(ctx, out, stats);
  if (out) out->wall_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return st;
}

vh_status vh_pd_step_ipc(vh_ctx* ctx, double dt, vh_state_f32* s, float* lambda_x, float* lambda_y,
                         float* lambda_z, vh_step_result* out) {
  if (!ctx) return VH_E_INVALID_ARG;
  if (!ctx->initialized) return fail(ctx, VH_E_NOT_INITIALIZED, "Solver not initialized. Call init() first.");
  if (!ctx->ipc.on) return fail(ctx, VH_E_NOT_INITIALIZED, "contact not initialized: call vh_ipc_init");
  if (!(dt > 0)) return fail(ctx, VH_E_INVALID_CONFIG, "dt must be > 0");
  if (!s || !s->pos_x || !s->pos_y || !s->pos_z || !s->vel_x || !s->vel_y || !s->vel_z || !s->prev_x ||
      !s->prev_y || !s->prev_z || !s->pred_x || !s->pred_y || !s->pred_z || !lambda_x || !lambda_y || !lambda_z)
    return fail(ctx, VH_E_INVALID_ARG, "null state array");
  const auto t0 = std::chrono::steady_clock::now();
  VH_HIP(hipSetDevice(ctx->device));
  {
    vh_status e = io_reserve(ctx);
    if (e == VH_OK) e = io_upload(ctx, s, lambda_x, lambda_y, lambda_z);
    if (e != VH_OK) return e;
  }
  VH_HIP(enqueue_ipc_step(ctx, dt));
  {
    const vh_status e = io_download(ctx, true);
    if (e != VH_OK) return e;
  }
  vh_status st = read_ipc_result(ctx, out, nullptr);  // synchronizes
  io_scatter(ctx, s, lambda_x, lambda_y, lambda_z);
  if (out) out->wall_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return st;
}

vh_status vh_ipc_detect(vh_ctx* ctx, const double* pos, double d_hat, double kappa, double* grad, double* normal,
                        uint8_t* in_contact, double* hessian_diag, vh_ipc_stats* stats) {
  if (!ctx) return VH_E_INVALID_ARG;
  if (!ctx->initialized) return fail(ctx, VH_E_NOT_INITIALIZED, "Solver not initialized. Call init() first.");
  if (!ctx->ipc.on) return fail(ctx, VH_E_NOT_INITIALIZED, "contact not initialized: call vh_ipc_init");
  if (!(d_hat > 0) || !(kappa >= 0)) return fail(ctx, VH_E_INVALID_ARG, "d_hat > 0 and kappa >= 0 required");
  VH_HIP(hipSetDevice(ctx->device));
  const vh::Setup& S = ctx->S;
  const int n = S.n;
  const double* x = ctx->x;
  if (pos) {
    to_dev_order(S, pos, pos + n, pos + 2 * (size_t)n, ctx->hbuf);
    VH_HIP(hipMemcpyAsync(ctx->ipc.scratch_a, ctx->hbuf.data(), ctx->hbuf.size() * 8, hipMemcpyHostToDevice,
                          ctx->stream));
    x = ctx->ipc.scratch_a;
  }
  vhc::Scene sc = ctx->ipc.sc;
  sc.dh = d_hat;
  sc.kappa = kappa;
  auto detect = [&] {
    hipLaunchKernelGGL(k_ipc_clear, dim3(1), dim3(64), 0, ctx->stream, ctx->ipc.ip);
    enqueue_detect(ctx, x, sc, false, nullptr, ctx->ipc.scratch_b);
  };
  detect();
  VH_HIP(hipGetLastError());
  if (sc.self_collision) VH_HIP(fit_pairs(ctx, detect, true));
  const size_t n3 = 3 * (size_t)n;
  std::vector<double> g(n3), nr(n3), hd(n);
  std::vector<uint8_t> ic(n);
  VH_HIP(hipMemcpyAsync(hd.data(), ctx->ipc.scratch_b, (size_t)n * 8, hipMemcpyDeviceToHost, ctx->stream));
  VH_HIP(hipMemcpyAsync(g.data(), ctx->ipc.grad, n3 * 8, hipMemcpyDeviceToHost, ctx->stream));
  VH_HIP(hipMemcpyAsync(nr.data(), ctx->ipc.nrm, n3 * 8, hipMemcpyDeviceToHost, ctx->stream));
  VH_HIP(hipMemcpyAsync(ic.data(), ctx->ipc.inc, (size_t)n, hipMemcpyDeviceToHost, ctx->stream));
  vh_status st = read_ipc_result(ctx, nullptr, stats);
  if (st != VH_OK && st != VH_E_NONFINITE) return st;
  if (grad) from_dev_order(S, g, grad, grad + n, grad + 2 * (size_t)n);
  if (normal) from_dev_order(S, nr, normal, normal + n, normal + 2 * (size_t)n);
  if (in_contact)
    for (int k = 0; k < n; ++k) in_contact[S.perm[k]] = ic[k];
  if (hessian_diag)
    for (int k = 0; k < n; ++k) hessian_diag[S.perm[k]] = hd[k];
  return VH_OK;
}

vh_status vh_ipc_ccd_step(vh_ctx* ctx, const double* p0, const double* p1, double* alpha) {
  if (!ctx) return VH_E_INVALID_ARG;
  if (!ctx->initialized) return fail(ctx, VH_E_NOT_INITIALIZED, "Solver not initialized. Call init() first.");
  if (!ctx->ipc.on) return fail(ctx, VH_E_NOT_INITIALIZED, "contact not initialized: call vh_ipc_init");
  if (!p0 || !p1 || !alpha) return fail(ctx, VH_E_INVALID_ARG, "null array");
  VH_HIP(hipSetDevice(ctx->device));
  const vh::Setup& S = ctx->S;
  const int n = S.n;
  IpcState& I = ctx->ipc;
  to_dev_order(S, p0, p0 + n, p0 + 2 * (size_t)n, ctx->hbuf);
  std::vector<double> a = ctx->hbuf;
  to_dev_order(S, p1, p1 + n, p1 + 2 * (size_t)n, ctx->hbuf);
  for (size_t k = 0; k < a.size(); ++k) ctx->hbuf[k] -= a[k];  // delta = p1 - p0
  VH_HIP(hipMemcpyAsync(I.scratch_a, a.data(), a.size() * 8, hipMemcpyHostToDevice, ctx->stream));
  VH_HIP(hipMemcpyAsync(I.scratch_b, ctx->hbuf.data(), ctx->hbuf.size() * 8, hipMemcpyHostToDevice, ctx->stream));
  auto ccd = [&] {
    hipLaunchKernelGGL(k_ipc_clear, dim3(1), dim3(64), 0, ctx->stream, I.ip);
    if (I.sc.self_collision) {
      enqueue_pairs(ctx, I.scratch_a, (float)I.sc.thickness, nullptr);
      enqueue_ccd_pairs(ctx, I.scratch_a, I.scratch_b);
    }
    if (I.sc.has_ground || I.sc.has_sphere)
      hipLaunchKernelGGL(vhc::k_ccd_vertices, dim3(nblk(n)), dim3(256), 0, ctx->stream, n, I.scratch_a,
                         I.scratch_b, I.sc, I.ip);
  };
  ccd();
  VH_HIP(hipGetLastError());
  if (I.sc.self_collision) VH_HIP(fit_pairs(ctx, ccd, false));
  vhc::IpcDev h{};
  unsigned int np = 0;
  VH_HIP(hipMemcpyAsync(&h, I.ip, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  VH_HIP(hipMemcpyAsync(&np, &I.bd->npairs, sizeof(np), hipMemcpyDeviceToHost, ctx->stream));
  VH_HIP(hipStreamSynchronize(ctx->stream));
  if (I.sc.self_collision && np > I.cap) return fail(ctx, VH_E_INTERNAL, "candidate pairs exceeded the pair buffer");
  const double toi = std::min(1.0, __builtin_bit_cast(double, h.toi));
  *alpha = std::max(toi, 1e-6);
  return VH_OK;
}

vh_status vh_ipc_crossings(vh_ctx* ctx, const double* pos, uint64_t* count) {
  if (!ctx) return VH_E_INVALID_ARG;
  if (!ctx->initialized) return fail(ctx, VH_E_NOT_INITIALIZED, "Solver not initialized. Call init() first.");
  if (!ctx->ipc.on) return fail(ctx, VH_E_NOT_INITIALIZED, "contact not initialized: call vh_ipc_init");
  if (!count) return fail(ctx, VH_E_INVALID_ARG, "null count");
  VH_HIP(hipSetDevice(ctx->device));
  const vh::Setup& S = ctx->S;
  IpcState& I = ctx->ipc;
  const double* x = ctx->x;
  if (pos) {
    to_dev_order(S, pos, pos + S.n, pos + 2 * (size_t)S.n, ctx->hbuf);
    VH_HIP(hipMemcpyAsync(I.scratch_a, ctx->hbuf.data(), ctx->hbuf.size() * 8, hipMemcpyHostToDevice, ctx->stream));
    x = I.scratch_a;
  }
  auto cross = [&] {
    enqueue_pairs(ctx, x, 0.0f, nullptr);
    hipLaunchKernelGGL(vhc::k_tri_crossings, dim3(pair_blocks(I.cap)), dim3(256), 0, ctx->stream,
                       I.pairs, &I.bd->npairs, I.cap, ctx->tri_idx, x, I.bd);
  };
  cross();
  VH_HIP(hipGetLastError());
  VH_HIP(fit_pairs(ctx, cross, false));
  vhc::BvhDev h{};
  VH_HIP(hipMemcpyAsync(&h, I.bd, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  VH_HIP(hipStreamSynchronize(ctx->stream));
  if (h.npairs > I.cap) return fail(ctx, VH_E_INTERNAL, "candidate pairs exceeded the pair buffer");
  *count = h.crossings;
  return VH_OK;
}

vh_status vh_bvh_pairs(vh_ctx* ctx, const double* pos, double margin, uint32_t* ta, uint32_t* tb, uint64_t cap,
                       uint64_t* count) {
  if (!ctx) return VH_E_INVALID_ARG;
  if (!ctx->initialized) return fail(ctx, VH_E_NOT_INITIALIZED, "Solver not initialized. Call init() first.");
  if (!ctx->ipc.on) return fail(ctx, VH_E_NOT_INITIALIZED, "contact not initialized: call vh_ipc_init");
  if (!pos || !count || (cap > 0 && (!ta || !tb))) return fail(ctx, VH_E_INVALID_ARG, "null array");
  VH_HIP(hipSetDevice(ctx->device));
  const vh::Setup& S = ctx->S;
  const int n = S.n;
  IpcState& I = ctx->ipc;
  to_dev_order(S, pos, pos + n, pos + 2 * (size_t)n, ctx->hbuf);
  VH_HIP(hipMemcpyAsync(I.scratch_a, ctx->hbuf.data(), ctx->hbuf.size() * 8, hipMemcpyHostToDevice, ctx->stream));
  enqueue_pairs(ctx, I.scratch_a, (float)margin, nullptr);
  VH_HIP(hipGetLastError());
  VH_HIP(fit_pairs(ctx, [&] { enqueue_pairs(ctx, I.scratch_a, (float)margin, nullptr); }, false));
  unsigned int np = 0;
  VH_HIP(hipMemcpyAsync(&np, &I.bd->npairs, sizeof(np), hipMemcpyDeviceToHost, ctx->stream));
  VH_HIP(hipStreamSynchronize(ctx->stream));
  if (np > I.cap) return fail(ctx, VH_E_INTERNAL, "candidate pairs exceeded the pair buffer");
  std::vector<int2> hp(np);
  if (np) VH_HIP(hipMemcpy(hp.data(), I.pairs, (size_t)np * sizeof(int2), hipMemcpyDeviceToHost));
  std::vector<std::pair<uint32_t, uint32_t>> pr(np);
  for (unsigned int k = 0; k < np; ++k) {
    const uint32_t a = (uint32_t)S.tri_orig[hp[k].x], b = (uint32_t)S.tri_orig[hp[k].y];
    pr[k] = {std::min(a, b), std::max(a, b)};
  }
  std::sort(pr.begin(), pr.end());
  *count = np;
  for (uint64_t k = 0; k < std::min<uint64_t>(cap, np); ++k) { ta[k] = pr[k].first; tb[k] = pr[k].second; }
  return VH_OK;
}

vh_status vh_state_lambda_set_f64(vh_ctx* ctx, const double* lambda) {
  if (!ctx) return VH_E_INVALID_ARG;
  if (!ctx->initialized || !ctx->ipc.on) return fail(ctx, VH_E_NOT_INITIALIZED, "contact not initialized: call vh_ipc_init");
  if (!lambda) return fail(ctx, VH_E_INVALID_ARG, "null array");
  VH_HIP(hipSetDevice(ctx->device));
  const int n = ctx->S.n;
  to_dev_order(ctx->S, lambda, lambda + n, lambda + 2 * (size_t)n, ctx->hbuf);
  VH_HIP(hipMemcpyAsync(ctx->ipc.lambda, ctx->hbuf.data(), ctx->hbuf.size() * 8, hipMemcpyHostToDevice, ctx->stream));
  VH_HIP(hipStreamSynchronize(ctx->stream));
  return VH_OK;
}

vh_status vh_backend_axpy(vh_ctx* ctx, const float* a, uint64_t na, const float* b, uint64_t nb, float scale,
                          float* out, uint64_t nout) {
  if (!ctx) return VH_E_INVALID_ARG;
  if (na != nb || na != nout) return fail(ctx, VH_E_INVALID_ARG, "AXPY buffer length mismatch");
  if (na == 0) return VH_OK;
  if (!a || !b || !out) return fail(ctx, VH_E_INVALID_ARG, "null buffer");
  VH_HIP(hipSetDevice(ctx->device));
  vh_status st = backend_reserve(ctx, na);
  if (st != VH_OK) return st;
  VH_HIP(hipMemcpyAsync(ctx->bk_a, a, na * 4, hipMemcpyHostToDevice, ctx->stream));
  VH_HIP(hipMemcpyAsync(ctx->bk_b, b, na * 4, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_backend_axpy, dim3(backend_blocks(na)), dim3(256), 0, ctx->stream, (size_t)na, ctx->bk_a,
                     ctx->bk_b, scale, ctx->bk_o);
  VH_HIP(hipGetLastError());
  VH_HIP(hipMemcpyAsync(out, ctx->bk_o, na * 4, hipMemcpyDeviceToHost, ctx->stream));
  VH_HIP(hipStreamSynchronize(ctx->stream));
  return VH_OK;
}

vh_status vh_backend_dot(vh_ctx* ctx, const float* a, uint64_t na, const float* b, uint64_t nb, double* result) {
  if (!ctx) return VH_E_INVALID_ARG;
  if (!result) return fail(ctx, VH_E_INVALID_ARG, "null result");
  if (na != nb) return fail(ctx, VH_E_INVALID_ARG, "Dot product buffer length mismatch");
  *result = 0.0;
  if (na == 0) return VH_OK;
  if (!a || !b) return fail(ctx, VH_E_INVALID_ARG, "null buffer");
  VH_HIP(hipSetDevice(ctx->device));
  vh_status st = backend_reserve(ctx, na);
  if (st != VH_OK) return st;
  VH_HIP(hipMemcpyAsync(ctx->bk_a, a, na * 4, hipMemcpyHostToDevice, ctx->stream));
  VH_HIP(hipMemcpyAsync(ctx->bk_b, b, na * 4, hipMemcpyHostToDevice, ctx->stream));
  const unsigned int nbk = backend_blocks(na);
  hipLaunchKernelGGL(k_backend_dot, dim3(nbk), dim3(256), 0, ctx->stream, (size_t)na, ctx->bk_a, ctx->bk_b,
                     ctx->bk_part);
  hipLaunchKernelGGL(k_backend_dot_sum, dim3(1), dim3(256), 0, ctx->stream, (int)nbk, ctx->bk_part);
  VH_HIP(hipGetLastError());
  VH_HIP(hipMemcpyAsync(result, ctx->bk_part + kBackendBlocks, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  VH_HIP(hipStreamSynchronize(ctx->stream));
  return VH_OK;
}

vh_status vh_backend_dot_noalloc(vh_ctx* ctx, const float* a, uint64_t na, const float* b, uint64_t nb,
                                 double* result) {
  if (!ctx) return VH_E_INVALID_ARG;
  if (!result) return fail(ctx, VH_E_INVALID_ARG, "null result");
  if (na != nb) return fail(ctx, VH_E_INVALID_ARG, "Dot product buffer length mismatch");
  if (na > kNoAllocMaxElems)
    return fail(ctx, VH_E_INVALID_ARG,
                "vh_backend_dot_noalloc: " + std::to_string(na) + " elements exceed its maximum of " +
                    std::to_string(kNoAllocMaxElems) + " (use vh_backend_dot)");
  *result = 0.0;
  if (na == 0) return VH_OK;
  if (!a || !b) return fail(ctx, VH_E_INVALID_ARG, "null buffer");
  VH_HIP(hipSetDevice(ctx->device));
  unsigned char* base = static_cast<unsigned char*>(g_noalloc_scratch[ctx->device]);
  if (!base) return fail(ctx, VH_E_INTERNAL, "scratch address not set up for this device");
  double* part = reinterpret_cast<double*>(base);
  float* da = reinterpret_cast<float*>(base + kNoAllocPartBytes);
  float* db = da + kNoAllocMaxElems;
  VH_HIP(hipMemcpyAsync(da, a, na * 4, hipMemcpyHostToDevice, ctx->stream));
  VH_HIP(hipMemcpyAsync(db, b, na * 4, hipMemcpyHostToDevice, ctx->stream));
  const unsigned int nbk = backend_blocks(na);
  // every partial the sum reads is written by this call's first kernel
  hipLaunchKernelGGL(k_backend_dot, dim3(nbk), dim3(256), 0, ctx->stream, (size_t)na, da, db, part);
  hipLaunchKernelGGL(k_backend_dot_sum, dim3(1), dim3(256), 0, ctx->stream, (int)nbk, part);
  VH_HIP(hipGetLastError());
  VH_HIP(hipMemcpyAsync(result, part + kBackendBlocks, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  VH_HIP(hipStreamSynchronize(ctx->stream));
  return VH_OK;
}

vh_status vh_backend_fill(vh_ctx* ctx, float* buf, uint64_t n, float value) {
  if (!ctx) return VH_E_INVALID_ARG;
  if (n == 0) return VH_OK;
  if (!buf) return fail(ctx, VH_E_INVALID_ARG, "null buffer");
  VH_HIP(hipSetDevice(ctx->device));
  vh_status st = backend_reserve(ctx, n);
  if (st != VH_OK) return st;
  hipLaunchKernelGGL(k_backend_fill, dim3(backend_blocks(n)), dim3(256), 0, ctx->stream, (size_t)n, ctx->bk_o, value);
  VH_HIP(hipGetLastError());
  VH_HIP(hipMemcpyAsync(buf, ctx->bk_o, n * 4, hipMemcpyDeviceToHost, ctx->stream));
  VH_HIP(hipStreamSynchronize(ctx->stream));
  return VH_OK;
}

vh_status vh_narrow_vertex_triangle(vh_ctx* ctx, uint32_t n, const float* px, const float* py, const float* pz,
                                    uint32_t t, const uint32_t* tri, uint64_t ncand, const uint32_t* cand_a,
                                    const uint32_t* cand_b, float thickness, vh_contact_pair* out, uint64_t cap,
                                    uint64_t* count) {
  if (!ctx) return VH_E_INVALID_ARG;
  if (!count) return fail(ctx, VH_E_INVALID_ARG, "count is null");
  *count = 0;
  if (ncand && (!cand_a || !cand_b)) return fail(ctx, VH_E_INVALID_ARG, "candidates are null");
  if (n && (!px || !py || !pz)) return fail(ctx, VH_E_INVALID_ARG, "positions are null");
  if (t && !tri) return fail(ctx, VH_E_INVALID_ARG, "triangles are null");
  if (cap && !out) return fail(ctx, VH_E_INVALID_ARG, "out is null");
  if (ncand > (1ull << 31)) return fail(ctx, VH_E_UNSUPPORTED, "more than 2^31 candidates");
  for (uint64_t k = 0; k < ncand; ++k)
    if (cand_a[k] >= n || cand_b[k] >= n) return fail(ctx, VH_E_INVALID_ARG, "candidate vertex index out of range");
  for (uint64_t i = 0; i < 3 * (uint64_t)t; ++i)
    if (tri[i] >= n) return fail(ctx, VH_E_INVALID_MESH, "triangle index out of range");
  if (ncand == 0 || t == 0) return VH_OK;
  VH_HIP(hipSetDevice(ctx->device));
  // vertex -> triangles, ascending, a triangle once per vertex it contains
  std::vector<int> vptr((size_t)n + 1, 0), vtri;
  auto each_vertex = [&](uint32_t e, auto&& f) {
    const uint32_t A = tri[3 * (size_t)e], B = tri[3 * (size_t)e + 1], C = tri[3 * (size_t)e + 2];
    f(A);
    if (B != A) f(B);
    if (C != A && C != B) f(C);
  };
  for (uint32_t e = 0; e < t; ++e) each_vertex(e, [&](uint32_t v) { ++vptr[(size_t)v + 1]; });
  for (uint32_t i = 0; i < n; ++i) vptr[(size_t)i + 1] += vptr[i];
  vtri.resize((size_t)vptr[n]);
  std::vector<int> cur(vptr.begin(), vptr.end() - 1);
  for (uint32_t e = 0; e < t; ++e) each_vertex(e, [&](uint32_t v) { vtri[(size_t)cur[v]++] = (int)e; });
  uint64_t bound = 0;  // the most contacts the candidates can give
  for (uint64_t k = 0; k < ncand; ++k) bound += (uint64_t)(vptr[(size_t)cand_b[k] + 1] - vptr[cand_b[k]]);
  if (bound > 0xffffffffull) return fail(ctx, VH_E_UNSUPPORTED, "candidates could give more than 2^32 contacts");
  const uint64_t ocap = std::min<uint64_t>(cap, bound);
  // one allocation: ticket | positions | tri | vptr | vtri | ca | cb | cnt | off | tile sums | contacts
  size_t at = 256;
  auto carve = [&](size_t bytes) {
    const size_t o = at;
    at += (bytes + 255) / 256 * 256;
    return o;
  };
  const size_t o_pos = carve(3 * (size_t)n * 4), o_tri = carve(3 * (size_t)t * 4), o_vptr = carve(((size_t)n + 1) * 4),
               o_vtri = carve(vtri.size() * 4 + 4), o_ca = carve(ncand * 4), o_cb = carve(ncand * 4),
               o_cnt = carve(ncand * 4), o_off = carve(ncand * 4),
               o_bsum = carve((size_t)vhs::scan_scratch_words((long long)ncand) * 4),
               o_out = carve(std::max<uint64_t>(ocap, 1) * sizeof(vh_contact_pair));
  if (at > ctx->nb_bytes) {
    if (ctx->nb_arena) VH_HIP(hipFree(ctx->nb_arena));
    ctx->nb_arena = nullptr;
    ctx->nb_bytes = 0;
    VH_HIP(hipMalloc(&ctx->nb_arena, at));
    ctx->nb_bytes = at;
    VH_HIP(hipMemsetAsync(ctx->nb_arena, 0, 256, ctx->stream));
  }
  char* base = static_cast<char*>(ctx->nb_arena);
  float* dpos = reinterpret_cast<float*>(base + o_pos);
  auto* dcnt = reinterpret_cast<unsigned int*>(base + o_cnt);
  auto* doff = reinterpret_cast<unsigned int*>(base + o_off);
  auto* dout = reinterpret_cast<vh_contact_pair*>(base + o_out);
  hipStream_t s = ctx->stream;
  VH_HIP(hipMemcpyAsync(dpos, px, (size_t)n * 4, hipMemcpyHostToDevice, s));
  VH_HIP(hipMemcpyAsync(dpos + n, py, (size_t)n * 4, hipMemcpyHostToDevice, s));
  VH_HIP(hipMemcpyAsync(dpos + 2 * (size_t)n, pz, (size_t)n * 4, hipMemcpyHostToDevice, s));
  VH_HIP(hipMemcpyAsync(base + o_tri, tri, 3 * (size_t)t * 4, hipMemcpyHostToDevice, s));
  VH_HIP(hipMemcpyAsync(base + o_vptr, vptr.data(), vptr.size() * 4, hipMemcpyHostToDevice, s));
  VH_HIP(hipMemcpyAsync(base + o_vtri, vtri.data(), vtri.size() * 4, hipMemcpyHostToDevice, s));
  VH_HIP(hipMemcpyAsync(base + o_ca, cand_a, ncand * 4, hipMemcpyHostToDevice, s));
  VH_HIP(hipMemcpyAsync(base + o_cb, cand_b, ncand * 4, hipMemcpyHostToDevice, s));
  vhn::NarrowArgs a{dpos, dpos + n, dpos + 2 * (size_t)n, reinterpret_cast<const unsigned int*>(base + o_tri),
                    reinterpret_cast<const int*>(base + o_vptr), reinterpret_cast<const int*>(base + o_vtri),
                    reinterpret_cast<const unsigned int*>(base + o_ca), reinterpret_cast<const unsigned int*>(base + o_cb),
                    (long long)ncand, thickness};
  const dim3 grid((unsigned int)((ncand + 255) / 256));
  hipLaunchKernelGGL(vhn::k_narrow_vt<false>, grid, dim3(256), 0, s, a, dcnt, (const unsigned int*)nullptr,
                     (vh_contact_pair*)nullptr, 0ull);
  vhs::scan_exclusive(s, dcnt, doff, (long long)ncand, reinterpret_cast<unsigned int*>(base + o_bsum),
                      reinterpret_cast<unsigned int*>(base), nullptr);
  hipLaunchKernelGGL(vhn::k_narrow_vt<true>, grid, dim3(256), 0, s, a, (unsigned int*)nullptr,
                     (const unsigned int*)doff, dout, (unsigned long long)ocap);
  VH_HIP(hipGetLastError());
  unsigned int last[2] = {0, 0};
  VH_HIP(hipMemcpyAsync(&last[0], doff + (ncand - 1), 4, hipMemcpyDeviceToHost, s));
  VH_HIP(hipMemcpyAsync(&last[1], dcnt + (ncand - 1), 4, hipMemcpyDeviceToHost, s));
  VH_HIP(hipStreamSynchronize(s));
  const uint64_t total = (uint64_t)last[0] + last[1];
  const uint64_t nout = std::min<uint64_t>(total, ocap);
  if (nout) VH_HIP(hipMemcpy(out, dout, nout * sizeof(vh_contact_pair), hipMemcpyDeviceToHost));
  *count = total;
  return VH_OK;
}

vh_status vh_state_lambda_get_f64(vh_ctx* ctx, double* lambda) {
  if (!ctx) return VH_E_INVALID_ARG;
  if (!ctx->initialized || !ctx->ipc.on) return fail(ctx, VH_E_NOT_INITIALIZED, "contact not initialized: call vh_ipc_init");
  if (!lambda) return fail(ctx, VH_E_INVALID_ARG, "null array");
  VH_HIP(hipSetDevice(ctx->device));
  const int n = ctx->S.n;
  ctx->hbuf.resize(3 * (size_t)n);
  VH_HIP(hipMemcpyAsync(ctx->hbuf.data(), ctx->ipc.lambda, ctx->hbuf.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
  VH_HIP(hipStreamSynchronize(ctx->stream));
  from_dev_order(ctx->S, ctx->hbuf, lambda, lambda + n, lambda + 2 * (size_t)n);
  return VH_OK;
}

}  // extern "C"
