// C ABI of the gfx950 Projective Dynamics step (declared in include/vistio_hip.h).
//
// One vh_ctx owns one HIP stream and all device buffers of one solver
// instance. The step runs entirely on the device:
//   predict -> { local -> gather -> solve (forward levels, backward levels)
//                [-> refinement sweeps] -> commit } x max_iterations -> finalize
// following ProjectiveDynamicsSolver::step (crates/vistio-solver/src/pd_solver.rs:259-476).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "vh_kernels.hpp"
#include "vh_setup.hpp"
#include "vistio_hip.h"

namespace {

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

struct Graph {
  hipGraphExec_t exec = nullptr;
  double dt = 0;
  bool valid = false;
};

}  // namespace

struct vh_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  bool initialized = false;
  vh_pd_params P{};
  vh::Setup S;
  double setup_seconds = 0;
  int prec = VH_PREC_F64;
  size_t tsz = 8;  // bytes of the compute type
  // device buffers
  double *x = nullptr, *v = nullptr, *prev = nullptr, *pred = nullptr, *r = nullptr, *acc = nullptr;
  double* mass = nullptr;
  uint8_t* pinned = nullptr;
  int* tri_idx = nullptr;
  double* tri_dm = nullptr;
  void* tri_w = nullptr;
  int* hinge_idx = nullptr;
  void *hinge_rest = nullptr, *hinge_w = nullptr;
  void *ctri = nullptr, *chinge = nullptr;
  int *inc_ptr = nullptr, *inc = nullptr;
  vhk::SnDev* sn = nullptr;
  int4 *fwd_tasks = nullptr, *bwd_tasks = nullptr;
  int *rowidx = nullptr, *emap = nullptr, *children = nullptr;
  void *M = nullptr, *y = nullptr, *d = nullptr, *U = nullptr;
  int *Ap = nullptr, *Ai = nullptr;
  double* Ax = nullptr;
  double* partials = nullptr;
  vhk::StepDev* st = nullptr;
  std::vector<void*> allocs;
  size_t fwd_lds = 0, bwd_lds = 0;
  std::vector<double> hbuf;  // host staging (3n)
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

void free_all(vh_ctx* ctx) {
  for (void* p : ctx->allocs) (void)hipFree(p);
  ctx->allocs.clear();
  ctx->initialized = false;
}

inline int nblk(int n) { return (n + 255) / 256; }

// ---- kernel sequences -----------------------------------------------------

template <class T>
void enqueue_solve(vh_ctx* c, const vhk::StepDev* st) {
  const vh::Setup& S = c->S;
  for (int l = 0; l <= S.max_depth; ++l) {
    const int b = S.fwd_level_off[l], e = S.fwd_level_off[l + 1];
    if (e > b)
      hipLaunchKernelGGL(vhk::k_fwd<T>, dim3(e - b), dim3(256), c->fwd_lds, c->stream, c->fwd_tasks + b, c->sn,
                         c->emap, c->children, (const T*)c->M, c->r, (T*)c->y, (T*)c->U, st);
  }
  for (int l = 0; l <= S.max_depth; ++l) {
    const int b = S.bwd_level_off[l], e = S.bwd_level_off[l + 1];
    if (e > b)
      hipLaunchKernelGGL(vhk::k_bwd<T>, dim3(e - b), dim3(256), c->bwd_lds, c->stream, c->bwd_tasks + b, c->sn,
                         c->rowidx, (const T*)c->M, (const T*)c->y, (T*)c->d, st);
  }
}

template <class T>
void enqueue_local_gather(vh_ctx* c, double dt, const vhk::StepDev* st) {
  const vh::Setup& S = c->S;
  const int tb = nblk(S.t), hb = nblk(S.h);
  if (tb + hb > 0)
    hipLaunchKernelGGL(vhk::k_local<T>, dim3(tb + hb), dim3(256), 0, c->stream, S.t, S.h, tb,
                       c->P.model == VH_MODEL_COROTATIONAL ? 1 : 0, c->x, c->tri_idx, c->tri_dm, (const T*)c->tri_w,
                       c->hinge_idx, (const T*)c->hinge_rest, (const T*)c->hinge_w, (T*)c->ctri, (T*)c->chinge, st);
  const double inv_dt2 = 1.0 / (dt * dt);
  const double inv_dtm2 = 1.0 / (c->P.dt_matrix * c->P.dt_matrix);
  hipLaunchKernelGGL(vhk::k_gather<T>, dim3(nblk(S.n)), dim3(256), 0, c->stream, S.n, c->inc_ptr, c->inc,
                     (const T*)c->ctri, (const T*)c->chinge, c->mass, c->pred, c->x, inv_dt2, inv_dtm2, c->r, st);
}

// One solve of A * delta = r in the context's precision; the solution is in
// d (FP64 mode) or acc (FP32 refinement mode). r is consumed.
void enqueue_correction(vh_ctx* c, int sweeps, const vhk::StepDev* st) {
  if (c->prec == VH_PREC_F64) {
    enqueue_solve<double>(c, st);
    return;
  }
  for (int s = 0; s < sweeps; ++s) {
    enqueue_solve<float>(c, st);
    hipLaunchKernelGGL(vhk::k_ir_update, dim3(nblk(c->S.n)), dim3(256), 0, c->stream, c->S.n, c->Ap, c->Ai, c->Ax,
                       (const float*)c->d, c->r, c->acc, s == 0 ? 1 : 0, s + 1 < sweeps ? 1 : 0, st);
  }
}

void enqueue_step(vh_ctx* c, double dt) {
  const vh::Setup& S = c->S;
  const vh_pd_params& P = c->P;
  hipLaunchKernelGGL(vhk::k_predict, dim3(nblk(S.n)), dim3(256), 0, c->stream, S.n, c->x, c->v, c->prev, c->pred,
                     c->pinned, dt, P.gravity[0], P.gravity[1], P.gravity[2], P.has_ground, P.ground_height, c->st);
  for (int it = 0; it < P.max_iterations; ++it) {
    const bool last = it + 1 == P.max_iterations;
    if (c->prec == VH_PREC_F64) enqueue_local_gather<double>(c, dt, c->st);
    else enqueue_local_gather<float>(c, dt, c->st);
    enqueue_correction(c, last ? P.ir_final_sweeps : P.ir_sweeps, c->st);
    if (c->prec == VH_PREC_F64)
      hipLaunchKernelGGL(vhk::k_commit<double>, dim3(nblk(S.n)), dim3(256), 0, c->stream, S.n, c->x,
                         (const double*)c->d, c->prev, c->pinned, P.has_ground, P.ground_height, P.tolerance, it,
                         c->partials, c->st);
    else
      hipLaunchKernelGGL(vhk::k_commit<double>, dim3(nblk(S.n)), dim3(256), 0, c->stream, S.n, c->x,
                         (const double*)c->acc, c->prev, c->pinned, P.has_ground, P.ground_height, P.tolerance, it,
                         c->partials, c->st);
  }
  const double frayl = P.rayleigh > 0 ? 1.0 / (1.0 + P.rayleigh * dt) : 0.0;
  hipLaunchKernelGGL(vhk::k_finalize, dim3(nblk(S.n)), dim3(256), 0, c->stream, S.n, c->x, c->prev, c->v, c->pinned,
                     1.0 / dt, P.has_ground, P.ground_height, 1.0 - P.damping, frayl);
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
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
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
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return VH_OK;
}

const char* vh_last_error(const vh_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

vh_status vh_pd_init(vh_ctx* ctx, uint32_t n, const float* px, const float* py, const float* pz, uint32_t t,
                     const uint32_t* tri, uint32_t h, const uint32_t* hinge, const uint8_t* pinned,
                     const vh_pd_params* params, float* lumped_mass_out) {
  if (!ctx) return VH_E_INVALID_ARG;
  if (!params) return fail(ctx, VH_E_INVALID_ARG, "params is null");
  VH_HIP(hipSetDevice(ctx->device));
  VH_HIP(hipStreamSynchronize(ctx->stream));
  free_all(ctx);
  const vh_pd_params& P = *params;
  if (P.model != VH_MODEL_ARAP && P.model != VH_MODEL_COROTATIONAL)
    return fail(ctx, VH_E_UNSUPPORTED, "unsupported constitutive model");
  if (P.bending != VH_BENDING_NONE && P.bending != VH_BENDING_DIHEDRAL)
    return fail(ctx, VH_E_UNSUPPORTED, "unsupported bending model");
  if (P.precision != VH_PREC_F64 && P.precision != VH_PREC_F32_IR)
    return fail(ctx, VH_E_UNSUPPORTED, "unsupported precision mode");
  if (P.max_iterations < 0) return fail(ctx, VH_E_INVALID_CONFIG, "max_iterations must be >= 0");
  if (P.precision == VH_PREC_F32_IR &&
      (P.ir_sweeps < 1 || P.ir_sweeps > 3 || P.ir_final_sweeps < 1 || P.ir_final_sweeps > 3))
    return fail(ctx, VH_E_INVALID_CONFIG, "ir_sweeps and ir_final_sweeps must be in 1..3");
  if (n == 0) return fail(ctx, VH_E_INVALID_MESH, "empty mesh");
  ctx->P = P;
  ctx->prec = P.precision;
  ctx->tsz = P.precision == VH_PREC_F64 ? 8 : 4;
  const auto t0 = std::chrono::steady_clock::now();
  vh::SetupInput in;
  in.n = (int)n; in.t = (int)t; in.h = P.bending == VH_BENDING_NONE ? 0 : (int)h;
  in.px = px; in.py = py; in.pz = pz; in.tri = tri; in.hinge = hinge; in.pinned = pinned;
  in.k_membrane = P.k_membrane; in.k_bending = P.k_bending; in.density = P.density; in.dt_matrix = P.dt_matrix;
  in.leaf_size = P.leaf_size > 0 ? P.leaf_size : 16;
  ctx->S = vh::Setup();
  if (!vh::build_setup(in, ctx->S)) return fail(ctx, (vh_status)ctx->S.status, ctx->S.err);
  vh::Setup& S = ctx->S;
  ctx->setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (lumped_mass_out)
    for (int i = 0; i < S.n; ++i) lumped_mass_out[i] = (float)S.mass_orig[i];

  vh_status s = VH_OK;
#define UP(expr)                 \
  do {                           \
    s = (expr);                  \
    if (s != VH_OK) return s;    \
  } while (0)
  const size_t n3 = 3 * (size_t)S.n;
  UP(dalloc(ctx, &ctx->x, n3)); UP(dalloc(ctx, &ctx->v, n3));
  UP(dalloc(ctx, &ctx->prev, n3)); UP(dalloc(ctx, &ctx->pred, n3));
  UP(dalloc(ctx, &ctx->r, n3)); UP(dalloc(ctx, &ctx->acc, n3));
  UP(upload(ctx, &ctx->mass, S.mass));
  UP(upload(ctx, &ctx->pinned, S.pinned));
  UP(upload(ctx, &ctx->tri_idx, S.tri_idx));
  UP(upload(ctx, &ctx->tri_dm, S.tri_dm));
  UP(upload_t(ctx, &ctx->tri_w, S.tri_w));
  UP(upload(ctx, &ctx->hinge_idx, S.hinge_idx));
  UP(upload_t(ctx, &ctx->hinge_rest, S.hinge_rest));
  UP(upload_t(ctx, &ctx->hinge_w, S.hinge_w));
  UP(dalloc_bytes(ctx, &ctx->ctri, 9 * (size_t)S.t * ctx->tsz));
  UP(dalloc_bytes(ctx, &ctx->chinge, 3 * (size_t)S.h * ctx->tsz));
  UP(upload(ctx, &ctx->inc_ptr, S.inc_ptr));
  UP(upload(ctx, &ctx->inc, S.inc));
  std::vector<vhk::SnDev> snd(S.sn.size());
  for (size_t k = 0; k < S.sn.size(); ++k) {
    const vh::Supernode& q = S.sn[k];
    snd[k] = vhk::SnDev{q.col0, q.w, q.r, q.rows_off, q.u_off, q.map_off, q.child_off, q.nchild, (long long)q.m_off};
  }
  UP(upload(ctx, &ctx->sn, snd));
  std::vector<int4> ft(S.fwd_tasks.size()), bt(S.bwd_tasks.size());
  for (size_t k = 0; k < ft.size(); ++k) ft[k] = make_int4(S.fwd_tasks[k].sn, S.fwd_tasks[k].b, S.fwd_tasks[k].e, S.fwd_tasks[k].g);
  for (size_t k = 0; k < bt.size(); ++k) bt[k] = make_int4(S.bwd_tasks[k].sn, S.bwd_tasks[k].b, S.bwd_tasks[k].e, S.bwd_tasks[k].g);
  UP(upload(ctx, &ctx->fwd_tasks, ft));
  UP(upload(ctx, &ctx->bwd_tasks, bt));
  UP(upload(ctx, &ctx->rowidx, S.rowidx));
  UP(upload(ctx, &ctx->emap, S.emap));
  UP(upload(ctx, &ctx->children, S.children));
  UP(upload_t(ctx, &ctx->M, S.M));
  UP(dalloc_bytes(ctx, &ctx->y, n3 * ctx->tsz));
  UP(dalloc_bytes(ctx, &ctx->d, n3 * ctx->tsz));
  UP(dalloc_bytes(ctx, &ctx->U, 3 * (size_t)std::max(1, S.u_total) * ctx->tsz));
  if (ctx->prec == VH_PREC_F32_IR) {
    UP(upload(ctx, &ctx->Ap, S.Ap));
    UP(upload(ctx, &ctx->Ai, S.Ai));
    UP(upload(ctx, &ctx->Ax, S.Ax));
  }
  UP(dalloc(ctx, &ctx->partials, 2 * (size_t)nblk(S.n)));
  UP(dalloc(ctx, &ctx->st, 1));
#undef UP
  VH_HIP(hipMemset(ctx->st, 0, sizeof(vhk::StepDev)));
  ctx->fwd_lds = 3 * (size_t)S.max_front * ctx->tsz;
  ctx->bwd_lds = (3 * (size_t)S.max_front + 3 * 256) * ctx->tsz;
  const size_t lds_max = 160 * 1024;
  if (ctx->bwd_lds > lds_max)
    return fail(ctx, VH_E_UNSUPPORTED, "supernode front exceeds the 160 KB LDS of one workgroup");
  if (ctx->fwd_lds > 64 * 1024) {
    if (ctx->prec == VH_PREC_F64) VH_HIP(hipFuncSetAttribute((const void*)vhk::k_fwd<double>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ctx->fwd_lds));
    else VH_HIP(hipFuncSetAttribute((const void*)vhk::k_fwd<float>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ctx->fwd_lds));
  }
  if (ctx->bwd_lds > 64 * 1024) {
    if (ctx->prec == VH_PREC_F64) VH_HIP(hipFuncSetAttribute((const void*)vhk::k_bwd<double>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ctx->bwd_lds));
    else VH_HIP(hipFuncSetAttribute((const void*)vhk::k_bwd<float>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ctx->bwd_lds));
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

vh_status vh_pd_step_resident(vh_ctx* ctx, double dt, uint32_t nsteps, vh_step_result* out) {
  if (!ctx) return VH_E_INVALID_ARG;
  if (!ctx->initialized) return fail(ctx, VH_E_NOT_INITIALIZED, "Solver not initialized. Call init() first.");
  if (!(dt > 0)) return fail(ctx, VH_E_INVALID_CONFIG, "dt must be > 0");
  VH_HIP(hipSetDevice(ctx->device));
  const auto t0 = std::chrono::steady_clock::now();
  for (uint32_t s = 0; s < nsteps; ++s) enqueue_step(ctx, dt);
  VH_HIP(hipGetLastError());
  vh_status st = read_step_result(ctx, out);
  if (out) out->wall_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return st;
}

vh_status vh_pd_step(vh_ctx* ctx, double dt, vh_state_f32* s, vh_step_result* out) {
  if (!ctx) return VH_E_INVALID_ARG;
  if (!ctx->initialized) return fail(ctx, VH_E_NOT_INITIALIZED, "Solver not initialized. Call init() first.");
  if (!s || !s->pos_x || !s->pos_y || !s->pos_z || !s->vel_x || !s->vel_y || !s->vel_z || !s->prev_x ||
      !s->prev_y || !s->prev_z || !s->pred_x || !s->pred_y || !s->pred_z)
    return fail(ctx, VH_E_INVALID_ARG, "null state array");
  const auto t0 = std::chrono::steady_clock::now();
  const vh::Setup& S = ctx->S;
  const size_t n3 = 3 * (size_t)S.n;
  std::vector<double> hp(n3), hv(n3);
  for (int k = 0; k < S.n; ++k) {
    const int o = S.perm[k];
    hp[3 * k] = s->pos_x[o]; hp[3 * k + 1] = s->pos_y[o]; hp[3 * k + 2] = s->pos_z[o];
    hv[3 * k] = s->vel_x[o]; hv[3 * k + 1] = s->vel_y[o]; hv[3 * k + 2] = s->vel_z[o];
  }
  VH_HIP(hipSetDevice(ctx->device));
  if (s->inv_mass) {
    vh_status e = vh_state_set_inv_mass(ctx, s->inv_mass);
    if (e != VH_OK) return e;
  }
  VH_HIP(hipMemcpyAsync(ctx->x, hp.data(), n3 * 8, hipMemcpyHostToDevice, ctx->stream));
  VH_HIP(hipMemcpyAsync(ctx->v, hv.data(), n3 * 8, hipMemcpyHostToDevice, ctx->stream));
  enqueue_step(ctx, dt);
  VH_HIP(hipGetLastError());
  std::vector<double> hq(n3), hr(n3);
  VH_HIP(hipMemcpyAsync(hp.data(), ctx->x, n3 * 8, hipMemcpyDeviceToHost, ctx->stream));
  VH_HIP(hipMemcpyAsync(hv.data(), ctx->v, n3 * 8, hipMemcpyDeviceToHost, ctx->stream));
  VH_HIP(hipMemcpyAsync(hq.data(), ctx->prev, n3 * 8, hipMemcpyDeviceToHost, ctx->stream));
  VH_HIP(hipMemcpyAsync(hr.data(), ctx->pred, n3 * 8, hipMemcpyDeviceToHost, ctx->stream));
  vh_status st = read_step_result(ctx, out);
  for (int k = 0; k < S.n; ++k) {
    const int o = S.perm[k];
    s->pos_x[o] = (float)hp[3 * k]; s->pos_y[o] = (float)hp[3 * k + 1]; s->pos_z[o] = (float)hp[3 * k + 2];
    s->vel_x[o] = (float)hv[3 * k]; s->vel_y[o] = (float)hv[3 * k + 1]; s->vel_z[o] = (float)hv[3 * k + 2];
    s->prev_x[o] = (float)hq[3 * k]; s->prev_y[o] = (float)hq[3 * k + 1]; s->prev_z[o] = (float)hq[3 * k + 2];
    s->pred_x[o] = (float)hr[3 * k]; s->pred_y[o] = (float)hr[3 * k + 1]; s->pred_z[o] = (float)hr[3 * k + 2];
  }
  if (out) out->wall_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return st;
}

vh_status vh_pd_residual(vh_ctx* ctx, double dt, double* out) {
  if (!ctx) return VH_E_INVALID_ARG;
  if (!ctx->initialized) return fail(ctx, VH_E_NOT_INITIALIZED, "Solver not initialized. Call init() first.");
  if (!out) return fail(ctx, VH_E_INVALID_ARG, "out is null");
  if (!(dt > 0)) return fail(ctx, VH_E_INVALID_CONFIG, "dt must be > 0");
  VH_HIP(hipSetDevice(ctx->device));
  if (ctx->prec == VH_PREC_F64) enqueue_local_gather<double>(ctx, dt, nullptr);
  else enqueue_local_gather<float>(ctx, dt, nullptr);
  VH_HIP(hipGetLastError());
  const size_t n3 = 3 * (size_t)ctx->S.n;
  ctx->hbuf.resize(n3);
  VH_HIP(hipMemcpyAsync(ctx->hbuf.data(), ctx->r, n3 * 8, hipMemcpyDeviceToHost, ctx->stream));
  VH_HIP(hipStreamSynchronize(ctx->stream));
  const int n = ctx->S.n;
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
  enqueue_correction(ctx, sweeps, nullptr);
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
  return VH_OK;
}

vh_status vh_pd_lumped_masses(vh_ctx* ctx, float* out) {
  if (!ctx || !out) return VH_E_INVALID_ARG;
  if (!ctx->initialized) return fail(ctx, VH_E_NOT_INITIALIZED, "Solver not initialized. Call init() first.");
  for (int i = 0; i < ctx->S.n; ++i) out[i] = (float)ctx->S.mass_orig[i];
  return VH_OK;
}

}  // extern "C"
