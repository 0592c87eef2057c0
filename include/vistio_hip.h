/* vistio_hip.h — C ABI of the MI355X (gfx950) Projective Dynamics step.
 *
 * Plain C types only (no torch, no HIP types). One context = one HIP stream
 * on one device; a context is not re-entrant (the Rust side holds it behind
 * `&mut self`). All host buffers are caller-owned; the context owns device
 * memory. No exception or panic crosses this boundary: every entry point
 * returns a vh_status and vh_last_error() gives the message.
 *
 * Reference interfaces each entry point replaces (file:line in the Vistio
 * tree; the Rust-side binding is the crate vistio_amd/rust/vistio-hip):
 *   SolverStrategy            crates/vistio-solver/src/strategy.rs:42-69
 *   ProjectiveDynamicsSolver  crates/vistio-solver/src/pd_solver.rs:53-476
 *   BroadPhase, NarrowPhase   crates/vistio-contact/src/broad.rs:39-56, narrow.rs:21-35
 *   GpuBackend                crates/vistio-gpu/src/backend.rs:19-46
 *   VistioError               crates/vistio-types/src/error.rs:9-44
 */
#ifndef VISTIO_HIP_H
#define VISTIO_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. Mapping to VistioError (crates/vistio-types/src/error.rs):
 *   VH_E_INVALID_CONFIG, VH_E_NOT_INITIALIZED, VH_E_NOT_SPD, VH_E_UNSUPPORTED
 *                         -> VistioError::InvalidConfig
 *   VH_E_INVALID_MESH     -> VistioError::InvalidMesh
 *   VH_E_GPU, VH_E_INTERNAL -> VistioError::Gpu
 *   VH_E_INVALID_ARG      -> VistioError::InvalidConfig (null pointer / bad size)
 *   VH_E_NONFINITE        -> VistioError::InvalidConfig (NaN/Inf in the state) */
typedef enum {
  VH_OK = 0,
  VH_E_INVALID_CONFIG = 1,
  VH_E_NOT_INITIALIZED = 2, /* pd_solver.rs:264-268 "Solver not initialized" */
  VH_E_NOT_SPD = 3,         /* pd_solver.rs:135-137 "Cholesky factorization failed" */
  VH_E_UNSUPPORTED = 4,     /* unknown ConstitutiveModel::name() / bending kind */
  VH_E_INVALID_MESH = 5,    /* state.rs:92-105 length mismatch, bad indices */
  VH_E_GPU = 6,             /* any HIP runtime error */
  VH_E_INTERNAL = 7,
  VH_E_INVALID_ARG = 8,
  VH_E_NONFINITE = 9
} vh_status;

/* Membrane projection, keyed by ConstitutiveModel::name()
 * (crates/vistio-material/src/):
 *   VH_MODEL_ARAP                       ElementData::project (element.rs:127-144), used by init()
 *   VH_MODEL_COROTATIONAL               "co_rotational" (corotational.rs:51-74)
 *   VH_MODEL_ISOTROPIC_LINEAR           "isotropic_linear" (isotropic.rs:40-64): target F = I
 *   VH_MODEL_ORTHOTROPIC_LINEAR         "orthotropic_linear" (orthotropic.rs:52-94),
 *                                       warp_stiffness / weft_stiffness
 *   VH_MODEL_ANISOTROPIC_COROTATIONAL   "anisotropic_corotational" (anisotropic.rs:78-143),
 *                                       warp_stiffness / weft_stiffness
 * Any other value returns VH_E_UNSUPPORTED. */
enum {
  VH_MODEL_ARAP = 0,
  VH_MODEL_COROTATIONAL = 1,
  VH_MODEL_ISOTROPIC_LINEAR = 2,
  VH_MODEL_ORTHOTROPIC_LINEAR = 3,
  VH_MODEL_ANISOTROPIC_COROTATIONAL = 4
};
/* Bending: dihedral springs (crates/vistio-solver/src/bending.rs) or Discrete
 * Shells (discrete_shells.rs: normalised cotangent stencil per hinge, the
 * same projection; init_with_material_tier3, pd_solver.rs:152-193). */
enum { VH_BENDING_NONE = 0, VH_BENDING_DIHEDRAL = 1, VH_BENDING_DISCRETE_SHELLS = 2 };
/* Precision of the local step and the triangular solves. State (positions,
 * velocities) is always held in FP64 on the device.
 *   VH_PREC_F64     FP64 local step, FP64 solve.
 *   VH_PREC_F32_IR  FP32 local step (delta form), FP32 solve, FP64 residual
 *                   r = b - A x with iterative refinement. */
enum { VH_PREC_F64 = 0, VH_PREC_F32_IR = 1 };

/* SolverConfig (crates/vistio-solver/src/config.rs:10-129) fields on the
 * path, plus the material values init_with_material derives
 * (element.rs:106 -> k_membrane, bending.rs:160 -> k_bending,
 * pd_solver.rs:1221 -> density) and the HIP-specific knobs. */
typedef struct {
  int32_t model;            /* VH_MODEL_* */
  int32_t bending;          /* VH_BENDING_* */
  double k_membrane;        /* element weight = k_membrane * rest area */
  double k_bending;         /* hinge weight = k_bending * rest edge length */
  double density;           /* kg/m^2 */
  double dt_matrix;         /* dt in A = M/dt^2 + ... (the reference fixes 1/60) */
  double gravity[3];
  double damping;           /* v *= 1 - damping */
  double rayleigh;          /* free vertices: v *= 1/(1 + rayleigh*dt) */
  int32_t max_iterations;
  double tolerance;         /* early exit when the relative update < tolerance */
  int32_t has_ground;
  double ground_height;
  int32_t precision;        /* VH_PREC_* */
  int32_t ir_sweeps;        /* VH_PREC_F32_IR: refinement sweeps per PD iteration (1..3) */
  int32_t ir_final_sweeps;  /* VH_PREC_F32_IR: sweeps on the last iteration (1..3) */
  int32_t leaf_size;        /* nested-dissection leaf size; 0 = default (16) */
  /* Solve schedule: consecutive dissection levels are merged into bands, one
   * dense block per band node, chosen to minimise
   * 2*(band_launch_us + block bytes / band_gbps) per band. 0 = defaults
   * (3 levels max in FP64, 4 in FP32-IR, 6 us, 4000 GB/s); max_band_levels
   * = 1 disables merging. */
  int32_t max_band_levels;
  double band_launch_us;
  double band_gbps;
  int32_t task_entries;     /* block entries per solve task; 0 = default (1024) */
  /* FabricProperties::stretch_stiffness_warp / _weft for the orthotropic and
   * anisotropic membrane models (their from_properties); unused otherwise. */
  double warp_stiffness, weft_stiffness;
  /* FabricProperties::bending_stiffness_warp / _weft (raw values): with
   * VH_BENDING_DISCRETE_SHELLS and a nonzero pair, each hinge weight is scaled
   * by its edge's warp/weft alignment as
   * DiscreteShellsBendingData::from_topology_with_anisotropic_material
   * (discrete_shells.rs:206-268); k_bending is then the average times 100. */
  double bend_warp, bend_weft;
} vh_pd_params;

/* StepResult, crates/vistio-solver/src/strategy.rs:15-24. */
typedef struct {
  uint32_t iterations;
  double final_residual;
  int32_t converged;
  double wall_time; /* seconds, host clock around the step(s) */
} vh_step_result;

/* SimulationState SoA buffers (crates/vistio-solver/src/state.rs:23-64). */
typedef struct {
  float* pos_x; float* pos_y; float* pos_z;
  float* vel_x; float* vel_y; float* vel_z;
  float* prev_x; float* prev_y; float* prev_z;
  float* pred_x; float* pred_y; float* pred_z;
  const float* inv_mass; /* pinned vertices have inv_mass == 0 */
} vh_state_f32;

typedef struct {
  uint32_t n, t, h;
  uint32_t supernodes, levels;
  int64_t nnz_a;      /* stored entries of A (full symmetric) */
  int64_t nnz_factor; /* entries of all solve blocks M_P (= supernodal L) */
  int32_t max_front;  /* max (w + r) over supernodes */
  double setup_seconds;   /* host set-up: rest data, matrix, ordering, symbolic structure */
  double factor_seconds;  /* device factorization into the solve blocks */
} vh_pd_info;

typedef struct vh_ctx vh_ctx;

/* Context lifetime. */
vh_status vh_create(int32_t device, vh_ctx** out);
vh_status vh_destroy(vh_ctx* ctx);
const char* vh_last_error(const vh_ctx* ctx);

/* ProjectiveDynamicsSolver::init / init_with_material (pd_solver.rs:106-143,
 * 221-257): rest data, lumped masses, system matrix, factorization, device
 * upload. hinge holds 4h indices (v0, v1, wing_a, wing_b) of the interior
 * edges (crates/vistio-mesh/src/topology.rs:36-51). lumped_mass_out (n, may
 * be NULL) receives lumped_masses() (pd_solver.rs:95-97). The device state is
 * initialised from px/py/pz with zero velocity (state.rs:119-139). */
vh_status vh_pd_init(vh_ctx* ctx, uint32_t n, const float* px, const float* py, const float* pz,
                     uint32_t t, const uint32_t* tri, uint32_t h, const uint32_t* hinge,
                     const uint8_t* pinned, const vh_pd_params* params, float* lumped_mass_out);

/* Per-element materials, for several garments stepped as one system (their
 * meshes concatenated; the system matrix is then block diagonal and one
 * launch sequence steps them all). Each array may be NULL (the uniform value
 * of vh_pd_params is used): k_membrane per triangle (t), k_bending per hinge
 * (h), density per triangle (t, kg/m^2; a vertex gets a third of each
 * incident triangle's area * density, pd_solver.rs:1219-1242). Every garment
 * corresponds to one ProjectiveDynamicsSolver::init_with_material
 * (pd_solver.rs:106-143) with its own FabricProperties. */
typedef struct {
  const double* k_membrane;
  const double* k_bending;
  const double* density;
} vh_material_arrays;
vh_status vh_pd_init_ex(vh_ctx* ctx, uint32_t n, const float* px, const float* py, const float* pz,
                        uint32_t t, const uint32_t* tri, uint32_t h, const uint32_t* hinge,
                        const uint8_t* pinned, const vh_pd_params* params, const vh_material_arrays* mat,
                        float* lumped_mass_out);

/* SolverStrategy::step (pd_solver.rs:259-476) on host SoA state: uploads
 * pos/vel and inv_mass, runs one step on the device, downloads
 * pos/vel/prev/pred. */
vh_status vh_pd_step(vh_ctx* ctx, double dt, vh_state_f32* state, vh_step_result* out);

/* Resident mode: the state stays on the device between calls. */
vh_status vh_pd_step_resident(vh_ctx* ctx, double dt, uint32_t nsteps, vh_step_result* out);
/* which: 0 = pos, 1 = vel, 2 = prev, 3 = pred. Arrays are SoA x[n], y[n], z[n]
 * in the caller's vertex order. */
vh_status vh_state_set_f64(vh_ctx* ctx, int32_t which, const double* x, const double* y, const double* z);
vh_status vh_state_get_f64(vh_ctx* ctx, int32_t which, double* x, double* y, double* z);
/* Pinning flags from SimulationState::inv_mass (inv_mass[i] == 0 -> pinned). */
vh_status vh_state_set_inv_mass(vh_ctx* ctx, const float* inv_mass);
/* SimulationState::ground_height (state.rs:52-55) for the plain step: when
 * set, predict clamps y to it (state.rs:160-165), every PD iteration's commit
 * enforces it (enforce_ground, state.rs:180-193, pd_solver.rs:437-438) and
 * the velocity update applies enforce_ground_velocities (state.rs:195-212,
 * pd_solver.rs:450-451). has_ground = 0 turns it off. Initialised from
 * vh_pd_params. */
vh_status vh_pd_set_ground(vh_ctx* ctx, int32_t has_ground, double ground_height);

/* Parity / inspection entry points. */
/* r = b - A x at the current device state (pred as set): the local step and
 * RHS assembly (pd_solver.rs:296-389, assembly.rs:157-240) in residual form.
 * out: SoA 3n, caller's vertex order. dt as in step. */
vh_status vh_pd_residual(vh_ctx* ctx, double dt, double* out);
/* x = A^{-1} b with the device solve (FaerSolver::solve x3,
 * crates/vistio-math/src/faer_solver.rs:95-129), SoA 3n in the caller's order.
 * In VH_PREC_F32_IR mode runs `sweeps` refinement sweeps (>= 1). */
vh_status vh_pd_solve(vh_ctx* ctx, const double* b, double* x, int32_t sweeps);
vh_status vh_pd_info_get(vh_ctx* ctx, vh_pd_info* out);
/* The solve blocks the device factorization built (FaerSolver::factorize,
 * crates/vistio-math/src/faer_solver.rs:67-93, as partitioned-inverse
 * blocks), widened to double: which = 0 M (each supernode's
 * [L_PP^{-1}; L_IP L_PP^{-1}] row-major, lower rows packed), 1 M^T, 2 the
 * roots' dense inverses (L L^T)^{-1}. Writes up to cap values and the total
 * count (cap = 0 queries the count). */
vh_status vh_pd_blocks_get(vh_ctx* ctx, int32_t which, double* out, uint64_t cap, uint64_t* count);

/* Per-kernel timing of PD iterations, measured with HIP events on the
 * context's stream around every launch (no graph). An iteration is the
 * local step plus `sweeps` solves (ir_sweeps in FP32 mode, 1 in FP64), as in
 * every iteration of a step but the last. A solve launches the forward
 * bands deepest first (fwd_*[0 .. bands-2]), the root band, whose forward
 * and backward sweeps are one launch (fwd_*[bands-1]), then the backward
 * bands below the root, shallowest first (bwd_*[1 .. bands-1]; bwd_*[0] is
 * 0). Times and bytes are per iteration, summed over its sweeps; bytes are
 * the algorithmic bytes the launches must move (block entries, element
 * records, vectors). */
#define VH_MAX_BANDS 16
typedef struct {
  uint32_t bands;
  uint32_t iterations;      /* PD iterations timed */
  uint32_t sweeps;          /* solves per iteration */
  uint32_t launches;        /* kernel launches per iteration */
  double local_us, local_bytes;
  double fwd_us[VH_MAX_BANDS], fwd_bytes[VH_MAX_BANDS];
  double bwd_us[VH_MAX_BANDS], bwd_bytes[VH_MAX_BANDS];
  double ir_us, ir_bytes;   /* refinement updates between sweeps (FP32 mode) */
  double iteration_us;      /* one PD iteration, all launches */
} vh_profile;
/* Times `iterations` PD iterations from the current device state (the state
 * advances like the loop of one step; call after a step). */
vh_status vh_pd_profile(vh_ctx* ctx, double dt, uint32_t iterations, vh_profile* out);
/* The same launches timed the way steps run: one solve of a PD iteration
 * (one sweep from the RHS slots, committing; 2 * bands - 1 launches) and
 * ten local steps are each captured into a HIP graph and replayed `reps`
 * times between two HIP events on the context's stream; the averages per
 * solve and per local step, and their algorithmic bytes (as vh_profile
 * counts them). The state advances with every replay: call after the steps you
 * measure. */
typedef struct {
  uint32_t reps;
  uint32_t solve_launches;
  double solve_us, solve_bytes;
  double local_us, local_bytes;
} vh_graph_profile;
vh_status vh_pd_profile_graph(vh_ctx* ctx, double dt, uint32_t reps, vh_graph_profile* out);
vh_status vh_pd_lumped_masses(vh_ctx* ctx, float* out);

/* ---- IPC contact (augmented-Lagrangian step) --------------------------------
 * Settings of ProjectiveDynamicsSolver::step_with_ipc
 * (crates/vistio-solver/src/config.rs:47-120) and of the CollisionPipeline the
 * benchmark runner builds (crates/vistio-bench/src/runner.rs:60-95, 118-122):
 * ground plane (also the state's ground height), analytic sphere, mesh
 * self-collision with n-ring exclusion. kappa is the barrier stiffness the
 * handler uses (the runner: estimate_initial_kappa, barrier.rs:102-121). */
typedef struct {
  int32_t has_ground;
  double ground_height;
  int32_t has_sphere;
  double sphere_center[3];
  double sphere_radius;
  int32_t self_collision;
  int32_t exclusion_depth;        /* TopologyExclusion depth (runner: 1) */
  double d_hat;                   /* barrier_d_hat, a squared distance */
  double kappa;
  double thickness;               /* broad-phase margin and self-contact thickness (runner: 1.5 d_hat) */
  int32_t al_max_iterations;
  double al_mu_initial, al_mu_growth, al_tolerance;
  int32_t compliant_contact;      /* effective d_hat = d_hat * compliant_d_hat_scale */
  double compliant_d_hat_scale;
  double friction_coefficient, contact_damping;
  int32_t adaptive_contact_damping;
  double contact_damping_max, contact_velocity_threshold;
  uint32_t max_pairs;             /* initial candidate-pair buffer (grows on demand); 0 = 32 per triangle */
  /* 1: the caller's handler ignores set_d_hat (IpcCollisionHandler's default
   * no-op, pd_solver.rs:488, as the handler of bench_tests.rs:181-199):
   * detection keeps d_hat while the CFL clamp uses the compliant-scaled one.
   * 0: detection follows set_d_hat, as the runner's handler (runner.rs:34-36). */
  int32_t detect_keeps_d_hat;
  /* 1: the CCD step also runs edge_edge_ccd (crates/vistio-contact/src/ccd.rs:138-229)
   * on the nine edge pairs (no shared vertex) of every candidate triangle
   * pair; 0: vertex-triangle only, as compute_max_step_size (ccd.rs:238-292). */
  int32_t ccd_edge_edge;
} vh_ipc_params;

/* Diagnostics of the last IPC step or detection. */
typedef struct {
  int32_t al_iterations;
  int32_t al_converged;
  double mu;
  double max_violation;
  int64_t active;                 /* active contacts (self + ground + sphere) */
  int64_t self_contacts;
  int64_t pairs;                  /* candidate triangle pairs */
  double last_alpha;              /* step fraction of the last PD iteration */
} vh_ipc_stats;

/* Enables the contact step on an initialized context (after vh_pd_init). */
vh_status vh_ipc_init(vh_ctx* ctx, const vh_ipc_params* params);
/* ProjectiveDynamicsSolver::step_with_ipc (crates/vistio-solver/src/pd_solver.rs:549-1157)
 * with the pipeline as IpcCollisionHandler, on caller f32 SoA state; the
 * augmented-Lagrangian multipliers (SimulationState.al_lambda_{x,y,z},
 * state.rs:55-63) are read and written through lambda_{x,y,z}. */
vh_status vh_pd_step_ipc(vh_ctx* ctx, double dt, vh_state_f32* state, float* lambda_x, float* lambda_y,
                         float* lambda_z, vh_step_result* out);
/* The same step with the state (and multipliers) resident in HBM. */
vh_status vh_pd_step_ipc_resident(vh_ctx* ctx, double dt, uint32_t nsteps, vh_step_result* out,
                                  vh_ipc_stats* stats);
/* CollisionPipeline::detect_ipc_contacts (crates/vistio-contact/src/collision_pipeline.rs:207-351)
 * at positions pos (3n SoA, caller order; NULL = the resident state) with the
 * given d_hat and kappa. Outputs (caller order), the fields of
 * IpcBarrierForces (crates/vistio-solver/src/pd_solver.rs:1160-1196): grad 3n
 * SoA, normal 3n SoA, in_contact n, hessian_diag n (ground and sphere terms
 * kappa |b''(d^2)| 4 d^2, collision_pipeline.rs:272-276, 301-305); any may be
 * NULL. */
vh_status vh_ipc_detect(vh_ctx* ctx, const double* pos, double d_hat, double kappa, double* grad, double* normal,
                        uint8_t* in_contact, double* hessian_diag, vh_ipc_stats* stats);
/* CollisionPipeline::compute_ccd_step (collision_pipeline.rs:355-400) from p0
 * to p1 (3n SoA each), with the candidate pairs of p0. */
vh_status vh_ipc_ccd_step(vh_ctx* ctx, const double* p0, const double* p1, double* alpha);
/* BroadPhase::query_triangle_pairs (crates/vistio-contact/src/bvh.rs:379-392):
 * every pair of triangle boxes (f32, inflated by margin) that overlap, as
 * (ta < tb) in the caller's triangle numbering, sorted. Writes up to cap
 * pairs and the total count. */
vh_status vh_bvh_pairs(vh_ctx* ctx, const double* pos, double margin, uint32_t* ta, uint32_t* tb, uint64_t cap,
                       uint64_t* count);
/* Intersection-freeness check of a contact scene: the number of triangle
 * pairs that share no vertex and properly cross (an edge of one passes
 * through the other) at pos (3n SoA, caller order; NULL = the resident
 * state). Candidates are the margin-0 box pairs of vh_bvh_pairs. The
 * reference has no such function; it checks what its contact scenes
 * promise (no interpenetration) on the GPU's own state. */
vh_status vh_ipc_crossings(vh_ctx* ctx, const double* pos, uint64_t* count);
/* ---- NarrowPhase (crates/vistio-contact/src/narrow.rs:21-35) -----------------
 * The pre-IPC narrow phase of the reference's collision pipeline
 * (CollisionPipeline::step with VertexTriangleTest, used by
 * BenchmarkRunner::run_with_collision, runner.rs:66). */

/* ContactPair (crates/vistio-contact/src/contact.rs:22-44) as
 * VertexTriangleTest fills it: indices = [vertex, tri_v0, tri_v1, tri_v2],
 * distance = -|plane distance|, normal from the triangle toward the vertex,
 * barycentric (u, v, w) of the vertex's projection; contact_type 0
 * (ContactType::VertexTriangle) and is_self 1 for every such contact. */
typedef struct {
  uint32_t indices[4];
  float distance;
  float normal[3];
  float barycentric[3];
  int32_t contact_type;
  int32_t is_self;
} vh_contact_pair;

/* VertexTriangleTest::detect (vertex_triangle.rs:25-81, with
 * point_triangle_proximity :88-152), in f32 as the reference: for each
 * candidate (cand_a = vertex vi, cand_b = vertex vj), in order, every
 * triangle of tri (t triangles, 3t indices, caller order), ascending, that
 * contains vj and not vi, whose plane lies within thickness of vi and whose
 * projection of vi passes the barycentric test (each coordinate >= -0.01).
 * px/py/pz hold n f32 positions (SimulationState's SoA). Writes up to cap
 * contacts and the total count. An index past n returns VH_E_INVALID_ARG
 * (candidates) or VH_E_INVALID_MESH (triangles). Needs only vh_create. */
vh_status vh_narrow_vertex_triangle(vh_ctx* ctx, uint32_t n, const float* px, const float* py, const float* pz,
                                    uint32_t t, const uint32_t* tri, uint64_t ncand, const uint32_t* cand_a,
                                    const uint32_t* cand_b, float thickness, vh_contact_pair* out, uint64_t cap,
                                    uint64_t* count);

/* Resident multipliers (3n SoA, caller order). */
vh_status vh_state_lambda_set_f64(vh_ctx* ctx, const double* lambda);
vh_status vh_state_lambda_get_f64(vh_ctx* ctx, double* lambda);

/* Mesh helpers (crates/vistio-mesh): quad_grid (generators.rs:28-100) and the
 * interior edges of Topology::build (topology.rs:58-120, :158-168) in a
 * canonical (v0, v1)-sorted order. */
vh_status vh_mesh_quad_grid(uint32_t cols, uint32_t rows, float width, float height,
                            float* px, float* py, float* pz, uint32_t* indices);
/* Writes up to cap interior edges (6 u32 each: v0, v1, wing_a, wing_b, tri_a,
 * tri_b) and their count. Call with cap = 0 to query the count. */
vh_status vh_mesh_interior_edges(uint32_t n, uint32_t t, const uint32_t* indices, uint32_t* out,
                                 uint32_t cap, uint32_t* count);

/* ---- GpuBackend (crates/vistio-gpu/src/backend.rs:19-46) --------------------
 * The trait's primitives over the caller's f32 buffers (ComputeBuffer holds
 * host data, buffers.rs:10-16): each call copies the operands to the device,
 * runs one kernel and copies the result back (PCIe-bound; the solver itself
 * keeps its state in HBM). Needs only vh_create. Length mismatches return
 * VH_E_INVALID_ARG with the reference's messages ("AXPY buffer length
 * mismatch", "Dot product buffer length mismatch"), which the binding maps
 * to VistioError::Gpu as CpuFallback does (backend.rs:88-116).
 *   axpy  out[i] = a[i] + scale * b[i], rounded as separate f32 multiply and add
 *   dot   sum of a[i] * b[i] in f64 (exact products; the summation order is a
 *         parallel tree, so the result can differ from the sequential sum by a
 *         few f64 ulps of the partial sums)
 *   fill  buf[i] = value */
vh_status vh_backend_axpy(vh_ctx* ctx, const float* a, uint64_t na, const float* b, uint64_t nb, float scale,
                          float* out, uint64_t nout);
vh_status vh_backend_dot(vh_ctx* ctx, const float* a, uint64_t na, const float* b, uint64_t nb, double* result);
vh_status vh_backend_fill(vh_ctx* ctx, float* buf, uint64_t n, float value);

/* vh_backend_dot without any device allocation. It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory (tight
 * serving loops, harnesses that forbid hipMalloc). Same arguments, result and
 * errors as vh_backend_dot, and the same kernels, so the same value bit for
 * bit. Its partial sums and operand copies live in one fixed device array
 * per GPU (2^21 floats per operand plus 8.2 KB, about 16 MB, held until the
 * process ends); more than 2^21 elements returns VH_E_INVALID_ARG. Every call
 * on one GPU uses that GPU's array, whichever context makes it: calls that
 * overlap on one GPU (other threads, other contexts) race. */
vh_status vh_backend_dot_noalloc(vh_ctx* ctx, const float* a, uint64_t na, const float* b, uint64_t nb,
                                 double* result);

#ifdef __cplusplus
}
#endif

#endif /* VISTIO_HIP_H */
