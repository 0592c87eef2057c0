//! Raw declarations of the C-ABI of libvistio_hip.so (include/vistio_hip.h),
//! one item per C declaration, in the header's order. The struct layouts are
//! `#[repr(C)]` copies of the C typedefs: same fields, same order, same types.

use std::os::raw::c_char;

/// `vh_status`.
pub type VhStatus = i32;
pub const VH_OK: VhStatus = 0;
pub const VH_E_INVALID_CONFIG: VhStatus = 1;
pub const VH_E_NOT_INITIALIZED: VhStatus = 2;
pub const VH_E_NOT_SPD: VhStatus = 3;
pub const VH_E_UNSUPPORTED: VhStatus = 4;
pub const VH_E_INVALID_MESH: VhStatus = 5;
pub const VH_E_GPU: VhStatus = 6;
pub const VH_E_INTERNAL: VhStatus = 7;
pub const VH_E_INVALID_ARG: VhStatus = 8;
pub const VH_E_NONFINITE: VhStatus = 9;

pub const VH_MODEL_ARAP: i32 = 0;
pub const VH_MODEL_COROTATIONAL: i32 = 1;
pub const VH_MODEL_ISOTROPIC_LINEAR: i32 = 2;
pub const VH_MODEL_ORTHOTROPIC_LINEAR: i32 = 3;
pub const VH_MODEL_ANISOTROPIC_COROTATIONAL: i32 = 4;

pub const VH_BENDING_NONE: i32 = 0;
pub const VH_BENDING_DIHEDRAL: i32 = 1;
pub const VH_BENDING_DISCRETE_SHELLS: i32 = 2;

pub const VH_PREC_F64: i32 = 0;
pub const VH_PREC_F32_IR: i32 = 1;

pub const VH_MAX_BANDS: usize = 16;

/// `vh_ctx`: opaque; one per solver, owns the device memory and one HIP stream.
#[repr(C)]
pub struct VhCtx {
    _opaque: [u8; 0],
}

/// `vh_pd_params`.
#[repr(C)]
#[derive(Debug, Clone, Copy, Default)]
pub struct VhPdParams {
    pub model: i32,
    pub bending: i32,
    pub k_membrane: f64,
    pub k_bending: f64,
    pub density: f64,
    pub dt_matrix: f64,
    pub gravity: [f64; 3],
    pub damping: f64,
    pub rayleigh: f64,
    pub max_iterations: i32,
    pub tolerance: f64,
    pub has_ground: i32,
    pub ground_height: f64,
    pub precision: i32,
    pub ir_sweeps: i32,
    pub ir_final_sweeps: i32,
    pub leaf_size: i32,
    pub max_band_levels: i32,
    pub band_launch_us: f64,
    pub band_gbps: f64,
    pub task_entries: i32,
    pub warp_stiffness: f64,
    pub weft_stiffness: f64,
    pub bend_warp: f64,
    pub bend_weft: f64,
}

/// `vh_step_result`.
#[repr(C)]
#[derive(Debug, Clone, Copy, Default)]
pub struct VhStepResult {
    pub iterations: u32,
    pub final_residual: f64,
    pub converged: i32,
    pub wall_time: f64,
}

/// `vh_state_f32`: the SoA arrays of a SimulationState, each of length n.
#[repr(C)]
#[derive(Debug, Clone, Copy)]
pub struct VhStateF32 {
    pub pos_x: *mut f32,
    pub pos_y: *mut f32,
    pub pos_z: *mut f32,
    pub vel_x: *mut f32,
    pub vel_y: *mut f32,
    pub vel_z: *mut f32,
    pub prev_x: *mut f32,
    pub prev_y: *mut f32,
    pub prev_z: *mut f32,
    pub pred_x: *mut f32,
    pub pred_y: *mut f32,
    pub pred_z: *mut f32,
    pub inv_mass: *const f32,
}

/// `vh_pd_info`.
#[repr(C)]
#[derive(Debug, Clone, Copy, Default)]
pub struct VhPdInfo {
    pub n: u32,
    pub t: u32,
    pub h: u32,
    pub supernodes: u32,
    pub levels: u32,
    pub nnz_a: i64,
    pub nnz_factor: i64,
    pub max_front: i32,
    pub setup_seconds: f64,
    pub factor_seconds: f64,
}

/// `vh_material_arrays`: per-element materials; a null pointer selects the
/// uniform value of `VhPdParams`.
#[repr(C)]
#[derive(Debug, Clone, Copy)]
pub struct VhMaterialArrays {
    pub k_membrane: *const f64,
    pub k_bending: *const f64,
    pub density: *const f64,
}

/// `vh_graph_profile`.
#[repr(C)]
#[derive(Debug, Clone, Copy, Default)]
pub struct VhGraphProfile {
    pub reps: u32,
    pub solve_launches: u32,
    pub solve_us: f64,
    pub solve_bytes: f64,
    pub local_us: f64,
    pub local_bytes: f64,
}

/// `vh_profile`.
#[repr(C)]
#[derive(Debug, Clone, Copy, Default)]
pub struct VhProfile {
    pub bands: u32,
    pub iterations: u32,
    pub sweeps: u32,
    pub launches: u32,
    pub local_us: f64,
    pub local_bytes: f64,
    pub fwd_us: [f64; VH_MAX_BANDS],
    pub fwd_bytes: [f64; VH_MAX_BANDS],
    pub bwd_us: [f64; VH_MAX_BANDS],
    pub bwd_bytes: [f64; VH_MAX_BANDS],
    pub ir_us: f64,
    pub ir_bytes: f64,
    pub iteration_us: f64,
}

/// `vh_ipc_params`.
#[repr(C)]
#[derive(Debug, Clone, Copy, Default)]
pub struct VhIpcParams {
    pub has_ground: i32,
    pub ground_height: f64,
    pub has_sphere: i32,
    pub sphere_center: [f64; 3],
    pub sphere_radius: f64,
    pub self_collision: i32,
    pub exclusion_depth: i32,
    pub d_hat: f64,
    pub kappa: f64,
    pub thickness: f64,
    pub al_max_iterations: i32,
    pub al_mu_initial: f64,
    pub al_mu_growth: f64,
    pub al_tolerance: f64,
    pub compliant_contact: i32,
    pub compliant_d_hat_scale: f64,
    pub friction_coefficient: f64,
    pub contact_damping: f64,
    pub adaptive_contact_damping: i32,
    pub contact_damping_max: f64,
    pub contact_velocity_threshold: f64,
    pub max_pairs: u32,
    pub detect_keeps_d_hat: i32,
    pub ccd_edge_edge: i32,
}

/// `vh_ipc_stats`.
#[repr(C)]
#[derive(Debug, Clone, Copy, Default)]
pub struct VhIpcStats {
    pub al_iterations: i32,
    pub al_converged: i32,
    pub mu: f64,
    pub max_violation: f64,
    pub active: i64,
    pub self_contacts: i64,
    pub pairs: i64,
    pub last_alpha: f64,
}

/// `vh_contact_pair` (ContactPair as VertexTriangleTest fills it).
#[repr(C)]
#[derive(Debug, Clone, Copy, Default)]
pub struct VhContactPair {
    pub indices: [u32; 4],
    pub distance: f32,
    pub normal: [f32; 3],
    pub barycentric: [f32; 3],
    pub contact_type: i32,
    pub is_self: i32,
}

#[link(name = "vistio_hip")]
extern "C" {
    pub fn vh_create(device: i32, out: *mut *mut VhCtx) -> VhStatus;
    pub fn vh_destroy(ctx: *mut VhCtx) -> VhStatus;
    pub fn vh_last_error(ctx: *const VhCtx) -> *const c_char;
    pub fn vh_pd_init(
        ctx: *mut VhCtx,
        n: u32,
        px: *const f32,
        py: *const f32,
        pz: *const f32,
        t: u32,
        tri: *const u32,
        h: u32,
        hinge: *const u32,
        pinned: *const u8,
        params: *const VhPdParams,
        lumped_mass_out: *mut f32,
    ) -> VhStatus;
    pub fn vh_pd_init_ex(
        ctx: *mut VhCtx,
        n: u32,
        px: *const f32,
        py: *const f32,
        pz: *const f32,
        t: u32,
        tri: *const u32,
        h: u32,
        hinge: *const u32,
        pinned: *const u8,
        params: *const VhPdParams,
        mat: *const VhMaterialArrays,
        lumped_mass_out: *mut f32,
    ) -> VhStatus;
    pub fn vh_pd_step(ctx: *mut VhCtx, dt: f64, state: *mut VhStateF32, out: *mut VhStepResult) -> VhStatus;
    pub fn vh_pd_step_resident(ctx: *mut VhCtx, dt: f64, nsteps: u32, out: *mut VhStepResult) -> VhStatus;
    pub fn vh_state_set_f64(ctx: *mut VhCtx, which: i32, x: *const f64, y: *const f64, z: *const f64) -> VhStatus;
    pub fn vh_state_get_f64(ctx: *mut VhCtx, which: i32, x: *mut f64, y: *mut f64, z: *mut f64) -> VhStatus;
    pub fn vh_state_set_inv_mass(ctx: *mut VhCtx, inv_mass: *const f32) -> VhStatus;
    pub fn vh_pd_set_ground(ctx: *mut VhCtx, has_ground: i32, ground_height: f64) -> VhStatus;
    pub fn vh_pd_residual(ctx: *mut VhCtx, dt: f64, out: *mut f64) -> VhStatus;
    pub fn vh_pd_solve(ctx: *mut VhCtx, b: *const f64, x: *mut f64, sweeps: i32) -> VhStatus;
    pub fn vh_pd_info_get(ctx: *mut VhCtx, out: *mut VhPdInfo) -> VhStatus;
    pub fn vh_pd_profile(ctx: *mut VhCtx, dt: f64, iterations: u32, out: *mut VhProfile) -> VhStatus;
    pub fn vh_pd_profile_graph(ctx: *mut VhCtx, dt: f64, reps: u32, out: *mut VhGraphProfile) -> VhStatus;
    pub fn vh_pd_lumped_masses(ctx: *mut VhCtx, out: *mut f32) -> VhStatus;
    pub fn vh_ipc_init(ctx: *mut VhCtx, params: *const VhIpcParams) -> VhStatus;
    pub fn vh_pd_step_ipc(
        ctx: *mut VhCtx,
        dt: f64,
        state: *mut VhStateF32,
        lambda_x: *mut f32,
        lambda_y: *mut f32,
        lambda_z: *mut f32,
        out: *mut VhStepResult,
    ) -> VhStatus;
    pub fn vh_pd_step_ipc_resident(
        ctx: *mut VhCtx,
        dt: f64,
        nsteps: u32,
        out: *mut VhStepResult,
        stats: *mut VhIpcStats,
    ) -> VhStatus;
    pub fn vh_ipc_detect(
        ctx: *mut VhCtx,
        pos: *const f64,
        d_hat: f64,
        kappa: f64,
        grad: *mut f64,
        normal: *mut f64,
        in_contact: *mut u8,
        hessian_diag: *mut f64,
        stats: *mut VhIpcStats,
    ) -> VhStatus;
    pub fn vh_ipc_ccd_step(ctx: *mut VhCtx, p0: *const f64, p1: *const f64, alpha: *mut f64) -> VhStatus;
    pub fn vh_bvh_pairs(
        ctx: *mut VhCtx,
        pos: *const f64,
        margin: f64,
        ta: *mut u32,
        tb: *mut u32,
        cap: u64,
        count: *mut u64,
    ) -> VhStatus;
    pub fn vh_ipc_crossings(ctx: *mut VhCtx, pos: *const f64, count: *mut u64) -> VhStatus;
    pub fn vh_narrow_vertex_triangle(
        ctx: *mut VhCtx,
        n: u32,
        px: *const f32,
        py: *const f32,
        pz: *const f32,
        t: u32,
        tri: *const u32,
        ncand: u64,
        cand_a: *const u32,
        cand_b: *const u32,
        thickness: f32,
        out: *mut VhContactPair,
        cap: u64,
        count: *mut u64,
    ) -> VhStatus;
    pub fn vh_state_lambda_set_f64(ctx: *mut VhCtx, lambda: *const f64) -> VhStatus;
    pub fn vh_state_lambda_get_f64(ctx: *mut VhCtx, lambda: *mut f64) -> VhStatus;
    pub fn vh_mesh_quad_grid(
        cols: u32,
        rows: u32,
        width: f32,
        height: f32,
        px: *mut f32,
        py: *mut f32,
        pz: *mut f32,
        indices: *mut u32,
    ) -> VhStatus;
    pub fn vh_mesh_interior_edges(
        n: u32,
        t: u32,
        indices: *const u32,
        out: *mut u32,
        cap: u32,
        count: *mut u32,
    ) -> VhStatus;
    pub fn vh_backend_axpy(
        ctx: *mut VhCtx,
        a: *const f32,
        na: u64,
        b: *const f32,
        nb: u64,
        scale: f32,
        out: *mut f32,
        nout: u64,
    ) -> VhStatus;
    pub fn vh_backend_dot(ctx: *mut VhCtx, a: *const f32, na: u64, b: *const f32, nb: u64, result: *mut f64)
        -> VhStatus;
    pub fn vh_backend_fill(ctx: *mut VhCtx, buf: *mut f32, n: u64, value: f32) -> VhStatus;
    pub fn vh_pd_blocks_get(ctx: *mut VhCtx, which: i32, out: *mut f64, cap: u64, count: *mut u64) -> VhStatus;
    pub fn vh_backend_dot_noalloc(ctx: *mut VhCtx, a: *const f32, na: u64, b: *const f32, nb: u64, result: *mut f64)
        -> VhStatus;
}
