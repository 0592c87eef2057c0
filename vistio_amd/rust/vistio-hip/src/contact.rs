//! The contact path on the GPU: `step_with_ipc` (pd_solver.rs:549-1157) as one
//! device step, an `IpcCollisionHandler` over the device detection and CCD,
//! and a `BroadPhase` over the device LBVH.

use vistio_contact::{
    broad::{BroadPhase, CandidatePair, CandidateTrianglePair},
    CollisionPipeline,
};
use vistio_mesh::TriangleMesh;
use vistio_solver::{
    config::SolverConfig,
    pd_solver::{IpcBarrierForces, IpcCollisionHandler},
    state::SimulationState,
    strategy::StepResult,
};
use vistio_types::{VistioError, VistioResult};

use crate::ctx::Ctx;
use crate::ffi::*;
use crate::solver::{to_step_result, HipPdSolver};

/// The colliders of a contact scene as the device step takes them: what
/// BenchmarkRunner wires into its CollisionPipeline (runner.rs:56-95).
#[derive(Debug, Clone, Copy, Default)]
pub struct IpcScene {
    pub ground: Option<f32>,
    pub sphere: Option<([f32; 3], f32)>,
    pub self_collision: bool,
    /// TopologyExclusion depth of the self-collision system (the runner's
    /// with_self_collision(&topology, 1) uses 1; the pipeline keeps it private).
    pub exclusion_depth: i32,
    /// Broad-phase margin and self-contact thickness (runner: 1.5 * barrier_d_hat).
    pub thickness: f32,
    /// The handler's `set_d_hat` is the trait's default no-op (pd_solver.rs:488),
    /// as in bench_tests.rs:181-199: detection keeps `barrier_d_hat` while the
    /// CFL clamp uses the compliant-scaled one. The runner's handler follows
    /// set_d_hat (runner.rs:34-36): false.
    pub handler_ignores_set_d_hat: bool,
    /// The CCD step also runs `edge_edge_ccd` (ccd.rs:138-229) on the nine
    /// edge pairs of every candidate triangle pair. false: vertex-triangle
    /// only, as `compute_max_step_size` (ccd.rs:238-292).
    pub ccd_edge_edge: bool,
}

/// The handler argument of [`HipPdSolver::step_with_ipc`]: the scene that
/// `init_ipc` registered on the device. Detection and CCD run inside the
/// device step, which cannot call back into a host `IpcCollisionHandler`;
/// taking this marker instead of any handler makes a caller's own handler a
/// compile error rather than an argument that would be silently ignored.
#[derive(Debug, Clone, Copy, Default)]
pub struct DeviceScene;

impl IpcScene {
    /// Ground, sphere, self-collision and thickness of a pipeline. Cylinder
    /// and box colliders have no device implementation: a pipeline with one
    /// is refused.
    pub fn from_pipeline(p: &CollisionPipeline, exclusion_depth: i32) -> VistioResult<Self> {
        if p.cylinder.is_some() || p.box_collider.is_some() {
            return Err(VistioError::InvalidConfig(
                "cylinder and box colliders are not supported by the GPU contact step".into(),
            ));
        }
        Ok(Self {
            ground: p.ground.as_ref().map(|g| g.height),
            sphere: p.sphere.as_ref().map(|s| ([s.center.x, s.center.y, s.center.z], s.radius)),
            self_collision: p.self_collision.is_some(),
            exclusion_depth,
            thickness: p.thickness,
            handler_ignores_set_d_hat: false,
            ccd_edge_edge: false,
        })
    }
}

fn widen(x: &[f32], y: &[f32], z: &[f32]) -> Vec<f64> {
    x.iter().chain(y).chain(z).map(|&v| v as f64).collect()
}

impl HipPdSolver {
    /// Enables `step_with_ipc`: the scene, SolverConfig's IPC fields
    /// (config.rs:47-128) and the handler's kappa (the runner's
    /// estimate_initial_kappa, barrier.rs:97-115).
    pub fn init_ipc(&mut self, scene: &IpcScene, config: &SolverConfig, kappa: f32) -> VistioResult<()> {
        self.require_init()?;
        let (has_sphere, center, radius) = match scene.sphere {
            Some((c, r)) => (1, c.map(|v| v as f64), r as f64),
            None => (0, [0.0; 3], 0.0),
        };
        let q = VhIpcParams {
            has_ground: scene.ground.is_some() as i32,
            ground_height: scene.ground.unwrap_or(0.0) as f64,
            has_sphere,
            sphere_center: center,
            sphere_radius: radius,
            self_collision: scene.self_collision as i32,
            exclusion_depth: scene.exclusion_depth,
            d_hat: config.barrier_d_hat as f64,
            kappa: kappa as f64,
            thickness: scene.thickness as f64,
            al_max_iterations: config.al_max_iterations as i32,
            al_mu_initial: config.al_mu_initial as f64,
            al_mu_growth: config.al_mu_growth as f64,
            al_tolerance: config.al_tolerance as f64,
            compliant_contact: config.compliant_contact as i32,
            compliant_d_hat_scale: config.compliant_d_hat_scale as f64,
            friction_coefficient: config.friction_coefficient as f64,
            contact_damping: config.contact_damping as f64,
            adaptive_contact_damping: config.adaptive_contact_damping as i32,
            contact_damping_max: config.contact_damping_max as f64,
            contact_velocity_threshold: config.contact_velocity_threshold as f64,
            max_pairs: 0,
            detect_keeps_d_hat: scene.handler_ignores_set_d_hat as i32,
            ccd_edge_edge: scene.ccd_edge_edge as i32,
        };
        self.ipc_ready = false;
        self.ctx.check(unsafe { vh_ipc_init(self.ctx.0, &q) })?;
        self.ipc_ready = true;
        self.ipc_ground = scene.ground;
        Ok(())
    }

    fn require_ipc(&self) -> VistioResult<()> {
        self.require_init()?;
        if self.ipc_ready {
            Ok(())
        } else {
            Err(VistioError::InvalidConfig("contact not initialized: call init_ipc".into()))
        }
    }

    /// step_with_ipc (pd_solver.rs:549-1157). Detection, CCD and the AL loop
    /// all run on the device with the scene given to `init_ipc`, so the
    /// handler is the [`DeviceScene`] marker (a caller's own
    /// IpcCollisionHandler does not type-check). The predict clamp and the
    /// ground-velocity fix use the scene's ground, so a `state.ground_height`
    /// that differs from it is refused (the reference reads the state's,
    /// state.rs:160-212). state.al_lambda_{x,y,z} carry the multipliers in
    /// and out.
    pub fn step_with_ipc(
        &mut self,
        state: &mut SimulationState,
        dt: f32,
        _scene: &mut DeviceScene,
    ) -> VistioResult<StepResult> {
        self.require_ipc()?;
        if state.ground_height != self.ipc_ground {
            return Err(VistioError::InvalidConfig(format!(
                "state.ground_height {:?} differs from the contact scene's ground {:?} given to init_ipc",
                state.ground_height, self.ipc_ground
            )));
        }
        let n = self.n;
        if state.al_lambda_x.len() != n || state.al_lambda_y.len() != n || state.al_lambda_z.len() != n {
            return Err(VistioError::InvalidMesh(format!("al_lambda arrays must have {n} entries")));
        }
        let mut st = self.state_ptrs(state)?;
        let mut r = VhStepResult::default();
        let rc = unsafe {
            vh_pd_step_ipc(
                self.ctx.0,
                dt as f64,
                &mut st,
                state.al_lambda_x.as_mut_ptr(),
                state.al_lambda_y.as_mut_ptr(),
                state.al_lambda_z.as_mut_ptr(),
                &mut r,
            )
        };
        self.ctx.check(rc)?;
        Ok(to_step_result(&r))
    }

    /// `nsteps` contact steps on the device-resident state, with the
    /// diagnostics of the last one.
    pub fn step_ipc_resident(&mut self, dt: f32, nsteps: u32) -> VistioResult<(StepResult, VhIpcStats)> {
        self.require_ipc()?;
        let mut r = VhStepResult::default();
        let mut s = VhIpcStats::default();
        self.ctx.check(unsafe { vh_pd_step_ipc_resident(self.ctx.0, dt as f64, nsteps, &mut r, &mut s) })?;
        Ok((to_step_result(&r), s))
    }

    /// Number of triangle pairs without a shared vertex whose surfaces
    /// cross, at the given positions (SoA) or, with None, the device state.
    pub fn count_crossings(&mut self, pos: Option<(&[f32], &[f32], &[f32])>) -> VistioResult<u64> {
        self.require_ipc()?;
        let p = pos.map(|(x, y, z)| widen(x, y, z));
        if let Some(ref v) = p {
            if v.len() != 3 * self.n {
                return Err(VistioError::InvalidMesh(format!("positions must have {} entries", self.n)));
            }
        }
        let mut count = 0u64;
        let ptr = p.as_ref().map_or(std::ptr::null(), |v| v.as_ptr());
        self.ctx.check(unsafe { vh_ipc_crossings(self.ctx.0, ptr, &mut count) })?;
        Ok(count)
    }
}

/// IpcCollisionHandler (pd_solver.rs:483-489) over the device detection and
/// CCD, for callers that drive the trait themselves. The trait's methods
/// cannot return errors: a device failure panics with the library's message.
pub struct HipIpcHandler<'a> {
    solver: &'a mut HipPdSolver,
    d_hat: f32,
    kappa: f32,
}

impl<'a> HipIpcHandler<'a> {
    /// `solver` has had `init_ipc`; d_hat and kappa as the runner's handler
    /// holds them (runner.rs:18-37).
    pub fn new(solver: &'a mut HipPdSolver, d_hat: f32, kappa: f32) -> VistioResult<Self> {
        solver.require_ipc()?;
        Ok(Self { solver, d_hat, kappa })
    }
}

impl IpcCollisionHandler for HipIpcHandler<'_> {
    /// CollisionPipeline::detect_ipc_contacts (collision_pipeline.rs:207-351).
    fn detect_contacts(&mut self, px: &[f32], py: &[f32], pz: &[f32]) -> IpcBarrierForces {
        let n = self.solver.n;
        assert!(px.len() == n && py.len() == n && pz.len() == n, "positions must have {n} entries");
        let pos = widen(px, py, pz);
        let (mut grad, mut nrm, mut hd) = (vec![0.0f64; 3 * n], vec![0.0f64; 3 * n], vec![0.0f64; n]);
        let mut inc = vec![0u8; n];
        let mut st = VhIpcStats::default();
        let rc = unsafe {
            vh_ipc_detect(
                self.solver.ctx.0,
                pos.as_ptr(),
                self.d_hat as f64,
                self.kappa as f64,
                grad.as_mut_ptr(),
                nrm.as_mut_ptr(),
                inc.as_mut_ptr(),
                hd.as_mut_ptr(),
                &mut st,
            )
        };
        if let Err(e) = self.solver.ctx.check(rc) {
            panic!("vh_ipc_detect: {e}");
        }
        let f = |v: &[f64]| v.iter().map(|&x| x as f32).collect::<Vec<f32>>();
        IpcBarrierForces {
            grad_x: f(&grad[..n]),
            grad_y: f(&grad[n..2 * n]),
            grad_z: f(&grad[2 * n..]),
            max_violation: st.max_violation as f32,
            active_contacts: st.active as usize,
            contact_nx: f(&nrm[..n]),
            contact_ny: f(&nrm[n..2 * n]),
            contact_nz: f(&nrm[2 * n..]),
            in_contact: inc.iter().map(|&b| b != 0).collect(),
            hessian_diag: f(&hd),
        }
    }

    /// CollisionPipeline::compute_ccd_step (collision_pipeline.rs:355-400).
    fn compute_ccd_step(
        &mut self,
        prev_x: &[f32],
        prev_y: &[f32],
        prev_z: &[f32],
        new_x: &[f32],
        new_y: &[f32],
        new_z: &[f32],
    ) -> f32 {
        let (p0, p1) = (widen(prev_x, prev_y, prev_z), widen(new_x, new_y, new_z));
        let n = self.solver.n;
        assert!(p0.len() == 3 * n && p1.len() == 3 * n, "positions must have {n} entries");
        let mut alpha = 1.0f64;
        let rc = unsafe { vh_ipc_ccd_step(self.solver.ctx.0, p0.as_ptr(), p1.as_ptr(), &mut alpha) };
        if let Err(e) = self.solver.ctx.check(rc) {
            panic!("vh_ipc_ccd_step: {e}");
        }
        alpha as f32
    }

    fn set_d_hat(&mut self, d_hat: f32) {
        self.d_hat = d_hat;
    }
}

/// BroadPhase (broad.rs:39-56) over the device LBVH. `update` runs the
/// query and keeps its result, as BvhBroadPhase::update does
/// (bvh.rs:305-373); the queries return copies. The pair set is every
/// overlapping pair of triangle boxes, a superset of the reference tree's
/// `query_self_pairs`, which drops pairs through its `ta < tb` subtree filter
/// (bvh.rs:207).
pub struct HipBroadPhase {
    ctx: Ctx,
    n: usize,
    indices: Vec<u32>,
    tri_pairs: Vec<CandidateTrianglePair>,
    pairs: Vec<CandidatePair>,
}

// The device context is touched only by `update` (&mut self); the &self
// queries read host vectors.
unsafe impl Sync for HipBroadPhase {}

impl HipBroadPhase {
    /// A broad phase for `mesh` on HIP device `device`. The library needs a
    /// set-up mesh context for its BVH; this one carries no physics (ARAP,
    /// nothing pinned) and costs one factorization of the mesh.
    pub fn new(device: i32, mesh: &TriangleMesh) -> VistioResult<Self> {
        let ctx = Ctx::new(device)?;
        let n = mesh.vertex_count();
        let p = VhPdParams {
            model: VH_MODEL_ARAP,
            bending: VH_BENDING_NONE,
            k_membrane: 1.0,
            density: 0.2,
            dt_matrix: 1.0 / 60.0,
            max_iterations: 1,
            precision: VH_PREC_F64,
            ir_sweeps: 1,
            ir_final_sweeps: 1,
            ..Default::default()
        };
        let pin = vec![0u8; n];
        let rc = unsafe {
            vh_pd_init(
                ctx.0,
                n as u32,
                mesh.pos_x.as_ptr(),
                mesh.pos_y.as_ptr(),
                mesh.pos_z.as_ptr(),
                mesh.triangle_count() as u32,
                mesh.indices.as_ptr(),
                0,
                std::ptr::null(),
                pin.as_ptr(),
                &p,
                std::ptr::null_mut(),
            )
        };
        ctx.check(rc)?;
        let q = VhIpcParams { self_collision: 1, exclusion_depth: 1, d_hat: 1e-3, al_max_iterations: 1, ..Default::default() };
        ctx.check(unsafe { vh_ipc_init(ctx.0, &q) })?;
        Ok(Self { ctx, n, indices: mesh.indices.clone(), tri_pairs: Vec::new(), pairs: Vec::new() })
    }
}

impl BroadPhase for HipBroadPhase {
    fn update(&mut self, pos_x: &[f32], pos_y: &[f32], pos_z: &[f32], thickness: f32) -> VistioResult<()> {
        let n = self.n;
        if pos_x.len() != n || pos_y.len() != n || pos_z.len() != n {
            return Err(VistioError::InvalidMesh(format!("positions must have {n} entries")));
        }
        let pos = widen(pos_x, pos_y, pos_z);
        let mut count = 0u64;
        let (mut ta, mut tb) = (Vec::<u32>::new(), Vec::<u32>::new());
        // The count first, then the pairs into buffers of that size.
        self.ctx.check(unsafe {
            vh_bvh_pairs(self.ctx.0, pos.as_ptr(), thickness as f64, std::ptr::null_mut(), std::ptr::null_mut(), 0, &mut count)
        })?;
        ta.resize(count as usize, 0);
        tb.resize(count as usize, 0);
        self.ctx.check(unsafe {
            vh_bvh_pairs(self.ctx.0, pos.as_ptr(), thickness as f64, ta.as_mut_ptr(), tb.as_mut_ptr(), count, &mut count)
        })?;
        self.tri_pairs = ta.iter().zip(&tb).map(|(&a, &b)| CandidateTrianglePair { ta: a, tb: b, is_self: true }).collect();
        // Vertex pairs of the triangle pairs that share no vertex, as
        // BvhBroadPhase::update emits them (bvh.rs:331-366).
        self.pairs.clear();
        for tp in &self.tri_pairs {
            let (a, b) = (3 * tp.ta as usize, 3 * tp.tb as usize);
            let va = [self.indices[a], self.indices[a + 1], self.indices[a + 2]];
            let vb = [self.indices[b], self.indices[b + 1], self.indices[b + 2]];
            if va.iter().any(|v| vb.contains(v)) {
                continue;
            }
            for &x in &va {
                for &y in &vb {
                    self.pairs.push(CandidatePair { a: x, b: y, is_self: true });
                }
            }
        }
        Ok(())
    }

    fn query_pairs(&self) -> Vec<CandidatePair> {
        self.pairs.clone()
    }

    fn query_triangle_pairs(&self) -> Vec<CandidateTrianglePair> {
        self.tri_pairs.clone()
    }

    fn name(&self) -> &str {
        "hip_lbvh"
    }
}
