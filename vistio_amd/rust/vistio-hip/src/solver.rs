//! `HipPdSolver`: ProjectiveDynamicsSolver (crates/vistio-solver/src/pd_solver.rs)
//! with its step on an MI355X. The set-up (rest data, lumped masses, system
//! matrix, nested-dissection factorization) and every PD iteration run in
//! libvistio_hip.so; this file maps the reference's init paths and
//! SolverStrategy onto the C-ABI.

use vistio_material::{ConstitutiveModel, FabricProperties};
use vistio_mesh::{topology::Topology, TriangleMesh};
use vistio_solver::{
    config::SolverConfig,
    state::SimulationState,
    strategy::{SolverStrategy, StepResult},
};
use vistio_types::{VistioError, VistioResult};

use crate::ctx::Ctx;
use crate::ffi::*;

/// Knobs of the GPU step that have no SolverConfig counterpart.
#[derive(Debug, Clone, Copy)]
pub struct HipSettings {
    /// VH_PREC_F64 (FP64 local step and solves) or VH_PREC_F32_IR (FP32
    /// local step and supernodal solves, FP64 residual refinement).
    pub precision: i32,
    /// Refinement sweeps per PD iteration and on the last one (1..3).
    pub ir_sweeps: i32,
    pub ir_final_sweeps: i32,
    /// Solve schedule; 0 keeps the library's defaults (leaf 16, at most 3
    /// dissection levels per band, 6 us and 4000 GB/s in the band cost model,
    /// 1024 block entries per task).
    pub leaf_size: i32,
    pub max_band_levels: i32,
    pub band_launch_us: f64,
    pub band_gbps: f64,
    pub task_entries: i32,
}

impl Default for HipSettings {
    /// FP32-refined: 1 sweep per PD iteration, the last included; positions
    /// stay within 1e-8 m of the FP64 step after 120 steps of the 256 x 256
    /// sheet.
    fn default() -> Self {
        Self {
            precision: VH_PREC_F32_IR,
            ir_sweeps: 1,
            ir_final_sweeps: 1,
            leaf_size: 0,
            max_band_levels: 0,
            band_launch_us: 0.0,
            band_gbps: 0.0,
            task_entries: 0,
        }
    }
}

/// ConstitutiveModel::name() -> the library's closed model enum.
pub(crate) fn model_id(name: &str) -> VistioResult<i32> {
    match name {
        "co_rotational" => Ok(VH_MODEL_COROTATIONAL),
        "isotropic_linear" => Ok(VH_MODEL_ISOTROPIC_LINEAR),
        "orthotropic_linear" => Ok(VH_MODEL_ORTHOTROPIC_LINEAR),
        "anisotropic_corotational" => Ok(VH_MODEL_ANISOTROPIC_COROTATIONAL),
        other => Err(VistioError::InvalidConfig(format!("unsupported constitutive model {other:?}"))),
    }
}

/// The material values one init path hands to vh_pd_init.
struct Material {
    model: i32,
    bending: i32,
    k_membrane: f64,
    k_bending: f64,
    density_kg_m2: f64,
    warp: f64,
    weft: f64,
    bend_warp: f64,
    bend_weft: f64,
}

/// ProjectiveDynamicsSolver on the GPU.
pub struct HipPdSolver {
    pub(crate) ctx: Ctx,
    settings: HipSettings,
    mass: Vec<f32>,
    pub(crate) n: usize,
    pub(crate) initialized: bool,
    ground: Option<f32>,
    pub(crate) ipc_ready: bool,
    /// Ground of the contact scene given to init_ipc.
    pub(crate) ipc_ground: Option<f32>,
}

impl HipPdSolver {
    /// A solver on HIP device `device` with the default settings.
    pub fn new(device: i32) -> VistioResult<Self> {
        Self::with_settings(device, HipSettings::default())
    }

    pub fn with_settings(device: i32, settings: HipSettings) -> VistioResult<Self> {
        Ok(Self {
            ctx: Ctx::new(device)?,
            settings,
            mass: Vec::new(),
            n: 0,
            initialized: false,
            ground: None,
            ipc_ready: false,
            ipc_ground: None,
        })
    }

    /// pd_solver.rs:95-97.
    pub fn lumped_masses(&self) -> &[f32] {
        &self.mass
    }

    pub(crate) fn require_init(&self) -> VistioResult<()> {
        if self.initialized {
            Ok(())
        } else {
            Err(VistioError::InvalidConfig("Solver not initialized. Call init() first.".into()))
        }
    }

    fn init_raw(
        &mut self,
        mesh: &TriangleMesh,
        topology: &Topology,
        config: &SolverConfig,
        pinned: &[bool],
        m: Material,
    ) -> VistioResult<()> {
        let n = mesh.vertex_count();
        if pinned.len() != n {
            return Err(VistioError::InvalidMesh(format!(
                "Pinned array length ({}) != vertex count ({n})",
                pinned.len()
            )));
        }
        // The library puts the hinges in a canonical order of its own, so the
        // HashMap order of Topology::build (topology.rs:88) does not matter.
        let hinges: Vec<u32> =
            topology.interior_edges.iter().flat_map(|e| [e.v0, e.v1, e.wing_a, e.wing_b]).collect();
        let pin: Vec<u8> = pinned.iter().map(|&p| p as u8).collect();
        let s = &self.settings;
        let p = VhPdParams {
            model: m.model,
            bending: m.bending,
            k_membrane: m.k_membrane,
            k_bending: m.k_bending,
            density: m.density_kg_m2,
            dt_matrix: 1.0 / 60.0, // the matrix dt every init path fixes (pd_solver.rs:244)
            gravity: config.gravity.map(|g| g as f64),
            damping: config.damping as f64,
            rayleigh: config.rayleigh_mass_damping as f64,
            max_iterations: config.max_iterations as i32,
            tolerance: config.tolerance,
            has_ground: 0,
            ground_height: 0.0,
            precision: s.precision,
            ir_sweeps: s.ir_sweeps,
            ir_final_sweeps: s.ir_final_sweeps,
            leaf_size: s.leaf_size,
            max_band_levels: s.max_band_levels,
            band_launch_us: s.band_launch_us,
            band_gbps: s.band_gbps,
            task_entries: s.task_entries,
            warp_stiffness: m.warp,
            weft_stiffness: m.weft,
            bend_warp: m.bend_warp,
            bend_weft: m.bend_weft,
        };
        let mut mass = vec![0.0f32; n];
        self.initialized = false;
        self.ipc_ready = false;
        self.ground = None;
        let rc = unsafe {
            vh_pd_init(
                self.ctx.0,
                n as u32,
                mesh.pos_x.as_ptr(),
                mesh.pos_y.as_ptr(),
                mesh.pos_z.as_ptr(),
                mesh.triangle_count() as u32,
                mesh.indices.as_ptr(),
                (hinges.len() / 4) as u32,
                hinges.as_ptr(),
                pin.as_ptr(),
                &p,
                mass.as_mut_ptr(),
            )
        };
        self.ctx.check(rc)?;
        self.mass = mass;
        self.n = n;
        self.initialized = true;
        Ok(())
    }

    /// init_with_material (pd_solver.rs:106-143): dihedral bending; element
    /// weight avg_stretch_stiffness * 10 * area (element.rs:106), hinge
    /// weight avg_bending_stiffness * 100 * edge length (bending.rs:160),
    /// mass from the areal density in g/m^2. The orthotropic and anisotropic
    /// models take warp/weft from the same properties their from_properties
    /// read.
    pub fn init_with_material(
        &mut self,
        mesh: &TriangleMesh,
        topology: &Topology,
        config: &SolverConfig,
        properties: &FabricProperties,
        model: Box<dyn ConstitutiveModel>,
        pinned: &[bool],
    ) -> VistioResult<()> {
        let m = Material {
            model: model_id(model.name())?,
            bending: VH_BENDING_DIHEDRAL,
            k_membrane: (properties.avg_stretch_stiffness() * 10.0) as f64,
            k_bending: (properties.avg_bending_stiffness() * 100.0) as f64,
            density_kg_m2: properties.density as f64 / 1000.0,
            warp: properties.stretch_stiffness_warp as f64,
            weft: properties.stretch_stiffness_weft as f64,
            bend_warp: 0.0,
            bend_weft: 0.0,
        };
        self.init_raw(mesh, topology, config, pinned, m)
    }

    /// init_with_material_tier3 (pd_solver.rs:152-193): Discrete Shells
    /// bending, per-hinge warp/weft modulation when the fabric is anisotropic
    /// (discrete_shells.rs:206-268).
    pub fn init_with_material_tier3(
        &mut self,
        mesh: &TriangleMesh,
        topology: &Topology,
        config: &SolverConfig,
        properties: &FabricProperties,
        model: Box<dyn ConstitutiveModel>,
        pinned: &[bool],
    ) -> VistioResult<()> {
        let aniso = properties.is_anisotropic();
        let m = Material {
            model: model_id(model.name())?,
            bending: VH_BENDING_DISCRETE_SHELLS,
            k_membrane: (properties.avg_stretch_stiffness() * 10.0) as f64,
            k_bending: (properties.avg_bending_stiffness() * 100.0) as f64,
            density_kg_m2: properties.density as f64 / 1000.0,
            warp: properties.stretch_stiffness_warp as f64,
            weft: properties.stretch_stiffness_weft as f64,
            bend_warp: if aniso { properties.bending_stiffness_warp as f64 } else { 0.0 },
            bend_weft: if aniso { properties.bending_stiffness_weft as f64 } else { 0.0 },
        };
        self.init_raw(mesh, topology, config, pinned, m)
    }

    /// init_with_material_tier4 (pd_solver.rs:201-214): the tier-3 set-up;
    /// contact is enabled with `init_ipc`.
    pub fn init_with_material_tier4(
        &mut self,
        mesh: &TriangleMesh,
        topology: &Topology,
        config: &SolverConfig,
        properties: &FabricProperties,
        model: Box<dyn ConstitutiveModel>,
        pinned: &[bool],
    ) -> VistioResult<()> {
        self.init_with_material_tier3(mesh, topology, config, properties, model, pinned)
    }

    /// The state arrays as the C-ABI takes them; checks every length.
    pub(crate) fn state_ptrs(&self, s: &mut SimulationState) -> VistioResult<VhStateF32> {
        let n = self.n;
        let lens = [
            s.pos_x.len(), s.pos_y.len(), s.pos_z.len(), s.vel_x.len(), s.vel_y.len(), s.vel_z.len(),
            s.prev_x.len(), s.prev_y.len(), s.prev_z.len(), s.pred_x.len(), s.pred_y.len(), s.pred_z.len(),
            s.inv_mass.len(),
        ];
        if s.vertex_count != n || lens.iter().any(|&l| l != n) {
            return Err(VistioError::InvalidMesh(format!(
                "state arrays must all have the mesh's vertex count ({n})"
            )));
        }
        Ok(VhStateF32 {
            pos_x: s.pos_x.as_mut_ptr(),
            pos_y: s.pos_y.as_mut_ptr(),
            pos_z: s.pos_z.as_mut_ptr(),
            vel_x: s.vel_x.as_mut_ptr(),
            vel_y: s.vel_y.as_mut_ptr(),
            vel_z: s.vel_z.as_mut_ptr(),
            prev_x: s.prev_x.as_mut_ptr(),
            prev_y: s.prev_y.as_mut_ptr(),
            prev_z: s.prev_z.as_mut_ptr(),
            pred_x: s.pred_x.as_mut_ptr(),
            pred_y: s.pred_y.as_mut_ptr(),
            pred_z: s.pred_z.as_mut_ptr(),
            inv_mass: s.inv_mass.as_ptr(),
        })
    }

    /// SimulationState::ground_height, handed to the device step when it
    /// changes.
    pub(crate) fn sync_ground(&mut self, ground: Option<f32>) -> VistioResult<()> {
        if ground != self.ground {
            let (on, h) = match ground {
                Some(h) => (1, h as f64),
                None => (0, 0.0),
            };
            self.ctx.check(unsafe { vh_pd_set_ground(self.ctx.0, on, h) })?;
            self.ground = ground;
        }
        Ok(())
    }

    /// `nsteps` steps on the device-resident state (no host copies): the
    /// state is the one `init` set up (positions of the mesh, zero velocity)
    /// or the last `upload_state`.
    pub fn step_resident(&mut self, dt: f32, nsteps: u32) -> VistioResult<StepResult> {
        self.require_init()?;
        let mut r = VhStepResult::default();
        self.ctx.check(unsafe { vh_pd_step_resident(self.ctx.0, dt as f64, nsteps, &mut r) })?;
        Ok(to_step_result(&r))
    }

    /// Positions of the device-resident state, SoA, in FP64.
    pub fn positions(&self) -> VistioResult<[Vec<f64>; 3]> {
        self.require_init()?;
        let (mut x, mut y, mut z) = (vec![0.0; self.n], vec![0.0; self.n], vec![0.0; self.n]);
        self.ctx.check(unsafe { vh_state_get_f64(self.ctx.0, 0, x.as_mut_ptr(), y.as_mut_ptr(), z.as_mut_ptr()) })?;
        Ok([x, y, z])
    }

    /// Replaces the device-resident state by a host SimulationState.
    pub fn upload_state(&mut self, s: &SimulationState) -> VistioResult<()> {
        self.require_init()?;
        let n = self.n;
        let soa = |a: &[f32]| a.iter().map(|&v| v as f64).collect::<Vec<f64>>();
        let sets = [
            (0, [&s.pos_x, &s.pos_y, &s.pos_z]),
            (1, [&s.vel_x, &s.vel_y, &s.vel_z]),
            (2, [&s.prev_x, &s.prev_y, &s.prev_z]),
            (3, [&s.pred_x, &s.pred_y, &s.pred_z]),
        ];
        for (which, [a, b, c]) in sets {
            if a.len() != n || b.len() != n || c.len() != n {
                return Err(VistioError::InvalidMesh(format!("state arrays must have {n} entries")));
            }
            let (x, y, z) = (soa(a.as_slice()), soa(b.as_slice()), soa(c.as_slice()));
            self.ctx.check(unsafe { vh_state_set_f64(self.ctx.0, which, x.as_ptr(), y.as_ptr(), z.as_ptr()) })?;
        }
        if s.inv_mass.len() != n {
            return Err(VistioError::InvalidMesh(format!("inv_mass must have {n} entries")));
        }
        self.ctx.check(unsafe { vh_state_set_inv_mass(self.ctx.0, s.inv_mass.as_ptr()) })?;
        self.sync_ground(s.ground_height)
    }
}

pub(crate) fn to_step_result(r: &VhStepResult) -> StepResult {
    StepResult {
        iterations: r.iterations,
        final_residual: r.final_residual,
        converged: r.converged != 0,
        wall_time: r.wall_time,
    }
}

impl SolverStrategy for HipPdSolver {
    /// pd_solver.rs:221-257: ARAP projection, element weight stretch_weight
    /// * 1000 * area, 200 g/m^2, dihedral hinges of weight bending_weight *
    /// 100 * edge length.
    fn init(
        &mut self,
        mesh: &TriangleMesh,
        topology: &Topology,
        config: &SolverConfig,
        pinned: &[bool],
    ) -> VistioResult<()> {
        let m = Material {
            model: VH_MODEL_ARAP,
            bending: VH_BENDING_DIHEDRAL,
            k_membrane: (config.stretch_weight * 1000.0) as f64,
            k_bending: (config.bending_weight * 100.0) as f64,
            density_kg_m2: 0.2,
            warp: 0.0,
            weft: 0.0,
            bend_warp: 0.0,
            bend_weft: 0.0,
        };
        self.init_raw(mesh, topology, config, pinned, m)
    }

    /// pd_solver.rs:259-476 on the host state: save_previous, predict, the
    /// PD iterations and the velocity update all run on the device; pos,
    /// vel, prev and pred come back. Pinning follows inv_mass == 0, the
    /// ground follows state.ground_height.
    fn step(&mut self, state: &mut SimulationState, dt: f32) -> VistioResult<StepResult> {
        self.require_init()?;
        self.sync_ground(state.ground_height)?;
        let mut st = self.state_ptrs(state)?;
        let mut r = VhStepResult::default();
        self.ctx.check(unsafe { vh_pd_step(self.ctx.0, dt as f64, &mut st, &mut r) })?;
        Ok(to_step_result(&r))
    }

    fn name(&self) -> &str {
        "ProjectiveDynamics"
    }
}
