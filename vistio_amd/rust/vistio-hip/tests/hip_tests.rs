//! Tests of the binding against the reference's own CPU implementations, in
//! the same workspace: the GpuBackend vectors of
//! crates/vistio-gpu/tests/gpu_tests.rs, the solver against
//! ProjectiveDynamicsSolver on the 20 x 20 hanging sheet, and the broad phase
//! against BvhBroadPhase. They need an MI355X and libvistio_hip.so.

use vistio_contact::broad::BroadPhase;
use vistio_contact::BvhBroadPhase;
use vistio_gpu::{ComputeBuffer, CpuFallback, GpuBackend};
use vistio_contact::{broad::CandidatePair, narrow::NarrowPhase, VertexTriangleTest};
use vistio_hip::{DeviceScene, HipBackend, HipBroadPhase, HipNarrowPhase, HipPdSolver, HipSettings, IpcScene};
use vistio_material::{CoRotationalModel, FabricProperties};
use vistio_mesh::{generators::quad_grid, topology::Topology};
use vistio_solver::{
    config::SolverConfig, state::SimulationState, strategy::SolverStrategy,
    ProjectiveDynamicsSolver,
};

#[test]
fn hip_backend_matches_cpu_fallback() {
    let mut gpu = HipBackend::new(0);
    gpu.init().unwrap();
    assert_eq!(gpu.name(), "hip_gfx950");
    assert!(gpu.is_gpu());
    let a = ComputeBuffer::from_data(vec![1.0, 2.0, 3.0]);
    let b = ComputeBuffer::from_data(vec![4.0, 5.0, 6.0]);
    let mut out = ComputeBuffer::zeros(3);
    gpu.axpy(&a, &b, 2.0, &mut out).unwrap();
    assert_eq!(out.as_slice(), &[9.0, 12.0, 15.0]);
    assert!((gpu.dot(&a, &b).unwrap() - 32.0).abs() < 1e-10);
    let short = ComputeBuffer::from_data(vec![1.0, 2.0]);
    assert!(gpu.axpy(&short, &b, 1.0, &mut out).is_err());
    assert!(gpu.dot(&short, &b).is_err());
    // bit-exact against the CPU loop on larger vectors
    let cpu = CpuFallback::new();
    let n = 100_003;
    let x = ComputeBuffer::from_data((0..n).map(|i| ((i * 7919) % 1000) as f32 * 1e-3 - 0.5).collect());
    let y = ComputeBuffer::from_data((0..n).map(|i| ((i * 104_729) % 997) as f32 * 1e-2).collect());
    let (mut og, mut oc) = (ComputeBuffer::zeros(n), ComputeBuffer::zeros(n));
    gpu.axpy(&x, &y, 0.37, &mut og).unwrap();
    cpu.axpy(&x, &y, 0.37, &mut oc).unwrap();
    assert_eq!(og.as_slice(), oc.as_slice());
    let (dg, dc) = (gpu.dot(&x, &y).unwrap(), cpu.dot(&x, &y).unwrap());
    assert!((dg - dc).abs() <= 1e-9 * dc.abs().max(1.0));
    let mut buf = ComputeBuffer::zeros(5);
    gpu.fill(&mut buf, std::f32::consts::PI);
    assert!(buf.as_slice().iter().all(|&v| v == std::f32::consts::PI));
}

fn sheet(c: usize) -> (vistio_mesh::TriangleMesh, Topology, Vec<bool>) {
    let mesh = quad_grid(c, c, 1.0, 1.0);
    let topo = Topology::build(&mesh);
    let pinned: Vec<bool> = (0..mesh.vertex_count()).map(|i| i <= c).collect();
    (mesh, topo, pinned)
}

fn config() -> SolverConfig {
    SolverConfig { max_iterations: 10, tolerance: 0.0, damping: 0.0, rayleigh_mass_damping: 0.01, ..Default::default() }
}

/// The hanging sheet of the headline workload at 20 x 20, 120 steps, both
/// solvers from init_with_material with the same fabric. The reference keeps
/// f32 state and solves in f64 after an f32 round trip; the HIP step keeps
/// FP64 state. Their trajectories agree to what that f32 rounding allows.
#[test]
fn hip_solver_follows_projective_dynamics_solver() {
    let (mesh, topo, pinned) = sheet(20);
    let cfg = config();
    let props = FabricProperties {
        density: 150.0,
        stretch_stiffness_warp: 100.0,
        stretch_stiffness_weft: 100.0,
        bending_stiffness_warp: 1e-6,
        bending_stiffness_weft: 1e-6,
        ..FabricProperties::cotton_twill()
    };
    let mut cpu = ProjectiveDynamicsSolver::new();
    cpu.init_with_material(&mesh, &topo, &cfg, &props, Box::new(CoRotationalModel::new()), &pinned).unwrap();
    for prec in [vistio_hip::ffi::VH_PREC_F64, vistio_hip::ffi::VH_PREC_F32_IR] {
        let mut gpu = HipPdSolver::with_settings(0, HipSettings { precision: prec, ..Default::default() }).unwrap();
        gpu.init_with_material(&mesh, &topo, &cfg, &props, Box::new(CoRotationalModel::new()), &pinned).unwrap();
        for (a, b) in gpu.lumped_masses().iter().zip(cpu.lumped_masses()) {
            assert!((a - b).abs() <= 1e-6 * b.abs().max(1.0), "lumped mass {a} vs {b}");
        }
        let mut sc = SimulationState::from_mesh_with_masses(&mesh, cpu.lumped_masses(), &pinned).unwrap();
        let mut sg = SimulationState::from_mesh_with_masses(&mesh, gpu.lumped_masses(), &pinned).unwrap();
        let mut cpu_run = ProjectiveDynamicsSolver::new();
        cpu_run.init_with_material(&mesh, &topo, &cfg, &props, Box::new(CoRotationalModel::new()), &pinned).unwrap();
        for _ in 0..120 {
            cpu_run.step(&mut sc, 1.0 / 60.0).unwrap();
            let r = gpu.step(&mut sg, 1.0 / 60.0).unwrap();
            assert_eq!(r.iterations, 10);
        }
        let mut worst = 0.0f32;
        for i in 0..mesh.vertex_count() {
            let d = ((sc.pos_x[i] - sg.pos_x[i]).powi(2) + (sc.pos_y[i] - sg.pos_y[i]).powi(2)
                + (sc.pos_z[i] - sg.pos_z[i]).powi(2))
            .sqrt();
            worst = worst.max(d);
        }
        assert!(worst < 1e-4, "max vertex distance {worst} m after 120 steps");
    }
}

#[test]
fn hip_broad_phase_contains_bvh_pairs() {
    let mut mesh = quad_grid(12, 12, 1.0, 1.0);
    for (i, z) in mesh.pos_z.iter_mut().enumerate() {
        *z = (((i * 2654435761usize) % 1000) as f32 / 1000.0 - 0.5) * 0.1;
    }
    let mut cpu = BvhBroadPhase::new(&mesh);
    let mut gpu = HipBroadPhase::new(0, &mesh).unwrap();
    cpu.update(&mesh.pos_x, &mesh.pos_y, &mesh.pos_z, 0.01).unwrap();
    gpu.update(&mesh.pos_x, &mesh.pos_y, &mesh.pos_z, 0.01).unwrap();
    let all: std::collections::HashSet<(u32, u32)> =
        gpu.query_triangle_pairs().iter().map(|p| (p.ta.min(p.tb), p.ta.max(p.tb))).collect();
    for p in cpu.query_triangle_pairs() {
        assert!(all.contains(&(p.ta.min(p.tb), p.ta.max(p.tb))), "pair ({}, {}) missing", p.ta, p.tb);
    }
    assert!(all.len() >= cpu.query_triangle_pairs().len());
}

#[test]
fn hip_narrow_phase_matches_vertex_triangle_test() {
    // a sheet folded back over itself 6 mm above, every vertex pair closer than 5 cm
    let mut mesh = quad_grid(16, 16, 1.0, 1.0);
    for i in 0..mesh.vertex_count() {
        mesh.pos_z[i] = (((i * 2654435761usize) % 1000) as f32 / 1000.0 - 0.5) * 0.008;
        if mesh.pos_x[i] > 0.0 {
            mesh.pos_x[i] = -mesh.pos_x[i];
            mesh.pos_z[i] += 0.006;
        }
    }
    let n = mesh.vertex_count();
    let pinned = vec![false; n];
    let state = SimulationState::from_mesh(&mesh, 0.15, &pinned).unwrap();
    let mut cand = Vec::new();
    for a in 0..n {
        for b in 0..n {
            let d = (mesh.pos_x[a] - mesh.pos_x[b]).powi(2) + (mesh.pos_y[a] - mesh.pos_y[b]).powi(2)
                + (mesh.pos_z[a] - mesh.pos_z[b]).powi(2);
            if a != b && d < 0.05 * 0.05 {
                cand.push(CandidatePair { a: a as u32, b: b as u32, is_self: true });
            }
        }
    }
    let cpu = VertexTriangleTest.detect(&cand, &state, &mesh, 0.01).unwrap();
    let hip = HipNarrowPhase::new(0).unwrap();
    let gpu = hip.detect(&cand, &state, &mesh, 0.01).unwrap();
    assert_eq!(hip.name(), "vertex_triangle_test");
    assert!(cpu.len() > 50);
    assert_eq!(gpu.len(), cpu.len());
    for (g, c) in gpu.iter().zip(&cpu) {
        assert_eq!(g.indices, c.indices);
        assert_eq!(g.distance.to_bits(), c.distance.to_bits());
        assert_eq!(g.normal.map(f32::to_bits), c.normal.map(f32::to_bits));
        assert_eq!(g.barycentric.map(f32::to_bits), c.barycentric.map(f32::to_bits));
        assert_eq!(g.is_self, c.is_self);
    }
}

#[test]
fn hip_step_with_ipc_rests_on_the_ground() {
    let mut mesh = quad_grid(16, 16, 0.5, 0.5);
    for i in 0..mesh.vertex_count() {
        mesh.pos_z[i] = mesh.pos_y[i];
        mesh.pos_y[i] = 0.03;
    }
    let topo = Topology::build(&mesh);
    let pinned = vec![false; mesh.vertex_count()];
    let mut cfg = config();
    cfg.ipc_enabled = true;
    cfg.barrier_d_hat = 1e-3;
    cfg.compliant_contact = false;
    let mut gpu = HipPdSolver::new(0).unwrap();
    gpu.init_with_material_tier4(&mesh, &topo, &cfg, &FabricProperties::cotton_twill(),
                                 Box::new(CoRotationalModel::new()), &pinned).unwrap();
    let scene = IpcScene { ground: Some(0.0), thickness: 1.5e-3, exclusion_depth: 1, ..Default::default() };
    gpu.init_ipc(&scene, &cfg, 100.0).unwrap();
    let mut st = SimulationState::from_mesh_with_masses(&mesh, gpu.lumped_masses(), &pinned).unwrap();
    st.ground_height = Some(0.0);
    for _ in 0..60 {
        gpu.step_with_ipc(&mut st, 1.0 / 60.0, &mut DeviceScene).unwrap();
    }
    assert!(st.pos_y.iter().all(|&y| y.is_finite() && y > -1e-3));
    assert!(st.pos_y.iter().any(|&y| y < 0.01));
    // the device step clamps to the scene's ground, so another state ground is refused
    st.ground_height = Some(0.5);
    assert!(gpu.step_with_ipc(&mut st, 1.0 / 60.0, &mut DeviceScene).is_err());
}
