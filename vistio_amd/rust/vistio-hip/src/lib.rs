//! # vistio-hip
//!
//! Vistio's Projective Dynamics step, IPC contact path and GPU backend on an
//! AMD MI355X (gfx950), behind the reference's own interfaces:
//!
//! - [`HipPdSolver`]: `SolverStrategy` (strategy.rs:42-69) plus the
//!   concrete solver's `init_with_material*`, `lumped_masses` and
//!   `step_with_ipc` (pd_solver.rs:95-214, 549);
//! - [`HipIpcHandler`]: `IpcCollisionHandler` (pd_solver.rs:483-489);
//! - [`HipBroadPhase`]: `BroadPhase` (broad.rs:39-56);
//! - [`HipNarrowPhase`]: `NarrowPhase` (narrow.rs:21-35), the
//!   `VertexTriangleTest` of the pre-IPC collision pipeline;
//! - [`HipBackend`]: `GpuBackend` (backend.rs:19-46).
//!
//! All of them call libvistio_hip.so through the C-ABI declared in
//! include/vistio_hip.h ([`ffi`]).

pub mod backend;
pub mod contact;
mod ctx;
pub mod ffi;
pub mod narrow;
pub mod solver;

pub use backend::HipBackend;
pub use contact::{DeviceScene, HipBroadPhase, HipIpcHandler, IpcScene};
pub use narrow::HipNarrowPhase;
pub use solver::{HipPdSolver, HipSettings};
