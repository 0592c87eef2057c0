"""The workloads the PD step is measured and checked on, as meshes plus
settings (no dataset: every mesh is a quad_grid, generators.rs:28-100).

* hanging_sheet: quad_grid(c, c, 1 m, 1 m) in the z = 0 plane, row 0 pinned
  (256 cells a side is the headline workload).
* sphere_drape: quad_grid(512, 512, 1.5 m, 1.5 m) laid flat at y = 0.8 m over
  an analytic sphere of radius 0.5 m, ground at y = -0.3 m.
* self_fold_ribbon: quad_grid(64, 640, 0.2 m, 2.0 m) turned 45 degrees about
  its short axis, lowest point 0.5 m above the ground at y = 0.
* reference_sphere_drape: the reference's own Scenario::sphere_drape with the
  set-up of its test verify_sphere_drape_no_penetration
  (crates/vistio-bench/src/scenarios.rs:120-148,
  crates/vistio-bench/tests/bench_tests.rs:138-224).
"""
from dataclasses import dataclass

import numpy as np

from .solver import IpcConfig, SolverConfig, Topology, TriangleMesh, estimate_initial_kappa, quad_grid


def hanging_sheet(cells: int):
    """quad_grid(cells, cells, 1, 1) in the z = 0 plane with row 0 (cells + 1
    vertices, y = +0.5) pinned: the mesh, its topology and the pin flags."""
    mesh = quad_grid(cells, cells, 1.0, 1.0)
    topo = Topology.build(mesh)
    pinned = np.zeros(mesh.vertex_count(), dtype=np.uint8)
    pinned[: cells + 1] = 1
    return mesh, topo, pinned


def contact_scene(name: str):
    """The two contact workloads of the benchmark: the sphere drape (d_hat
    1e-3, kappa 1e4, AL <= 5, 180 steps) and the self-fold ribbon
    (self-contact with 1-ring exclusion, d_hat 5e-4, AL <= 8, 300 steps).
    Both have self-collision on and compliant contact off. Returns (mesh,
    IpcConfig, steps, description).

    d_hat is a distance here: the barrier acts on squared distances, so it
    gets barrier_d_hat = d_hat^2, and the broad-phase margin / self-contact
    thickness is 1.5 d_hat (the reference runner's thickness = 1.5 *
    barrier_d_hat, read in metres). Read as m^2 instead, 1e-3 and 5e-4 would
    switch the self-contact barrier on between vertices two rings apart on
    the flat sheet (cells of 2.9 and 3.1 mm vs 3.2 and 2.2 cm activation
    radii)."""
    if name == "sphere_drape":
        cells, w = 512, 1.5
        m = quad_grid(cells, cells, w, w)
        px = m.pos_x.astype(np.float32)
        pz = m.pos_y.astype(np.float32)
        py = np.full_like(px, 0.8)
        dh, al, steps = 1e-3, 5, 180
        ipc = IpcConfig(ground_height=-0.3, sphere=((0.0, 0.0, 0.0), 0.5), self_collision=True,
                        barrier_d_hat=dh * dh, barrier_kappa=1e4, thickness=1.5 * dh, al_max_iterations=al,
                        compliant_contact=False)
        desc = (f"sphere_drape {cells}x{cells} cells, 1.5 m sheet flat at y 0.8 m over a sphere r 0.5 m, "
                f"ground -0.3 m, self-collision on (1-ring exclusion), d_hat 1e-3 m (barrier_d_hat 1e-6 m^2, "
                f"thickness 1.5e-3 m), kappa 1e4, AL <= 5")
    elif name == "self_fold_ribbon":
        m = quad_grid(64, 640, 0.2, 2.0)
        c = np.float32(np.sqrt(0.5))
        px = m.pos_x.astype(np.float32)
        y = m.pos_y.astype(np.float64) * c
        pz = (m.pos_y.astype(np.float64) * c).astype(np.float32)
        py = (y + 0.5 - y.min()).astype(np.float32)
        dh, al, steps = 5e-4, 8, 300
        ipc = IpcConfig(ground_height=0.0, self_collision=True, exclusion_depth=1, barrier_d_hat=dh * dh,
                        barrier_kappa=1e4, thickness=1.5 * dh, al_max_iterations=al, compliant_contact=False)
        desc = ("self_fold_ribbon 64x640 cells, 0.2 m x 2.0 m at 45 degrees, lowest point 0.5 m above the "
                "ground y 0, self-contact (1-ring exclusion), d_hat 5e-4 m (barrier_d_hat 2.5e-7 m^2, "
                "thickness 7.5e-4 m), kappa 1e4, AL <= 8")
    else:
        raise ValueError(f"unknown contact workload {name!r}")
    mesh = TriangleMesh(px, py, pz, m.indices)
    return mesh, ipc, steps, desc


@dataclass
class ReferenceDrape:
    mesh: TriangleMesh
    topology: Topology
    pinned: np.ndarray
    config: SolverConfig
    ipc: IpcConfig
    vertex_mass: float
    timesteps: int
    sphere_radius: float


def reference_sphere_drape(timesteps: int = 40) -> ReferenceDrape:
    """Scenario::sphere_drape (scenarios.rs:120-148) as
    verify_sphere_drape_no_penetration sets it up (bench_tests.rs:148-199):
    quad_grid(20, 20, 1.5, 1.5) turned into the XZ plane at y = 1, nothing
    pinned, SolverConfig::default() with ipc_enabled, the solver's init()
    (ARAP, 200 g/m^2, k = 1000 and 50), a pipeline with a sphere of radius
    0.3 at the origin and a ground at -0.3 (no self-collision; thickness
    barrier_d_hat), kappa = estimate_initial_kappa(0, 0.002, 9.81, 1e-3), and
    a handler that ignores set_d_hat: detection uses barrier_d_hat = 1e-3
    while the CFL clamp uses the compliant 2e-3. The state's ground height
    is -0.3."""
    m = quad_grid(20, 20, 1.5, 1.5)
    px = m.pos_x.astype(np.float32)
    pz = m.pos_y.astype(np.float32)
    py = np.full_like(px, np.float32(1.0))
    mesh = TriangleMesh(px, py, pz, m.indices)
    cfg = SolverConfig()
    d_hat = 1e-3
    kappa = estimate_initial_kappa(0.0, 0.002, 9.81, d_hat)
    ipc = IpcConfig(ground_height=-0.3, sphere=((0.0, 0.0, 0.0), 0.3), self_collision=False, barrier_d_hat=d_hat,
                    barrier_kappa=kappa, thickness=d_hat, al_max_iterations=10, al_mu_initial=1.0,
                    al_mu_growth=10.0, al_tolerance=1e-4, compliant_contact=True, compliant_d_hat_scale=2.0,
                    friction_coefficient=0.4, contact_damping=0.3, adaptive_contact_damping=True,
                    contact_damping_max=0.5, contact_velocity_threshold=0.1, detect_keeps_d_hat=True)
    return ReferenceDrape(mesh, Topology.build(mesh), np.zeros(mesh.vertex_count(), np.uint8), cfg, ipc, 0.002,
                          timesteps, 0.3)
