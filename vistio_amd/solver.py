"""Host-side mirror of Vistio's solver interface over the HIP C ABI.

Mirrors, with the same names, argument meaning and error behaviour:
  SolverStrategy (init / step / name)      crates/vistio-solver/src/strategy.rs:42-69
  ProjectiveDynamicsSolver                 crates/vistio-solver/src/pd_solver.rs:53-476
    init()               -> ARAP, k = stretch_weight*1000, 200 g/m^2, k_b = bending_weight*100 (:221-257)
    init_with_material() -> model by ConstitutiveModel::name(), k = avg_stretch*10,
                            k_b = avg_bending*100, density from properties (:106-143)
    lumped_masses()      (:95-97)
  SolverConfig                             crates/vistio-solver/src/config.rs:10-166
  SimulationState                          crates/vistio-solver/src/state.rs:23-140
  StepResult                               crates/vistio-solver/src/strategy.rs:15-24
  VistioError                              crates/vistio-types/src/error.rs:9-44
  quad_grid / Topology                     crates/vistio-mesh/src/generators.rs:28-100, topology.rs:58-120

All numerical work runs in libvistio_hip.so on the GPU; this module only
moves arrays across the C ABI.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _native as nat


class VistioError(Exception):
    """Base error (VistioError)."""


class InvalidConfig(VistioError):
    pass


class InvalidMesh(VistioError):
    pass


class GpuError(VistioError):
    pass


_STATUS_TO_ERROR = {1: InvalidConfig, 2: InvalidConfig, 3: InvalidConfig, 4: InvalidConfig, 5: InvalidMesh,
                    6: GpuError, 7: GpuError, 8: InvalidConfig, 9: InvalidConfig}


def _check(status: int, ctx=None, what: str = "", err_cls=None):
    """Raises the VistioError variant of a vh_status (err_cls overrides it:
    the GpuBackend calls report every failure as VistioError::Gpu)."""
    if status == nat.VH_OK:
        return
    msg = ""
    if ctx:
        raw = nat.load().vh_last_error(ctx)
        msg = raw.decode() if raw else ""
    cls = err_cls or _STATUS_TO_ERROR.get(status, VistioError)
    raise cls(f"{what}: {nat.STATUS_NAMES.get(status, status)}: {msg}".strip(": "))


@dataclass
class SolverConfig:
    """SolverConfig fields used by the PD step (config.rs:10-166 defaults)."""
    max_iterations: int = 10
    tolerance: float = 1e-6
    gravity: Sequence[float] = (0.0, -9.81, 0.0)
    damping: float = 0.01
    stretch_weight: float = 1.0
    bending_weight: float = 0.5
    rayleigh_mass_damping: float = 2.0
    # HIP-specific knobs (no reference counterpart)
    precision: str = "f64"          # "f64" or "f32_ir"
    ir_sweeps: int = 1
    ir_final_sweeps: int = 3
    leaf_size: int = 16
    max_band_levels: int = 0        # 0 = default; 1 = one solve level per dissection level
    band_launch_us: float = 0.0     # 0 = default
    band_gbps: float = 0.0          # 0 = default
    task_entries: int = 0           # 0 = default (1024 block entries per solve task)


@dataclass
class FabricProperties:
    """The FabricProperties fields init_with_material reads (properties.rs:23-98)."""
    name: str = "custom"
    density: float = 150.0                  # g/m^2
    stretch_stiffness_warp: float = 100.0
    stretch_stiffness_weft: float = 100.0
    bending_stiffness_warp: float = 1e-6
    bending_stiffness_weft: float = 1e-6

    def avg_stretch_stiffness(self) -> float:
        return (self.stretch_stiffness_warp + self.stretch_stiffness_weft) / 2.0

    def avg_bending_stiffness(self) -> float:
        return (self.bending_stiffness_warp + self.bending_stiffness_weft) / 2.0

    def is_anisotropic(self) -> bool:
        ratio = self.stretch_stiffness_warp / self.stretch_stiffness_weft if self.stretch_stiffness_weft > 1e-8 else 1.0
        return abs(ratio - 1.0) > 0.05


class MaterialDatabase:
    """MaterialDatabase::with_defaults (database.rs:20-33): the five fabric
    presets (database.rs:80-160), the fields the solver reads."""

    _PRESETS = (  # name, density g/m^2, stretch warp, weft, bending warp, weft
        ("cotton_twill", 200.0, 0.90, 0.85, 0.50, 0.45),
        ("silk_charmeuse", 80.0, 0.70, 0.65, 0.10, 0.08),
        ("denim_14oz", 400.0, 0.98, 0.92, 0.85, 0.75),
        ("jersey_knit", 180.0, 0.50, 0.60, 0.20, 0.25),
        ("chiffon", 50.0, 0.60, 0.55, 0.05, 0.04),
    )

    def __init__(self):
        self.materials = {}

    @classmethod
    def with_defaults(cls) -> "MaterialDatabase":
        db = cls()
        for name, dens, sw, sf, bw, bf in cls._PRESETS:
            db.register(FabricProperties(name, dens, sw, sf, bw, bf))
        return db

    def register(self, props: "FabricProperties") -> None:
        self.materials[props.name] = props

    def get(self, name: str):
        return self.materials.get(name)


@dataclass
class TriangleMesh:
    pos_x: np.ndarray
    pos_y: np.ndarray
    pos_z: np.ndarray
    indices: np.ndarray

    def vertex_count(self) -> int:
        return int(self.pos_x.size)

    def triangle_count(self) -> int:
        return int(self.indices.size // 3)


def quad_grid(cols: int, rows: int, width: float, height: float) -> TriangleMesh:
    n = (cols + 1) * (rows + 1)
    px, py, pz = (np.empty(n, dtype=np.float32) for _ in range(3))
    idx = np.empty(6 * cols * rows, dtype=np.uint32)
    _check(nat.load().vh_mesh_quad_grid(cols, rows, width, height, nat.ptr(px, C.c_float), nat.ptr(py, C.c_float),
                                        nat.ptr(pz, C.c_float), nat.ptr(idx, C.c_uint32)), what="quad_grid")
    return TriangleMesh(px, py, pz, idx)


@dataclass
class Topology:
    """Interior edges (v0, v1, wing_a, wing_b, tri_a, tri_b), sorted by (v0, v1)."""
    interior_edges: np.ndarray

    @staticmethod
    def build(mesh: TriangleMesh) -> "Topology":
        L = nat.load()
        n, t = mesh.vertex_count(), mesh.triangle_count()
        idx = np.ascontiguousarray(mesh.indices, dtype=np.uint32)
        cnt = C.c_uint32(0)
        _check(L.vh_mesh_interior_edges(n, t, nat.ptr(idx, C.c_uint32), None, 0, C.byref(cnt)), what="topology")
        out = np.empty((cnt.value, 6), dtype=np.uint32)
        _check(L.vh_mesh_interior_edges(n, t, nat.ptr(idx, C.c_uint32), nat.ptr(out, C.c_uint32), cnt.value,
                                        C.byref(cnt)), what="topology")
        return Topology(out)


@dataclass
class SimulationState:
    vertex_count: int
    pos_x: np.ndarray
    pos_y: np.ndarray
    pos_z: np.ndarray
    vel_x: np.ndarray
    vel_y: np.ndarray
    vel_z: np.ndarray
    prev_x: np.ndarray
    prev_y: np.ndarray
    prev_z: np.ndarray
    pred_x: np.ndarray
    pred_y: np.ndarray
    pred_z: np.ndarray
    mass: np.ndarray
    inv_mass: np.ndarray
    # augmented-Lagrangian multipliers of the contact step (state.rs:55-63)
    al_lambda_x: np.ndarray = None
    al_lambda_y: np.ndarray = None
    al_lambda_z: np.ndarray = None
    # ground plane height (Y) enforced by the plain step (state.rs:52-55)
    ground_height: Optional[float] = None

    def __post_init__(self):
        for nm in ("al_lambda_x", "al_lambda_y", "al_lambda_z"):
            if getattr(self, nm) is None:
                setattr(self, nm, np.zeros(self.vertex_count, dtype=np.float32))

    @staticmethod
    def from_mesh_with_masses(mesh: TriangleMesh, masses, pinned) -> "SimulationState":
        n = mesh.vertex_count()
        pinned = np.asarray(pinned)
        masses = np.asarray(masses, dtype=np.float32)
        if pinned.size != n:
            raise InvalidMesh(f"Pinned array length ({pinned.size}) != vertex count ({n})")
        if masses.size != n:
            raise InvalidMesh(f"Mass array length ({masses.size}) != vertex count ({n})")
        mass = masses.copy()
        inv = np.zeros(n, dtype=np.float32)
        free = ~pinned.astype(bool)
        mass[~free] = np.finfo(np.float32).max
        ok = free & (mass > 1e-12)
        inv[ok] = (1.0 / mass[ok]).astype(np.float32)
        z = lambda: np.zeros(n, dtype=np.float32)
        return SimulationState(n, mesh.pos_x.copy(), mesh.pos_y.copy(), mesh.pos_z.copy(), z(), z(), z(),
                               mesh.pos_x.copy(), mesh.pos_y.copy(), mesh.pos_z.copy(), z(), z(), z(), mass, inv)


@dataclass
class StepResult:
    iterations: int
    final_residual: float
    converged: bool
    wall_time: float


_MODEL_BY_NAME = {"arap": nat.MODEL_ARAP, "co_rotational": nat.MODEL_COROTATIONAL,
                  "isotropic_linear": nat.MODEL_ISOTROPIC_LINEAR, "orthotropic_linear": nat.MODEL_ORTHOTROPIC_LINEAR,
                  "anisotropic_corotational": nat.MODEL_ANISOTROPIC_COROTATIONAL}


@dataclass
class IpcConfig:
    """Contact settings: SolverConfig's IPC fields (config.rs:47-120, defaults
    config.rs:113-133) plus the collision pipeline the runner builds
    (runner.rs:60-95): ground plane, analytic sphere, self-collision."""
    ground_height: Optional[float] = None
    sphere: Optional[tuple] = None            # ((cx, cy, cz), radius)
    self_collision: bool = False
    exclusion_depth: int = 1
    barrier_d_hat: float = 1e-3
    barrier_kappa: float = 1e4                # the handler's kappa (see estimate_initial_kappa)
    thickness: Optional[float] = None         # None: 1.5 * barrier_d_hat (runner.rs:60)
    al_max_iterations: int = 10
    al_mu_initial: float = 1.0
    al_mu_growth: float = 10.0
    al_tolerance: float = 1e-4
    compliant_contact: bool = True
    compliant_d_hat_scale: float = 2.0
    friction_coefficient: float = 0.4
    contact_damping: float = 0.3
    adaptive_contact_damping: bool = True
    contact_damping_max: float = 0.5
    contact_velocity_threshold: float = 0.1
    max_pairs: int = 0                        # initial pair buffer, grows on demand; 0: 32 per triangle
    # The handler ignores set_d_hat (the trait's no-op, pd_solver.rs:488):
    # detection keeps barrier_d_hat, the CFL clamp uses the compliant one.
    detect_keeps_d_hat: bool = False
    # The CCD step adds edge_edge_ccd (ccd.rs:138-229) on each candidate pair.
    ccd_edge_edge: bool = False


def estimate_initial_kappa(configured: float, avg_mass: float, gravity_mag: float, d_hat: float) -> float:
    """barrier.rs:102-121: kappa = m g / (2 sqrt(d_hat/2) |b'(d_hat/2)|) in
    [1, 1e6]; a configured kappa > 0 is returned as is."""
    if configured > 0:
        return float(configured)
    dm = 0.5 * d_hat
    diff = dm - d_hat
    g = abs(-diff * (2.0 * np.log(dm / d_hat) + diff / dm))
    if g < 1e-12:
        return 1e4
    return float(min(max(avg_mass * gravity_mag / (2.0 * np.sqrt(dm) * g), 1.0), 1e6))


class HipPdSolver:
    """ProjectiveDynamicsSolver on the GPU (a SolverStrategy implementor)."""

    def __init__(self, device: int = 0):
        self._L = nat.load()
        self._ctx = C.c_void_p()
        _check(self._L.vh_create(device, C.byref(self._ctx)), what="vh_create")
        self._n = 0
        self._initialized = False
        self._mass = None
        self.config: Optional[SolverConfig] = None

    def close(self):
        if self._ctx:
            self._L.vh_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def name(self) -> str:
        return "ProjectiveDynamics"

    # -- init -------------------------------------------------------------
    def _init(self, mesh: TriangleMesh, topology: Topology, config: SolverConfig, pinned, model: int,
              k_membrane, k_bending, density_kgm2, bending: int = nat.BENDING_DIHEDRAL, warp: float = 0.0,
              weft: float = 0.0, bend_warp: float = 0.0, bend_weft: float = 0.0):
        """k_membrane, k_bending and density_kgm2 are each a scalar or a
        per-element array (triangles, interior edges, triangles)."""
        n = mesh.vertex_count()
        pin = np.ascontiguousarray(np.asarray(pinned).astype(np.uint8))
        if pin.size != n:
            raise InvalidMesh(f"Pinned array length ({pin.size}) != vertex count ({n})")
        if config.precision not in ("f64", "f32_ir"):
            raise InvalidConfig(f"unknown precision {config.precision!r}")
        p = nat.PdParams()
        p.model = model
        p.bending = bending
        p.warp_stiffness, p.weft_stiffness = float(warp), float(weft)
        p.bend_warp, p.bend_weft = float(bend_warp), float(bend_weft)
        per = {}
        for key, val, count in (("k_membrane", k_membrane, mesh.triangle_count()),
                                ("k_bending", k_bending, topology.interior_edges.shape[0]),
                                ("density", density_kgm2, mesh.triangle_count())):
            if np.ndim(val) == 0:
                setattr(p, key, float(val))
            else:
                a = np.ascontiguousarray(val, dtype=np.float64)
                if a.size != count:
                    raise InvalidMesh(f"{key} has {a.size} entries, expected {count}")
                per[key] = a
                setattr(p, key, float(a.mean()) if a.size else 0.0)
        p.dt_matrix = 1.0 / 60.0  # pd_solver.rs:130,180,244
        for k in range(3):
            p.gravity[k] = float(config.gravity[k])
        p.damping, p.rayleigh = float(config.damping), float(config.rayleigh_mass_damping)
        p.max_iterations, p.tolerance = int(config.max_iterations), float(config.tolerance)
        p.has_ground, p.ground_height = 0, 0.0
        p.precision = nat.PREC_F64 if config.precision == "f64" else nat.PREC_F32_IR
        p.ir_sweeps, p.ir_final_sweeps, p.leaf_size = int(config.ir_sweeps), int(config.ir_final_sweeps), int(config.leaf_size)
        p.max_band_levels = int(config.max_band_levels)
        p.band_launch_us, p.band_gbps = float(config.band_launch_us), float(config.band_gbps)
        p.task_entries = int(config.task_entries)
        px, py, pz = (np.ascontiguousarray(a, dtype=np.float32) for a in (mesh.pos_x, mesh.pos_y, mesh.pos_z))
        tri = np.ascontiguousarray(mesh.indices, dtype=np.uint32)
        hg = np.ascontiguousarray(topology.interior_edges[:, :4], dtype=np.uint32)
        self._mass = np.empty(n, dtype=np.float32)
        self._params = p
        mat = nat.MaterialArrays(*[nat.ptr(per[k], C.c_double) if k in per else None
                                   for k in ("k_membrane", "k_bending", "density")])
        _check(self._L.vh_pd_init_ex(self._ctx, n, nat.ptr(px, C.c_float), nat.ptr(py, C.c_float),
                                     nat.ptr(pz, C.c_float), tri.size // 3, nat.ptr(tri, C.c_uint32), hg.shape[0],
                                     nat.ptr(hg, C.c_uint32), nat.ptr(pin, C.c_uint8), C.byref(p),
                                     C.byref(mat), nat.ptr(self._mass, C.c_float)),
               self._ctx, "init")
        self._n = n
        self.config = config
        self._initialized = True

    def init(self, mesh: TriangleMesh, topology: Topology, config: SolverConfig, pinned):
        """SolverStrategy::init (pd_solver.rs:221-257): ARAP, 200 g/m^2."""
        self._init(mesh, topology, config, pinned, nat.MODEL_ARAP, float(np.float32(config.stretch_weight) * np.float32(1000.0)),
                   float(np.float32(config.bending_weight) * np.float32(100.0)), 0.2)

    def init_with_material(self, mesh: TriangleMesh, topology: Topology, config: SolverConfig,
                           properties: FabricProperties, model: str, pinned):
        """init_with_material (pd_solver.rs:106-143). model = ConstitutiveModel::name()."""
        if model not in _MODEL_BY_NAME:
            raise InvalidConfig(f"unsupported constitutive model {model!r}")
        self._init(mesh, topology, config, pinned, _MODEL_BY_NAME[model], properties.avg_stretch_stiffness() * 10.0,
                   properties.avg_bending_stiffness() * 100.0, properties.density / 1000.0,
                   warp=properties.stretch_stiffness_warp, weft=properties.stretch_stiffness_weft)

    def init_with_material_tier3(self, mesh: TriangleMesh, topology: Topology, config: SolverConfig,
                                 properties: FabricProperties, model: str, pinned):
        """init_with_material_tier3 (pd_solver.rs:152-193): Discrete Shells
        bending, warp/weft-modulated for anisotropic fabrics
        (discrete_shells.rs:206-268); the orthotropic and anisotropic models
        take their warp/weft stiffness from the properties (from_properties)."""
        if model not in _MODEL_BY_NAME:
            raise InvalidConfig(f"unsupported constitutive model {model!r}")
        aniso = properties.is_anisotropic()
        self._init(mesh, topology, config, pinned, _MODEL_BY_NAME[model], properties.avg_stretch_stiffness() * 10.0,
                   properties.avg_bending_stiffness() * 100.0, properties.density / 1000.0,
                   bending=nat.BENDING_DISCRETE_SHELLS, warp=properties.stretch_stiffness_warp,
                   weft=properties.stretch_stiffness_weft,
                   bend_warp=properties.bending_stiffness_warp if aniso else 0.0,
                   bend_weft=properties.bending_stiffness_weft if aniso else 0.0)

    def init_with_material_tier4(self, mesh: TriangleMesh, topology: Topology, config: SolverConfig,
                                 properties: FabricProperties, model: str, pinned):
        """init_with_material_tier4 (pd_solver.rs:203-214) = tier 3; contact
        is then enabled with init_ipc."""
        self.init_with_material_tier3(mesh, topology, config, properties, model, pinned)

    def init_raw(self, mesh: TriangleMesh, topology: Topology, config: SolverConfig, pinned, model: str,
                 k_membrane, k_bending, density_kgm2, bending: str = "dihedral", warp: float = 0.0,
                 weft: float = 0.0, bend_warp: float = 0.0, bend_weft: float = 0.0):
        """Init with explicit membrane/bending weights and areal density
        (kg/m^2): scalars, or per-element arrays (per triangle, per interior
        edge, per triangle) for several garments in one system."""
        if model not in _MODEL_BY_NAME:
            raise InvalidConfig(f"unsupported constitutive model {model!r}")
        kinds = {"dihedral": nat.BENDING_DIHEDRAL, "discrete_shells": nat.BENDING_DISCRETE_SHELLS}
        if bending not in kinds:
            raise InvalidConfig(f"unsupported bending model {bending!r}")
        self._init(mesh, topology, config, pinned, _MODEL_BY_NAME[model], k_membrane, k_bending, density_kgm2,
                   bending=kinds[bending], warp=warp, weft=weft, bend_warp=bend_warp, bend_weft=bend_weft)

    def lumped_masses(self) -> np.ndarray:
        if not self._initialized:
            raise InvalidConfig("Solver not initialized. Call init() first.")
        return self._mass

    # -- stepping -----------------------------------------------------------
    def step(self, state: SimulationState, dt: float) -> StepResult:
        """SolverStrategy::step on host state (uploads, steps on the GPU, downloads)."""
        if not self._initialized:
            raise InvalidConfig("Solver not initialized. Call init() first.")
        n = self._n
        arrs = [state.pos_x, state.pos_y, state.pos_z, state.vel_x, state.vel_y, state.vel_z,
                state.prev_x, state.prev_y, state.prev_z, state.pred_x, state.pred_y, state.pred_z, state.inv_mass]
        for a in arrs:
            if a.dtype != np.float32 or a.size != n or not a.flags.c_contiguous:
                raise InvalidMesh("state arrays must be contiguous float32 of the mesh's vertex count")
        self.set_ground_height(state.ground_height)
        s = nat.StateF32(*[nat.ptr(a, C.c_float) for a in arrs])
        r = nat.StepResultC()
        _check(self._L.vh_pd_step(self._ctx, float(dt), C.byref(s), C.byref(r)), self._ctx, "step")
        return StepResult(r.iterations, r.final_residual, bool(r.converged), r.wall_time)

    def set_ground_height(self, height: Optional[float]) -> None:
        """The ground plane (Y = height, None = none) of the plain step: what
        SimulationState.ground_height does for the reference's step (predict
        clamp, enforce_ground, enforce_ground_velocities)."""
        if not self._initialized:
            raise InvalidConfig("Solver not initialized. Call init() first.")
        on = height is not None
        _check(self._L.vh_pd_set_ground(self._ctx, int(on), float(np.float32(height)) if on else 0.0), self._ctx,
               "set_ground_height")

    def step_resident(self, dt: float, nsteps: int = 1) -> StepResult:
        if not self._initialized:
            raise InvalidConfig("Solver not initialized. Call init() first.")
        r = nat.StepResultC()
        _check(self._L.vh_pd_step_resident(self._ctx, float(dt), int(nsteps), C.byref(r)), self._ctx, "step_resident")
        return StepResult(r.iterations, r.final_residual, bool(r.converged), r.wall_time)

    # -- contact ------------------------------------------------------------
    def init_ipc(self, cfg: IpcConfig) -> None:
        """Enables step_with_ipc on this solver (after init)."""
        if not self._initialized:
            raise InvalidConfig("Solver not initialized. Call init() first.")
        q = nat.IpcParams()
        q.has_ground = 0 if cfg.ground_height is None else 1
        q.ground_height = 0.0 if cfg.ground_height is None else float(cfg.ground_height)
        q.has_sphere = 0 if cfg.sphere is None else 1
        if cfg.sphere is not None:
            (cx, cy, cz), r = cfg.sphere
            q.sphere_center[0], q.sphere_center[1], q.sphere_center[2] = cx, cy, cz
            q.sphere_radius = float(r)
        q.self_collision = int(bool(cfg.self_collision))
        q.exclusion_depth = int(cfg.exclusion_depth)
        q.d_hat, q.kappa = float(cfg.barrier_d_hat), float(cfg.barrier_kappa)
        q.thickness = 1.5 * cfg.barrier_d_hat if cfg.thickness is None else float(cfg.thickness)
        q.al_max_iterations = int(cfg.al_max_iterations)
        q.al_mu_initial, q.al_mu_growth, q.al_tolerance = cfg.al_mu_initial, cfg.al_mu_growth, cfg.al_tolerance
        q.compliant_contact, q.compliant_d_hat_scale = int(cfg.compliant_contact), cfg.compliant_d_hat_scale
        q.friction_coefficient, q.contact_damping = cfg.friction_coefficient, cfg.contact_damping
        q.adaptive_contact_damping = int(cfg.adaptive_contact_damping)
        q.contact_damping_max, q.contact_velocity_threshold = cfg.contact_damping_max, cfg.contact_velocity_threshold
        q.max_pairs = int(cfg.max_pairs)
        q.detect_keeps_d_hat = int(bool(cfg.detect_keeps_d_hat))
        q.ccd_edge_edge = int(bool(cfg.ccd_edge_edge))
        _check(self._L.vh_ipc_init(self._ctx, C.byref(q)), self._ctx, "init_ipc")

    def step_with_ipc(self, state: SimulationState, dt: float) -> StepResult:
        """ProjectiveDynamicsSolver::step_with_ipc on host state."""
        n = self._n
        arrs = [state.pos_x, state.pos_y, state.pos_z, state.vel_x, state.vel_y, state.vel_z,
                state.prev_x, state.prev_y, state.prev_z, state.pred_x, state.pred_y, state.pred_z, state.inv_mass]
        lam = [state.al_lambda_x, state.al_lambda_y, state.al_lambda_z]
        for a in arrs + lam:
            if a.dtype != np.float32 or a.size != n or not a.flags.c_contiguous:
                raise InvalidMesh("state arrays must be contiguous float32 of the mesh's vertex count")
        s = nat.StateF32(*[nat.ptr(a, C.c_float) for a in arrs])
        r = nat.StepResultC()
        _check(self._L.vh_pd_step_ipc(self._ctx, float(dt), C.byref(s), *[nat.ptr(a, C.c_float) for a in lam],
                                      C.byref(r)), self._ctx, "step_with_ipc")
        return StepResult(r.iterations, r.final_residual, bool(r.converged), r.wall_time)

    def step_ipc_resident(self, dt: float, nsteps: int = 1):
        """step_with_ipc with the state in HBM; returns (StepResult, stats)."""
        r, st = nat.StepResultC(), nat.IpcStats()
        _check(self._L.vh_pd_step_ipc_resident(self._ctx, float(dt), int(nsteps), C.byref(r), C.byref(st)),
               self._ctx, "step_ipc_resident")
        return (StepResult(r.iterations, r.final_residual, bool(r.converged), r.wall_time),
                {k: getattr(st, k) for k, _ in nat.IpcStats._fields_})

    def detect_contacts(self, pos, d_hat: float, kappa: float) -> dict:
        """CollisionPipeline::detect_ipc_contacts at positions (3, n), or the
        resident state when pos is None (IpcBarrierForces, pd_solver.rs:1160-1196)."""
        n = self._n
        p = None if pos is None else np.ascontiguousarray(np.asarray(pos, dtype=np.float64).reshape(3, n))
        g, nr = np.empty((3, n)), np.empty((3, n))
        ic = np.empty(n, dtype=np.uint8)
        hd = np.empty(n)
        st = nat.IpcStats()
        _check(self._L.vh_ipc_detect(self._ctx, None if p is None else nat.ptr(p, C.c_double), float(d_hat),
                                     float(kappa), nat.ptr(g, C.c_double), nat.ptr(nr, C.c_double),
                                     nat.ptr(ic, C.c_uint8), nat.ptr(hd, C.c_double), C.byref(st)),
               self._ctx, "detect_contacts")
        return dict(grad=g, normal=nr, in_contact=ic, hessian_diag=hd, max_violation=st.max_violation,
                    active=st.active, self_contacts=st.self_contacts, pairs=st.pairs)

    def compute_ccd_step(self, p0, p1) -> float:
        """CollisionPipeline::compute_ccd_step from p0 to p1 ((3, n) each)."""
        n = self._n
        a = np.ascontiguousarray(np.asarray(p0, dtype=np.float64).reshape(3, n))
        b = np.ascontiguousarray(np.asarray(p1, dtype=np.float64).reshape(3, n))
        out = C.c_double(0)
        _check(self._L.vh_ipc_ccd_step(self._ctx, nat.ptr(a, C.c_double), nat.ptr(b, C.c_double), C.byref(out)),
               self._ctx, "compute_ccd_step")
        return out.value

    def bvh_pairs(self, pos, margin: float) -> np.ndarray:
        """Every overlapping pair of triangle boxes (ta < tb), sorted, (k, 2) uint32."""
        n = self._n
        p = np.ascontiguousarray(np.asarray(pos, dtype=np.float64).reshape(3, n))
        cnt = C.c_uint64(0)
        _check(self._L.vh_bvh_pairs(self._ctx, nat.ptr(p, C.c_double), float(margin), None, None, 0, C.byref(cnt)),
               self._ctx, "bvh_pairs")
        k = int(cnt.value)
        ta, tb = np.empty(max(k, 1), dtype=np.uint32), np.empty(max(k, 1), dtype=np.uint32)
        _check(self._L.vh_bvh_pairs(self._ctx, nat.ptr(p, C.c_double), float(margin), nat.ptr(ta, C.c_uint32),
                                    nat.ptr(tb, C.c_uint32), k, C.byref(cnt)), self._ctx, "bvh_pairs")
        return np.stack([ta[:k], tb[:k]], 1)

    def count_crossings(self, pos=None) -> int:
        """Triangle pairs with no shared vertex that properly cross, at pos
        (3 x n) or at the resident positions: 0 for an intersection-free
        contact scene."""
        cnt = C.c_uint64(0)
        p = None
        if pos is not None:
            p = np.ascontiguousarray(np.asarray(pos, dtype=np.float64).reshape(3, self._n))
        _check(self._L.vh_ipc_crossings(self._ctx, None if p is None else nat.ptr(p, C.c_double), C.byref(cnt)),
               self._ctx, "count_crossings")
        return int(cnt.value)

    def set_lambda(self, xyz) -> None:
        a = np.ascontiguousarray(np.asarray(xyz, dtype=np.float64).reshape(3, self._n))
        _check(self._L.vh_state_lambda_set_f64(self._ctx, nat.ptr(a, C.c_double)), self._ctx, "set_lambda")

    def get_lambda(self) -> np.ndarray:
        a = np.empty((3, self._n))
        _check(self._L.vh_state_lambda_get_f64(self._ctx, nat.ptr(a, C.c_double)), self._ctx, "get_lambda")
        return a

    def set_state(self, which: int, xyz) -> None:
        a = np.ascontiguousarray(np.asarray(xyz, dtype=np.float64).reshape(3, self._n))
        _check(self._L.vh_state_set_f64(self._ctx, which, nat.ptr(a[0], C.c_double), nat.ptr(a[1], C.c_double),
                                        nat.ptr(a[2], C.c_double)), self._ctx, "set_state")

    def get_state(self, which: int = 0) -> np.ndarray:
        a = np.empty((3, self._n))
        _check(self._L.vh_state_get_f64(self._ctx, which, nat.ptr(a[0], C.c_double), nat.ptr(a[1], C.c_double),
                                        nat.ptr(a[2], C.c_double)), self._ctx, "get_state")
        return a

    def positions(self) -> np.ndarray:
        return self.get_state(0)

    def residual(self, dt: float) -> np.ndarray:
        out = np.empty(3 * self._n)
        _check(self._L.vh_pd_residual(self._ctx, float(dt), nat.ptr(out, C.c_double)), self._ctx, "residual")
        return out.reshape(3, self._n)

    def solve(self, b, sweeps: int = 3) -> np.ndarray:
        bb = np.ascontiguousarray(np.asarray(b, dtype=np.float64).reshape(3 * self._n))
        x = np.empty(3 * self._n)
        _check(self._L.vh_pd_solve(self._ctx, nat.ptr(bb, C.c_double), nat.ptr(x, C.c_double), int(sweeps)),
               self._ctx, "solve")
        return x.reshape(3, self._n)

    def profile(self, dt: float, iterations: int = 10) -> dict:
        """Kernel times (HIP events on the solver's stream) and algorithmic
        bytes per PD iteration, averaged over `iterations` iterations."""
        p = nat.Profile()
        _check(self._L.vh_pd_profile(self._ctx, float(dt), int(iterations), C.byref(p)), self._ctx, "profile")
        nb = p.bands
        return dict(bands=nb, iterations=p.iterations, sweeps=p.sweeps, launches=p.launches,
                    local_us=p.local_us, local_bytes=p.local_bytes,
                    fwd_us=list(p.fwd_us[:nb]), fwd_bytes=list(p.fwd_bytes[:nb]), bwd_us=list(p.bwd_us[:nb]),
                    bwd_bytes=list(p.bwd_bytes[:nb]), ir_us=p.ir_us, ir_bytes=p.ir_bytes,
                    iteration_us=p.iteration_us)

    def profile_graph(self, dt: float, reps: int = 50) -> dict:
        """vh_pd_profile_graph: one solve sweep and the local step, each
        captured into a graph and replayed `reps` times between HIP events."""
        g = nat.GraphProfile()
        _check(self._L.vh_pd_profile_graph(self._ctx, float(dt), int(reps), C.byref(g)), self._ctx, "profile_graph")
        return {k: getattr(g, k) for k, _ in nat.GraphProfile._fields_}

    def blocks(self, which: int) -> np.ndarray:
        """The device-built solve blocks as float64: 0 = M, 1 = M^T, 2 = the
        roots' dense inverses (vh_pd_blocks_get)."""
        cnt = C.c_uint64(0)
        _check(self._L.vh_pd_blocks_get(self._ctx, int(which), None, 0, C.byref(cnt)), self._ctx, "blocks")
        out = np.empty(cnt.value, dtype=np.float64)
        _check(self._L.vh_pd_blocks_get(self._ctx, int(which), nat.ptr(out, C.c_double), out.size, C.byref(cnt)),
               self._ctx, "blocks")
        return out

    def info(self) -> dict:
        i = nat.PdInfo()
        _check(self._L.vh_pd_info_get(self._ctx, C.byref(i)), self._ctx, "info")
        return {k: getattr(i, k) for k, _ in nat.PdInfo._fields_}
