"""State and mesh I/O on either side of the GPU step.

- StateSnapshot (crates/vistio-debug/src/snapshot.rs:14-71): timestep,
  simulation time, interleaved f32 positions and velocities. Serialised in
  bincode 1.x's default layout (little-endian, fixed-width integers, u64
  sequence lengths), the layout `bincode::serialize` writes, so snapshots
  taken from the GPU state can be read by the reference and vice versa.
  from_solver / apply_to_solver move the resident state of a HipPdSolver.
- JsonFrameExporter (crates/vistio-render/src/json_exporter.rs:39-104):
  per-frame positions for the reference's Three.js viewer.
- SimulationInput / SimulationParams / validate_input
  (crates/vistio-io/src/contract.rs:13-61, validator.rs:17-90): the run
  description as JSON, in serde's field layout.
"""
import json
import struct
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from .solver import InvalidConfig, InvalidMesh, TriangleMesh, VistioError

_HEAD = struct.Struct("<Id")   # timestep u32, sim_time f64
_LEN = struct.Struct("<Q")     # sequence length / usize


class SerializationError(VistioError):
    """VistioError::Serialization."""


@dataclass
class StateSnapshot:
    timestep: int
    sim_time: float
    positions: np.ndarray     # f32 [x0, y0, z0, x1, ...]
    velocities: np.ndarray    # f32, same layout
    vertex_count: int

    @staticmethod
    def from_soa(timestep, sim_time, pos_x, pos_y, pos_z, vel_x, vel_y, vel_z) -> "StateSnapshot":
        """snapshot.rs:28-59: SoA buffers interleaved per vertex."""
        pos = np.stack([np.asarray(a, dtype=np.float32) for a in (pos_x, pos_y, pos_z)], 1).reshape(-1)
        vel = np.stack([np.asarray(a, dtype=np.float32) for a in (vel_x, vel_y, vel_z)], 1).reshape(-1)
        return StateSnapshot(int(timestep), float(sim_time), pos, vel, int(len(pos_x)))

    @staticmethod
    def from_solver(solver, timestep: int, sim_time: float) -> "StateSnapshot":
        """The resident positions and velocities of a HipPdSolver, rounded
        to f32 as the reference's SimulationState holds them."""
        x, v = solver.get_state(0), solver.get_state(1)
        return StateSnapshot.from_soa(timestep, sim_time, x[0], x[1], x[2], v[0], v[1], v[2])

    def apply_to_solver(self, solver) -> None:
        """Restores positions and velocities (prev = positions) for replay."""
        x = self.positions.reshape(-1, 3).T.astype(np.float64)
        v = self.velocities.reshape(-1, 3).T.astype(np.float64)
        solver.set_state(0, x)
        solver.set_state(2, x)
        solver.set_state(1, v)

    def to_bytes(self) -> bytes:
        pos = np.ascontiguousarray(self.positions, dtype="<f4")
        vel = np.ascontiguousarray(self.velocities, dtype="<f4")
        return b"".join([_HEAD.pack(self.timestep, self.sim_time), _LEN.pack(pos.size), pos.tobytes(),
                         _LEN.pack(vel.size), vel.tobytes(), _LEN.pack(self.vertex_count)])

    @staticmethod
    def from_bytes(data: bytes) -> "StateSnapshot":
        try:
            off = 0
            ts, t = _HEAD.unpack_from(data, off)
            off += _HEAD.size
            arrays = []
            for _ in range(2):
                (k,) = _LEN.unpack_from(data, off)
                off += _LEN.size
                if off + 4 * k > len(data):
                    raise struct.error("sequence runs past the end of the data")
                arrays.append(np.frombuffer(data, dtype="<f4", count=k, offset=off).astype(np.float32))
                off += 4 * k
            (n,) = _LEN.unpack_from(data, off)
            off += _LEN.size
        except struct.error as e:
            raise SerializationError(f"Snapshot deserialization failed: {e}") from None
        return StateSnapshot(ts, t, arrays[0], arrays[1], int(n))


class JsonFrameExporter:
    """Renderer that collects frames and writes them as one JSON document."""

    def __init__(self, output_path: str):
        self.output_path = output_path
        self.indices: List[int] = []
        self.vertex_count = 0
        self.triangle_count = 0
        self.frames = []

    def name(self) -> str:
        return "json_exporter"

    def init(self, mesh: TriangleMesh) -> None:
        self.vertex_count = mesh.vertex_count()
        self.triangle_count = mesh.triangle_count()
        self.indices = [int(i) for i in np.asarray(mesh.indices).reshape(-1)]

    def submit_frame(self, timestep: int, pos_x, pos_y, pos_z) -> None:
        pos = np.stack([np.asarray(a, dtype=np.float32) for a in (pos_x, pos_y, pos_z)], 1).reshape(-1)
        self.frames.append((int(timestep), pos))

    def submit_solver(self, solver, timestep: int) -> None:
        x = solver.positions()
        self.submit_frame(timestep, x[0], x[1], x[2])

    def frame_count(self) -> int:
        return len(self.frames)

    def finalize(self) -> None:
        # f32 values written with the shortest repr that reads back to the same f32
        frames = [{"timestep": ts, "positions": [float(repr_f32) for repr_f32 in map(_f32_repr, pos)]}
                  for ts, pos in self.frames]
        data = {"vertex_count": self.vertex_count, "triangle_count": self.triangle_count,
                "indices": self.indices, "frames": frames}
        try:
            text = json.dumps(data, separators=(",", ":"))
        except (TypeError, ValueError) as e:
            raise SerializationError(f"JSON serialization failed: {e}") from None
        with open(self.output_path, "w") as f:
            f.write(text)
        self.frames = []


def _f32_repr(v: np.float32) -> str:
    return np.format_float_positional(np.float32(v), unique=True, trim="0")


@dataclass
class SimulationParams:
    """contract.rs:30-61 with the defaults of vistio_types::constants."""
    duration: float = 2.0
    dt: float = float(np.float32(1.0 / 60.0))
    iterations: int = 10
    gravity: float = float(np.float32(9.81))
    gravity_direction: List[float] = field(default_factory=lambda: [0.0, -1.0, 0.0])
    contact_thickness: float = float(np.float32(0.005))
    material_name: str = "cotton_twill"


def _mesh_dict(m: TriangleMesh) -> dict:
    """TriangleMesh in serde's layout (crates/vistio-mesh/src/mesh.rs:19-43);
    normals and UVs as quad_grid makes them when absent here."""
    n, t = m.vertex_count(), m.triangle_count()
    f = lambda a: [float(v) for v in np.asarray(a, dtype=np.float32)]
    extra = getattr(m, "_extra", None) or {}
    return {
        "pos_x": f(m.pos_x), "pos_y": f(m.pos_y), "pos_z": f(m.pos_z),
        "normal_x": extra.get("normal_x", [0.0] * n), "normal_y": extra.get("normal_y", [0.0] * n),
        "normal_z": extra.get("normal_z", [1.0] * n), "uv_u": extra.get("uv_u", [0.0] * n),
        "uv_v": extra.get("uv_v", [0.0] * n),
        "indices": [int(i) for i in np.asarray(m.indices).reshape(-1)],
        "material_ids": extra.get("material_ids", [0] * t),
    }


def _mesh_from(d: dict) -> TriangleMesh:
    m = TriangleMesh(np.asarray(d["pos_x"], dtype=np.float32), np.asarray(d["pos_y"], dtype=np.float32),
                     np.asarray(d["pos_z"], dtype=np.float32), np.asarray(d["indices"], dtype=np.uint32))
    m._extra = {k: d[k] for k in ("normal_x", "normal_y", "normal_z", "uv_u", "uv_v", "material_ids") if k in d}
    return m


@dataclass
class SimulationInput:
    garment: TriangleMesh
    body: Optional[TriangleMesh]
    pinned: List[bool]
    params: SimulationParams = field(default_factory=SimulationParams)

    def to_json(self) -> str:
        return json.dumps({
            "garment": _mesh_dict(self.garment),
            "body": None if self.body is None else _mesh_dict(self.body),
            "pinned": [bool(p) for p in self.pinned],
            "params": dict(self.params.__dict__),
        })

    @staticmethod
    def from_json(text: str) -> "SimulationInput":
        try:
            d = json.loads(text)
            return SimulationInput(_mesh_from(d["garment"]), None if d["body"] is None else _mesh_from(d["body"]),
                                   list(d["pinned"]), SimulationParams(**d["params"]))
        except (KeyError, TypeError, ValueError) as e:
            raise SerializationError(f"SimulationInput deserialization failed: {e}") from None


def _validate_mesh(m: TriangleMesh, what: str) -> None:
    """TriangleMesh::validate (mesh.rs:118-170) for the fields held here."""
    n = m.vertex_count()
    if np.asarray(m.pos_y).size != n or np.asarray(m.pos_z).size != n:
        raise InvalidMesh(f"{what}: Position arrays have inconsistent lengths")
    idx = np.asarray(m.indices).reshape(-1)
    if idx.size % 3 != 0:
        raise InvalidMesh(f"{what}: Index count is not divisible by 3")
    bad = np.nonzero(idx >= n)[0]
    if bad.size:
        raise InvalidMesh(f"{what}: Index {int(idx[bad[0]])} at position {int(bad[0])} is out of range "
                          f"(vertex count: {n})")


def validate_input(inp: SimulationInput) -> None:
    """validator.rs:17-90."""
    _validate_mesh(inp.garment, "Garment mesh")
    if inp.body is not None:
        _validate_mesh(inp.body, "Body mesh")
    if len(inp.pinned) != inp.garment.vertex_count():
        raise InvalidMesh(f"Pinned array length ({len(inp.pinned)}) != garment vertex count "
                          f"({inp.garment.vertex_count()})")
    p = inp.params
    if p.dt <= 0.0:
        raise InvalidConfig("Timestep dt must be positive")
    if p.dt > 1.0:
        raise InvalidConfig("Timestep dt > 1.0 is unreasonably large")
    if p.duration <= 0.0:
        raise InvalidConfig("Duration must be positive")
    if p.iterations == 0:
        raise InvalidConfig("Solver iterations must be >= 1")
    if p.gravity < 0.0:
        raise InvalidConfig("Gravity magnitude must be non-negative")
    if p.contact_thickness < 0.0:
        raise InvalidConfig("Contact thickness must be non-negative")
    gd = np.asarray(p.gravity_direction, dtype=np.float32)
    ln = float(np.sqrt((gd * gd).sum()))
    if abs(ln - 1.0) > 0.01:
        raise InvalidConfig(f"Gravity direction should be unit length, got magnitude {ln}")
