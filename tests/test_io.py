"""State and mesh I/O (vistio_amd.io): the reference's own tests
(crates/vistio-debug/tests/debug_tests.rs:38-68,
crates/vistio-io/tests/io_tests.rs), the bincode 1.x byte layout of a
snapshot written out field by field, the JSON frame document, and (GPU) a
replay from a snapshot of the resident state."""
import json
import struct

import numpy as np
import pytest


def test_snapshot_round_trip_and_interleaving():
    from vistio_amd.io import StateSnapshot
    snap = StateSnapshot.from_soa(42, 0.7, [1.0, 2.0, 3.0], [4.0, 5.0, 6.0], [7.0, 8.0, 9.0],
                                  [0.1, 0.2, 0.3], [0.4, 0.5, 0.6], [0.7, 0.8, 0.9])   # debug_tests.rs:38-56
    rec = StateSnapshot.from_bytes(snap.to_bytes())
    assert rec.timestep == 42 and rec.vertex_count == 3 and abs(rec.sim_time - 0.7) < 1e-10
    assert rec.positions.size == 9 and rec.velocities.size == 9
    snap = StateSnapshot.from_soa(0, 0.0, [1.0, 2.0], [3.0, 4.0], [5.0, 6.0], [0.0] * 2, [0.0] * 2, [0.0] * 2)
    assert snap.positions.tolist() == [1.0, 3.0, 5.0, 2.0, 4.0, 6.0]                 # :58-67


def test_snapshot_bincode_layout():
    """bincode 1.x default encoding of StateSnapshot: u32, f64, (u64 len,
    f32 LE...) twice, u64 vertex_count."""
    from vistio_amd.io import SerializationError, StateSnapshot
    snap = StateSnapshot.from_soa(7, 0.25, [1.5], [2.5], [-3.0], [0.0], [1.0], [2.0])
    b = snap.to_bytes()
    want = (struct.pack("<I", 7) + struct.pack("<d", 0.25) + struct.pack("<Q", 3) + struct.pack("<3f", 1.5, 2.5, -3.0)
            + struct.pack("<Q", 3) + struct.pack("<3f", 0.0, 1.0, 2.0) + struct.pack("<Q", 1))
    assert b == want and len(b) == 4 + 8 + 8 + 12 + 8 + 12 + 8
    for cut in (0, 3, 11, 19, 25, len(b) - 1):
        with pytest.raises(SerializationError):
            StateSnapshot.from_bytes(b[:cut])


def test_json_frame_exporter(tmp_path):
    import vistio_amd as va
    from vistio_amd.io import JsonFrameExporter
    m = va.quad_grid(2, 2, 1.0, 1.0)
    ex = JsonFrameExporter(str(tmp_path / "out.json"))
    ex.init(m)
    assert ex.name() == "json_exporter"
    for k in range(3):
        ex.submit_frame(k, m.pos_x, m.pos_y + np.float32(0.1 * k), m.pos_z)
    assert ex.frame_count() == 3
    ex.finalize()
    d = json.loads((tmp_path / "out.json").read_text())
    assert list(d) == ["vertex_count", "triangle_count", "indices", "frames"]
    assert d["vertex_count"] == 9 and d["triangle_count"] == 8 and d["indices"] == m.indices.tolist()
    assert [f["timestep"] for f in d["frames"]] == [0, 1, 2]
    p = np.asarray(d["frames"][2]["positions"], dtype=np.float32).reshape(-1, 3)
    assert np.array_equal(p[:, 1], m.pos_y + np.float32(0.2))   # every f32 reads back exactly


def test_simulation_input_kats():
    import vistio_amd as va
    from vistio_amd.io import SimulationInput, SimulationParams, validate_input
    p = SimulationParams()
    assert abs(p.gravity - 9.81) < 1e-3 and p.iterations == 10                      # io_tests.rs:9-14
    m = va.quad_grid(2, 2, 1.0, 1.0)
    inp = SimulationInput(m, None, [False] * m.vertex_count())
    rec = SimulationInput.from_json(inp.to_json())                                    # :16-30
    assert rec.garment.vertex_count() == 9 and np.array_equal(rec.garment.indices, m.indices)
    d = json.loads(inp.to_json())
    assert list(d) == ["garment", "body", "pinned", "params"]
    assert list(d["params"]) == ["duration", "dt", "iterations", "gravity", "gravity_direction",
                                 "contact_thickness", "material_name"]
    m4 = va.quad_grid(4, 4, 1.0, 1.0)
    good = lambda: SimulationInput(m4, None, [False] * m4.vertex_count())
    validate_input(good())                                                            # :44-47
    bad = good(); bad.pinned = [False] * 3
    with pytest.raises(va.InvalidMesh):                                               # :49-54
        validate_input(bad)
    for field, value in (("dt", -0.01), ("iterations", 0), ("gravity_direction", [0.0, -2.0, 0.0])):
        bad = good()
        setattr(bad.params, field, value)
        with pytest.raises(va.InvalidConfig):                                         # :56-75
            validate_input(bad)
    inp = good()
    inp.body = va.quad_grid(2, 2, 0.5, 0.5)
    validate_input(inp)                                                               # :77-82


@pytest.mark.gpu
def test_snapshot_replay_of_resident_state():
    """A snapshot of the GPU state after 20 steps, restored into a fresh
    solver, continues like the original run up to the f32 rounding of the
    snapshot."""
    from conftest import DT, gpu_solver, max_vertex_err, sheet
    from vistio_amd.io import StateSnapshot
    px, py, pz, idx, hg, pin = sheet(20, rotate_xz=True)
    a = gpu_solver(px, py, pz, idx, hg, pin)
    a.step_resident(DT, 20)
    snap = StateSnapshot.from_bytes(StateSnapshot.from_solver(a, 20, 20 * DT).to_bytes())
    assert snap.vertex_count == px.size and snap.timestep == 20
    a.step_resident(DT, 20)
    b = gpu_solver(px, py, pz, idx, hg, pin)
    snap.apply_to_solver(b)
    b.step_resident(DT, 20)
    err = max_vertex_err(a.positions(), b.positions())
    print("replay vs continued run:", err, flush=True)
    assert err < 1e-5
    a.close(); b.close()


@pytest.mark.gpu
def test_snapshot_of_gpu_state_matches_oracle_snapshot():
    """StateSnapshot (snapshot.rs:14-71) of the GPU state after 30 steps and
    of the FP64 oracle's state after the same steps: the bincode bytes decode
    to the same f32 positions and velocities, to one f32 rounding (the two
    FP64 states are within 1e-9 m). A snapshot written from the oracle and
    restored into the GPU solver then continues like the oracle restored
    from the same snapshot."""
    from conftest import DT, gpu_solver, max_vertex_err, oracle_solver, sheet
    from vistio_amd.io import StateSnapshot
    px, py, pz, idx, hg, pin = sheet(20, rotate_xz=True)
    g = gpu_solver(px, py, pz, idx, hg, pin)
    o = oracle_solver(px, py, pz, idx, hg, pin)
    g.step_resident(DT, 30)
    o.step(DT, 30)
    sg = StateSnapshot.from_bytes(StateSnapshot.from_solver(g, 30, 30 * DT).to_bytes())
    xo, vo = o.get(0), o.get(1)
    so = StateSnapshot.from_bytes(StateSnapshot.from_soa(30, 30 * DT, *xo, *vo).to_bytes())
    assert (sg.timestep, sg.vertex_count) == (so.timestep, so.vertex_count) == (30, px.size)
    # the bytes hold exactly the f32 rounding of the GPU's own FP64 state
    xg, vg = g.get_state(0), g.get_state(1)
    assert np.array_equal(sg.positions, xg.T.reshape(-1).astype(np.float32))
    assert np.array_equal(sg.velocities, vg.T.reshape(-1).astype(np.float32))
    # and agree with the oracle's state within the FP64 parity bound (1e-9 m,
    # so 1e-9 / dt = 6e-8 m/s for velocities) plus one f32 rounding
    ulp_p = np.spacing(np.abs(so.positions).astype(np.float32)).astype(np.float64)
    ulp_v = np.spacing(np.abs(so.velocities).astype(np.float32)).astype(np.float64)
    dp = np.abs(sg.positions.astype(np.float64) - xo.T.reshape(-1))
    dv = np.abs(sg.velocities.astype(np.float64) - vo.T.reshape(-1))
    print(f"snapshot vs oracle state: positions max {dp.max():.2e} m, velocities max {dv.max():.2e} m/s", flush=True)
    assert np.all(dp <= 1e-9 + ulp_p) and np.all(dv <= 1e-9 / DT + ulp_v)
    # replay: the oracle's snapshot restored into both
    so.apply_to_solver(g)
    x = so.positions.reshape(-1, 3).T.astype(np.float64)
    v = so.velocities.reshape(-1, 3).T.astype(np.float64)
    o.set(0, x); o.set(2, x); o.set(1, v)
    g.step_resident(DT, 20)
    o.step(DT, 20)
    err = max_vertex_err(g.positions(), o.positions())
    print(f"replayed from the oracle's snapshot: max vertex err after 20 steps {err:.3e} m", flush=True)
    assert err < 1e-9
    g.close()
