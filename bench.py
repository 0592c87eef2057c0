#!/usr/bin/env python3
"""PD steps/s of the 256x256 hanging sheet on one MI355X: FP64 state, an FP32
local step in residual form and FP32 supernodal solves of the correction
(each PD iteration re-forms r = b - A x from the FP64 state, so the PD loop
itself refines; extra sweeps per iteration refine with the FP64 CSR
residual), next to the FP64 GPU path and the CPU restatement of the
reference step.

Workload: quad_grid(256, 256, 1 m, 1 m) in the z = 0
plane (66,049 vertices, 131,072 triangles, 196,096 hinges), row 0 pinned,
gravity -9.81 m/s^2 along y, 0.15 kg/m^2, co-rotational membrane weight
1e3 N/m, dihedral bending weight 1e-4 N m, Rayleigh 0.01, dt = 1/60 s,
10 PD iterations per step (tolerance 0), synthetic mesh (no dataset).

Timing: W warm-up steps, the state is reset to rest, then exactly K steps
are timed from a device synchronisation to the next (state resident in
HBM), as replays of the step's HIP graph. The roofline figures time the
same kernels the same way: one solve sweep and the local step, each
captured into a graph and replayed (vh_pd_profile_graph).

Accuracy: the max per-vertex distance of the FP32-refined and of the FP64
GPU positions from the FP64 oracle after the same K steps from rest (the
oracle runs after the timed region, on all host threads; its results do not
depend on the thread count), and between the two GPU paths.
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "PD steps/s at 256² sheet, FP32-refined vs FP64 GPU, at ≤1e-6 m err"
HBM_PEAK_GBS = 8000.0


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def make_sheet(cells):
    from vistio_amd.scenes import hanging_sheet
    return hanging_sheet(cells)


def make_solver(mesh, topo, pinned, precision, ir_sweeps, ir_final_sweeps):
    import vistio_amd as va
    cfg = va.SolverConfig(max_iterations=10, tolerance=0.0, gravity=(0.0, -9.81, 0.0), damping=0.0,
                          rayleigh_mass_damping=0.01, precision=precision, ir_sweeps=ir_sweeps,
                          ir_final_sweeps=ir_final_sweeps)
    s = va.HipPdSolver(0)
    t0 = time.perf_counter()
    s.init_raw(mesh, topo, cfg, pinned, "co_rotational", 1e3, 1e-4, 0.15)
    return s, time.perf_counter() - t0


def reset(s, mesh):
    rest = np.stack([mesh.pos_x, mesh.pos_y, mesh.pos_z]).astype(np.float64)
    s.set_state(0, rest)
    s.set_state(2, rest)
    s.set_state(1, np.zeros_like(rest))


def run_path(mesh, topo, pinned, precision, steps, warmup, dt, ir_sweeps=1, ir_final_sweeps=3):
    s, setup = make_solver(mesh, topo, pinned, precision, ir_sweeps, ir_final_sweeps)
    if warmup:
        s.step_resident(dt, warmup)
    reset(s, mesh)
    s.step_resident(dt, 0)  # device synchronisation before the timed region
    t0 = time.perf_counter()
    r = s.step_resident(dt, steps)  # ends with a device synchronisation
    wall = time.perf_counter() - t0
    x = s.positions()
    graph = s.profile_graph(dt, 100)
    prof = s.profile(dt, 10)
    info = s.info()
    s.close()
    return dict(wall=wall, x=x, prof=prof, graph=graph, info=info, setup=setup, iterations=r.iterations,
                final_residual=r.final_residual)


def run_drop_in(mesh, topo, pinned, steps, dt, ir_sweeps, ir_final_sweeps):
    """The drop-in entry point (vh_pd_step): the f32 host SimulationState is
    copied to the device and back around every step, as SolverStrategy::step
    hands it over. Reported beside `value`, which keeps the state in HBM."""
    import vistio_amd as va
    s, _ = make_solver(mesh, topo, pinned, "f32_ir", ir_sweeps, ir_final_sweeps)
    state = va.SimulationState.from_mesh_with_masses(mesh, s.lumped_masses(), pinned)
    for _ in range(2):
        s.step(state, dt)
    t0 = time.perf_counter()
    for _ in range(steps):
        s.step(state, dt)
    wall = time.perf_counter() - t0
    s.close()
    return steps / wall


def roofline_of(graph, ms_per_step, sweeps_per_step, traffic=None):
    """Dominant kernels: the band launches of the supernodal solve (forward
    bands, the root band, backward bands; 2 * bands - 1 launches a sweep),
    timed as the steps run them: one sweep captured into a HIP graph and
    replayed (vh_pd_profile_graph), HIP events on the solver's stream around
    the replays. achieved = algorithmic bytes per launch / average launch
    time. The band launches of one step cannot take longer than the step:
    launch_avg_us x launches_per_step <= ms_per_step is checked and reported."""
    launches = graph["solve_launches"]
    us = graph["solve_us"] / launches
    by = graph["solve_bytes"] / launches
    achieved = by / (us * 1e-6) / 1e9
    per_step = launches * sweeps_per_step
    band_ms = us * per_step / 1000.0
    if band_ms > ms_per_step:
        print(f"roofline check failed: {per_step} band launches x {us:.3f} us = {band_ms:.4f} ms > "
              f"{ms_per_step:.4f} ms per step", file=sys.stderr, flush=True)
    return dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic,
                kernel="k_fwd_band/k_root_band/k_bwd_band (supernodal sweeps)",
                launch_avg_us=round(us, 3), bytes_per_launch=round(by), launches_per_step=per_step,
                band_ms_per_step=round(band_ms, 4), fits_in_step=bool(band_ms <= ms_per_step),
                timing="graph replay of one solve sweep, HIP events")


def sweeps_per_step(args, precision):
    """Solve sweeps in one step: one per PD iteration in FP64; in FP32-IR
    ir_sweeps per iteration and ir_final_sweeps on the last."""
    if precision == "f64":
        return 10
    return 9 * args.ir_sweeps + args.ir_final_sweeps


def local_step_of(graph):
    """The local step (k_local) timed in graph replay: algorithmic bytes per
    launch over the average launch time."""
    return dict(us=round(graph["local_us"], 3), bytes=round(graph["local_bytes"]),
                gbs=round(graph["local_bytes"] / (graph["local_us"] * 1e-6) / 1e9, 1),
                frac=round(graph["local_bytes"] / (graph["local_us"] * 1e-6) / 1e9 / HBM_PEAK_GBS, 4))


def oracle_errors(cells, steps, dt, positions):
    """Max per-vertex distance of each GPU result from the FP64 oracle after
    the same number of steps from rest (the oracle, run after the timed
    region on all host threads)."""
    from oracle import mesh as omesh
    from oracle import ref
    ref.lib()
    threads = min(16, int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1))
    ref.set_threads(threads)
    px, py, pz, idx = omesh.quad_grid(cells, cells, 1.0, 1.0)
    o = ref.Oracle(ref.F64, px, py, pz, idx, omesh.hinges4(idx, px.size), omesh.pinned_top_row(cells, cells),
                   ref.params())
    t0 = time.perf_counter()
    o.step(dt, steps)
    wall = time.perf_counter() - t0
    ref.set_threads(1)
    xo = o.positions()
    out = {k: float(np.sqrt(((x - xo) ** 2).sum(0)).max()) for k, x in positions.items()}
    out["oracle_seconds"] = round(wall, 1)
    out["oracle_threads"] = threads
    return out


def copy_probe(mb):
    """A kernel with known HBM traffic (reads mb MB, writes mb MB), run three
    times, so that counter totals can be scaled to bytes."""
    import torch
    n = mb * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        torch.add(a, 1.0, out=b)
    torch.cuda.synchronize()
    del a, b


def launch_floor_probe(nodes=200, reps=20):
    """The launch floor of a graph replay: microseconds per node of a captured
    HIP graph whose nodes are dependent one-element kernels (an in-place add
    each), replayed reps times between HIP events. Every band launch of the
    solve pays at least this. The tensor's final value is checked against the
    count of adds, so a replay that skipped or repeated nodes would show."""
    import torch
    x = torch.zeros(1, dtype=torch.float64, device="cuda")
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(3):
            x.add_(1.0)
    torch.cuda.current_stream().wait_stream(side)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        for _ in range(nodes):
            x.add_(1.0)
    g.replay()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        g.replay()
    e1.record()
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) * 1000.0 / (reps * nodes)
    # three eager adds, then nodes per replay (the capture itself runs nothing)
    ok = float(x.item()) == float(3 + nodes * (reps + 1))
    return dict(us_per_node=round(us, 3), nodes=nodes, replays=reps, value_check=ok)


def traffic_from_profiles(prec):
    """HBM bytes per band launch measured with rocprofv3 FETCH_SIZE and
    WRITE_SIZE (profiles/band_kernels_256_traffic.md), or None."""
    path = os.path.join(ROOT, "profiles", "band_kernels_256_traffic.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(f"traffic_bytes_per_launch_{prec}:"):
                    return int(float(line.split(":", 1)[1]))
    except OSError:
        pass
    return None


def cpu_baseline(cells, sample_steps, dt, threads=1):
    """The oracle in the reference's precision split (f32 state, f64 solve,
    three single-RHS solves per iteration) on a bounded sample. threads = 1
    is the reference's single thread; more threads split the element loops,
    the three RHS axes and the three solves (same results)."""
    from oracle import mesh as omesh
    from oracle import ref
    ref.lib()
    ref.set_threads(threads)
    px, py, pz, idx = omesh.quad_grid(cells, cells, 1.0, 1.0)
    hg = omesh.hinges4(idx, px.size)
    pin = omesh.pinned_top_row(cells, cells)
    o = ref.Oracle(ref.F32_REF, px, py, pz, idx, hg, pin, ref.params())
    t0 = time.perf_counter()
    o.step(dt, sample_steps)
    wall = time.perf_counter() - t0
    ref.set_threads(1)
    how = "one thread" if threads == 1 else f"{threads} threads (elements, RHS axes and solves in parallel)"
    return dict(value=round(sample_steps / wall, 4), unit="steps/s", cores=threads, kind="port",
                sample=f"{sample_steps} steps of the {cells}x{cells} hanging sheet (10 PD iterations each, "
                       f"f32 state + f64 LL^T with three single-RHS solves, as the reference), {how}, "
                       f"{wall:.1f} s; CPU: {cpu_model()} ({os.cpu_count()} logical CPUs on the host)")


def run_contact(args):
    """Steps/s of the IPC step (state resident in HBM). Every step is one
    call that ends with a device synchronisation; the intersection check
    between steps is outside the timed intervals."""
    import vistio_amd as va
    dt = 1.0 / 60.0
    from vistio_amd.scenes import contact_scene
    mesh, ipc, steps, desc = contact_scene(args.workload)
    steps = args.steps if args.steps_given else steps
    topo = va.Topology.build(mesh)
    pinned = np.zeros(mesh.vertex_count(), dtype=np.uint8)
    res = {}
    for prec in ("f32_ir", "f64"):
        cfg = va.SolverConfig(max_iterations=10, tolerance=0.0, gravity=(0.0, -9.81, 0.0), damping=0.0,
                              rayleigh_mass_damping=0.01, precision=prec, ir_sweeps=args.ir_sweeps,
                              ir_final_sweeps=args.ir_final_sweeps)
        s = va.HipPdSolver(0)
        t0 = time.perf_counter()
        s.init_raw(mesh, topo, cfg, pinned, "co_rotational", 1e3, 1e-4, 0.15)
        s.init_ipc(ipc)
        setup = time.perf_counter() - t0
        rest = np.stack([mesh.pos_x, mesh.pos_y, mesh.pos_z]).astype(np.float64)
        s.step_ipc_resident(dt, max(args.warmup, 1))
        s.set_state(0, rest); s.set_state(2, rest); s.set_state(1, np.zeros_like(rest))
        s.set_lambda(np.zeros_like(rest))
        wall, al, most_self, most_act, cross_steps, worst = 0.0, 0, 0, 0, 0, 0
        check = prec == "f32_ir" or args.check_all
        for k in range(steps):
            t0 = time.perf_counter()
            r, st = s.step_ipc_resident(dt, 1)
            wall += time.perf_counter() - t0
            al += st["al_iterations"]
            most_self = max(most_self, st["self_contacts"])
            most_act = max(most_act, st["active"])
            if check and (k % args.check_every == 0 or k == steps - 1):
                c = s.count_crossings()
                cross_steps += c > 0
                worst = max(worst, c)
            if k % 30 == 29:
                print(f"[{args.workload} {prec}] step {k + 1}/{steps}: {1000 * wall / (k + 1):.2f} ms/step, "
                      f"AL {st['al_iterations']}, contacts {st['active']} (self {st['self_contacts']})",
                      file=sys.stderr, flush=True)
        x = s.positions()
        res[prec] = dict(wall=wall, x=x, al=al / steps, most_self=most_self, most_active=most_act,
                         crossing_steps=cross_steps if check else None, worst_crossings=worst if check else None,
                         setup=setup, last=st)
        s.close()
    f32, f64 = res["f32_ir"], res["f64"]
    value = steps / f32["wall"]
    x = f32["x"]
    out = {
        "metric": f"IPC steps/s ({args.workload})",
        "value": round(value, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * f32["wall"] / steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": desc + f"; {mesh.vertex_count()} verts, {mesh.triangle_count()} tris; co-rotational "
                               "1e3 N/m, bending 1e-4 N m, 0.15 kg/m^2, Rayleigh 0.01, dt 1/60 s, 10 PD "
                               "iterations per AL iteration; FP64 contact terms and CCD, FP32-refined solve"},
        "extra": {
            "fp64_gpu_steps_per_s": round(steps / f64["wall"], 3),
            "al_iterations_per_step": {"fp32_ir": round(f32["al"], 3), "fp64": round(f64["al"], 3)},
            "most_active_contacts": f32["most_active"],
            "most_self_contacts": f32["most_self"],
            "crossing_check_every": args.check_every,
            "steps_with_crossings": f32["crossing_steps"],
            "most_crossing_pairs": f32["worst_crossings"],
            "y_min": float(x[1].min()),
            "fp32_ir_vs_fp64_gpu_max_vertex_m": float(np.sqrt(((f32["x"] - f64["x"]) ** 2).sum(0)).max()),
            "setup_s": round(f32["setup"], 2),
        },
    }
    if args.workload == "sphere_drape":
        out["extra"]["min_dist_to_sphere_minus_r"] = float(np.sqrt((x ** 2).sum(0)).min() - 0.5)
    print(json.dumps(out))


def run_batch(args):
    """The batched-sheets workload: 64 pinned sheets (grids 32..256 cells a
    side, per-sheet density and stiffness; vistio_amd.batch.batched_sheet_specs, seeds
    SplitMix64(0x5EED ^ sheet id)) stepped as one block-diagonal system in one
    graph-captured launch sequence. value = batch steps/s (every sheet
    advances one step); the error is the max vertex distance between the
    FP32-refined and the FP64 GPU batch after the timed steps from rest."""
    import vistio_amd as va
    from vistio_amd.batch import build_batch, batched_sheet_specs
    dt = 1.0 / 60.0
    specs = batched_sheet_specs(args.sheets)
    b = build_batch(specs)
    steps = args.steps
    res, tuned = {}, {}
    for prec in ("f32_ir", "f64"):
        cfg = va.SolverConfig(max_iterations=10, tolerance=0.0, gravity=(0.0, -9.81, 0.0), damping=0.0,
                              rayleigh_mass_damping=0.01, precision=prec, ir_sweeps=args.ir_sweeps,
                              ir_final_sweeps=args.ir_final_sweeps)
        if args.autotune_sheets > 0:
            # schedule knobs tuned on a proxy: the first sheets of the same draw
            from vistio_amd.autotune import autotune
            pb = build_batch(specs[: args.autotune_sheets])
            t0 = time.perf_counter()
            cfg, table = autotune(pb.mesh, pb.topology, cfg, pb.pinned, "co_rotational", pb.k_membrane,
                                  pb.k_bending, pb.density, dt, steps=10,
                                  log=lambda row: print(f"[autotune {prec}] {row}", file=sys.stderr, flush=True))
            tuned[prec] = dict(proxy_sheets=args.autotune_sheets, seconds=round(time.perf_counter() - t0, 1),
                               chosen={k: getattr(cfg, k) for k in ("max_band_levels", "band_gbps",
                                                                    "band_launch_us", "task_entries")},
                               candidates=table)
        s = va.HipPdSolver(0)
        t0 = time.perf_counter()
        s.init_raw(b.mesh, b.topology, cfg, b.pinned, "co_rotational", b.k_membrane, b.k_bending, b.density)
        setup = time.perf_counter() - t0
        print(f"[batched_sheets {prec}] set-up {setup:.1f} s", file=sys.stderr, flush=True)
        if args.warmup:
            s.step_resident(dt, args.warmup)
        reset(s, b.mesh)
        s.step_resident(dt, 0)
        t0 = time.perf_counter()
        s.step_resident(dt, steps)
        wall = time.perf_counter() - t0
        res[prec] = dict(wall=wall, x=s.positions(), graph=s.profile_graph(dt, 20), prof=s.profile(dt, 10),
                         info=s.info(), setup=setup)
        s.close()
    f32, f64 = res["f32_ir"], res["f64"]
    value = steps / f32["wall"]
    err = float(np.sqrt(((f32["x"] - f64["x"]) ** 2).sum(0)).max())
    cells = np.array([sp.cells for sp in specs])
    out = {
        "metric": "batch steps/s (batched sheets)",
        "value": round(value, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * f32["wall"] / steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"batched_sheets: {len(specs)} pinned sheets, grids {cells.min()}..{cells.max()} "
                               f"cells a side (mean {cells.mean():.0f}), {b.mesh.vertex_count()} verts, "
                               f"{b.mesh.triangle_count()} tris, per-sheet density 0.08-0.40 kg/m^2, membrane "
                               "3e2-3e3 N/m, bending 1e-5-1e-3 N m (SplitMix64 seeds 0x5EED ^ id), co-rotational, "
                               "Rayleigh 0.01, dt 1/60 s, 10 PD iterations; FP64 state, FP32 residual-form local "
                               "step + one FP32 solve per PD iteration; one block-diagonal system, one graph per "
                               "step",
                   "sheets": len(specs)},
        "roofline": roofline_of(f32["graph"], 1000.0 * f32["wall"] / steps, sweeps_per_step(args, "f32_ir")),
        "extra": {
            "sheet_steps_per_s": round(value * len(specs), 1),
            "vertex_steps_per_s": round(value * b.mesh.vertex_count(), 0),
            "max_vertex_err_vs_fp64_gpu_m": err,
            "fp64_gpu_steps_per_s": round(steps / f64["wall"], 3),
            "fp32_ir_over_fp64_gpu": round(value / (steps / f64["wall"]), 3),
            "iteration_us": {"fp32_ir": round(f32["prof"]["iteration_us"], 2),
                             "fp64": round(f64["prof"]["iteration_us"], 2)},
            "local_step": local_step_of(f32["graph"]),
            "solve_bands": f32["info"]["levels"],
            "factor_entries": f32["info"]["nnz_factor"],
            "setup_s": {"fp32_ir": round(f32["setup"], 1), "fp64": round(f64["setup"], 1),
                        "host_symbolic_fp32_ir": round(f32["info"]["setup_seconds"], 2),
                        "device_factor_fp32_ir": round(f32["info"]["factor_seconds"], 3)},
            "autotune": tuned or None,
        },
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=120)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cells", type=int, default=256)
    ap.add_argument("--ir-sweeps", type=int, default=1)
    ap.add_argument("--ir-final-sweeps", type=int, default=1)
    ap.add_argument("--cpu-sample-steps", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="first run a plain copy-like kernel (b = a + 1, float32) over this many MB, "
                         "to calibrate rocprofv3 FETCH_SIZE / WRITE_SIZE")
    ap.add_argument("--workload", default="hanging_sheet",
                    choices=("hanging_sheet", "sphere_drape", "self_fold_ribbon", "batched_sheets"),
                    help="hanging_sheet is the headline metric; the contact workloads print their own line")
    ap.add_argument("--check-every", type=int, default=1,
                    help="contact workloads: count crossing triangle pairs every this many steps")
    ap.add_argument("--check-all", action="store_true", help="contact workloads: also check the FP64 run")
    ap.add_argument("--sheets", type=int, default=64, help="batched_sheets: number of sheets")
    ap.add_argument("--autotune-sheets", type=int, default=8,
                    help="batched_sheets: tune the solve schedule on this many sheets of the draw first (0: off)")
    args = ap.parse_args()
    args.steps_given = any(a == "--steps" or a.startswith("--steps=") for a in sys.argv[1:])
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU: the PD step of one sheet runs on a single MI355X; "
              f"--gpus {args.gpus} (WORLD_SIZE {world}) is not supported", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe_mb > 0:
        copy_probe(args.copy_probe_mb)
    # torch's HIP runtime is initialised before the library's, as in copy_probe
    floor = launch_floor_probe() if args.workload == "hanging_sheet" else None
    if args.workload == "batched_sheets":
        run_batch(args)
        return
    if args.workload != "hanging_sheet":
        run_contact(args)
        return
    dt = 1.0 / 60.0
    mesh, topo, pinned = make_sheet(args.cells)
    f64 = run_path(mesh, topo, pinned, "f64", args.steps, args.warmup, dt)
    f32 = run_path(mesh, topo, pinned, "f32_ir", args.steps, args.warmup, dt, args.ir_sweeps, args.ir_final_sweeps)
    err = float(np.sqrt(((f32["x"] - f64["x"]) ** 2).sum(0)).max())
    drop_in = run_drop_in(mesh, topo, pinned, min(args.steps, 60), dt, args.ir_sweeps, args.ir_final_sweeps)
    value = args.steps / f32["wall"]
    v64 = args.steps / f64["wall"]
    ms32 = 1000.0 * f32["wall"] / args.steps
    ms64 = 1000.0 * f64["wall"] / args.steps
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * f32["wall"] / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"hanging_sheet {args.cells}x{args.cells} cells ({mesh.vertex_count()} verts, "
                        f"{mesh.triangle_count()} tris), co-rotational membrane 1e3 N/m + dihedral bending "
                        f"1e-4 N m, 0.15 kg/m^2, Rayleigh 0.01, dt 1/60 s, 10 PD iterations/step; FP64 state; "
                        f"each PD iteration forms r = b - A x from the FP64 state (element terms in FP32 "
                        f"residual form after FP64 F, inertia in FP64) and solves A d = r with FP32 supernodal "
                        f"sweeps ({args.ir_sweeps}/iteration, {args.ir_final_sweeps} on the last; further "
                        f"sweeps refine with the FP64 CSR residual)",
            "cells": args.cells,
            "pd_iterations_per_step": 10,
        },
        "roofline": roofline_of(f32["graph"], ms32, sweeps_per_step(args, "f32_ir"),
                                traffic_from_profiles("fp32") if args.cells == 256 else None),
        "extra": {
            "pd_iterations_per_s": round(10 * value, 1),
            "max_vertex_err_vs_fp64_gpu_m": err,
            "fp64_gpu_steps_per_s": round(v64, 3),
            "fp32_ir_over_fp64_gpu": round(value / v64, 3),
            "fp64_roofline": roofline_of(f64["graph"], ms64, sweeps_per_step(args, "f64"),
                                         traffic_from_profiles("fp64") if args.cells == 256 else None),
            "local_step": {"fp32": local_step_of(f32["graph"]), "fp64": local_step_of(f64["graph"])},
            "event_profile_us": {"fp32_ir_iteration": round(f32["prof"]["iteration_us"], 2),
                                 "fp64_iteration": round(f64["prof"]["iteration_us"], 2),
                                 "note": "one event after every launch, no graph (vh_pd_profile)"},
            "solve_bands": f32["info"]["levels"],
            "factor_entries": f32["info"]["nnz_factor"],
            "setup_s": {"fp32_ir": round(f32["setup"], 2), "fp64": round(f64["setup"], 2),
                        "host_symbolic": round(f32["info"]["setup_seconds"], 3),
                        "device_factor_fp32_ir": round(f32["info"]["factor_seconds"], 4),
                        "device_factor_fp64": round(f64["info"]["factor_seconds"], 4)},
            "final_residual": f32["final_residual"],
            "drop_in_host_state_steps_per_s": round(drop_in, 3),
            "launch_floor": floor,
        },
    }
    if not args.no_cpu:
        # the FP64 oracle over the same steps from rest
        oe = oracle_errors(args.cells, args.steps, dt, {"fp32_ir": f32["x"], "fp64": f64["x"]})
        out["extra"]["max_vertex_err_vs_fp64_oracle_m"] = oe
        cb = cpu_baseline(args.cells, args.cpu_sample_steps, dt)
        out["cpu_baseline"] = cb
        out["extra"]["fp32_ir_over_cpu"] = round(value / cb["value"], 1)
        threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
        if threads > 1:
            ca = cpu_baseline(args.cells, args.cpu_sample_steps, dt, threads)
            out["extra"]["cpu_all_cores"] = ca
            out["extra"]["fp32_ir_over_cpu_all_cores"] = round(value / ca["value"], 1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
