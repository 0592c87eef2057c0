"""Schedule autotuner: picks the solve schedule of a mesh by timing it.

The set-up's band schedule comes from a cost model (launch time and block
bytes per band; vh_pd_params.band_launch_us, band_gbps, max_band_levels) and
its work partition from task_entries (block entries per solve task). The
model's constants are the MI355X's measured ones for one sheet; for other
sizes and mixes of garments this tunes them on the device: a coordinate
search, one knob at a time from the defaults, each candidate built and timed
over a few resident steps. The result feeds SolverConfig.

Large batches can be tuned on a proxy (a subset of their garments drawn the
same way): the knobs describe the hardware and the size mix, not one mesh.
"""
import time
from dataclasses import replace
from typing import Dict, List, Tuple

from .solver import HipPdSolver, SolverConfig

KNOBS = (
    ("max_band_levels", (2, 3, 4)),
    ("band_gbps", (2000.0, 4000.0, 8000.0)),
    ("band_launch_us", (3.0, 6.0, 12.0)),
    ("task_entries", (512, 1024, 2048)),
)
DEFAULTS = {"max_band_levels": 3, "band_gbps": 4000.0, "band_launch_us": 6.0, "task_entries": 1024}


def time_config(mesh, topology, config: SolverConfig, pinned, model, k_membrane, k_bending, density,
                dt: float = 1.0 / 60.0, steps: int = 10, warmup: int = 2) -> Tuple[float, dict]:
    """Resident steps/s of one configuration (set-up excluded), and its info."""
    s = HipPdSolver()
    try:
        s.init_raw(mesh, topology, config, pinned, model, k_membrane, k_bending, density)
        s.step_resident(dt, warmup)
        s.step_resident(dt, 0)
        t0 = time.perf_counter()
        s.step_resident(dt, steps)
        wall = time.perf_counter() - t0
        return steps / wall, s.info()
    finally:
        s.close()


def autotune(mesh, topology, config: SolverConfig, pinned, model: str, k_membrane, k_bending, density,
             dt: float = 1.0 / 60.0, steps: int = 10, log=None) -> Tuple[SolverConfig, List[Dict]]:
    """Returns (the fastest SolverConfig found, every timed candidate)."""
    best = dict(DEFAULTS)
    table: List[Dict] = []
    seen: Dict[tuple, float] = {}

    def score(knobs):
        key = tuple(sorted(knobs.items()))
        if key not in seen:
            cfg = replace(config, **knobs)
            rate, info = time_config(mesh, topology, cfg, pinned, model, k_membrane, k_bending, density, dt, steps)
            seen[key] = rate
            row = dict(knobs, steps_per_s=round(rate, 3), bands=int(info["levels"]),
                       factor_entries=int(info["nnz_factor"]))
            table.append(row)
            if log:
                log(row)
        return seen[key]

    best_rate = score(best)
    for name, values in KNOBS:
        for v in values:
            if v == best[name]:
                continue
            trial = dict(best, **{name: v})
            try:
                rate = score(trial)
            except Exception as e:  # a schedule the set-up rejects (front over the LDS budget)
                table.append(dict(trial, error=str(e)))
                continue
            if rate > best_rate:
                best, best_rate = trial, rate
    return replace(config, **best), table
