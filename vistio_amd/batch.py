"""Many garments stepped as one system.

The meshes of several sheets are concatenated into one mesh: the system
matrix is then block diagonal, the nested dissection separates the
components for free (each sheet is placed beside the previous one, so the
first coordinate cuts fall between sheets), and one launch sequence, one
HIP graph, steps every sheet. Each sheet keeps its own material through the
per-element arrays of vh_pd_init_ex; each is what one
ProjectiveDynamicsSolver::init_with_material (pd_solver.rs:106-143) would
simulate on its own.

`batched_sheet_specs` draws the batched-sheets workload: 64 pinned sheets with grids
uniform in 32..256 cells a side, densities uniform in [0.08, 0.40] kg/m^2,
membrane weights log-uniform in [3e2, 3e3] N/m and bending weights
log-uniform in [1e-5, 1e-3] N m, from SplitMix64 streams seeded
0x5EED xor sheet id.
"""
from dataclasses import dataclass
from typing import List

import numpy as np

from .solver import Topology, TriangleMesh, quad_grid

_MASK = (1 << 64) - 1


def splitmix64(state: int):
    """One SplitMix64 step: returns (new state, 64-bit output)."""
    state = (state + 0x9E3779B97F4A7C15) & _MASK
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
    return state, z ^ (z >> 31)


def _uniform01(state: int):
    state, z = splitmix64(state)
    return state, (z >> 11) * (1.0 / (1 << 53))


@dataclass
class SheetSpec:
    cells: int          # quad_grid(cells, cells, size, size)
    density: float      # kg/m^2
    k_membrane: float   # N/m (element weight = k * area)
    k_bending: float    # N m (hinge weight = k * edge length)


def batched_sheet_specs(count: int = 64, seed: int = 0x5EED, min_cells: int = 32, max_cells: int = 256) -> List[SheetSpec]:
    specs = []
    for sid in range(count):
        st = (seed ^ sid) & _MASK
        st, u = _uniform01(st)
        cells = min_cells + min(int(u * (max_cells - min_cells + 1)), max_cells - min_cells)
        st, u = _uniform01(st)
        density = 0.08 + u * (0.40 - 0.08)
        st, u = _uniform01(st)
        k_mem = float(np.exp(np.log(3e2) + u * (np.log(3e3) - np.log(3e2))))
        st, u = _uniform01(st)
        k_b = float(np.exp(np.log(1e-5) + u * (np.log(1e-3) - np.log(1e-5))))
        specs.append(SheetSpec(cells, density, k_mem, k_b))
    return specs


@dataclass
class SheetBatch:
    mesh: TriangleMesh
    topology: Topology
    pinned: np.ndarray       # row 0 of every sheet
    k_membrane: np.ndarray   # per triangle
    k_bending: np.ndarray    # per interior edge
    density: np.ndarray      # per triangle
    vert_off: np.ndarray     # sheet s owns vertices vert_off[s] .. vert_off[s + 1]
    tri_off: np.ndarray
    hinge_off: np.ndarray
    specs: List[SheetSpec]

    def sheet_slice(self, s: int) -> slice:
        return slice(int(self.vert_off[s]), int(self.vert_off[s + 1]))


def build_batch(specs: List[SheetSpec], size: float = 1.0, gap: float = 0.5) -> SheetBatch:
    """Concatenates quad_grid(c, c, size, size) sheets, sheet s shifted by
    s * (size + gap) along x, each pinned along its row 0 (the hanging-sheet
    pose of the single-sheet workload)."""
    xs, ys, zs, tris, edges, pins = [], [], [], [], [], []
    km, kb, dens = [], [], []
    voff, toff, hoff = [0], [0], [0]
    for s, sp in enumerate(specs):
        m = quad_grid(sp.cells, sp.cells, size, size)
        topo = Topology.build(m)
        n, t = m.vertex_count(), m.triangle_count()
        xs.append(m.pos_x + np.float32(s * (size + gap)))
        ys.append(m.pos_y)
        zs.append(m.pos_z)
        tris.append(m.indices.astype(np.int64) + voff[-1])
        e = topo.interior_edges.astype(np.int64).copy()
        e[:, :4] += voff[-1]
        e[:, 4:] += toff[-1]
        edges.append(e)
        p = np.zeros(n, dtype=np.uint8)
        p[: sp.cells + 1] = 1
        pins.append(p)
        km.append(np.full(t, sp.k_membrane))
        dens.append(np.full(t, sp.density))
        kb.append(np.full(e.shape[0], sp.k_bending))
        voff.append(voff[-1] + n)
        toff.append(toff[-1] + t)
        hoff.append(hoff[-1] + e.shape[0])
    mesh = TriangleMesh(np.concatenate(xs).astype(np.float32), np.concatenate(ys).astype(np.float32),
                        np.concatenate(zs).astype(np.float32), np.concatenate(tris).astype(np.uint32))
    topo = Topology(np.concatenate(edges).astype(np.uint32))
    return SheetBatch(mesh, topo, np.concatenate(pins), np.concatenate(km), np.concatenate(kb), np.concatenate(dens),
                      np.array(voff), np.array(toff), np.array(hoff), list(specs))
