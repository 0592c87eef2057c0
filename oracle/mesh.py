"""Mesh generators and topology for the oracle (numpy restatement).

TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the CPU-baseline
leg of bench.py to build oracle inputs. The product has its own
implementation (vistio_amd/csrc/vh_mesh.cpp).

* ``quad_grid``  follows crates/vistio-mesh/src/generators.rs:28-100 with the
  same float32 arithmetic (u = i/cols in f32, x = -w/2 + u*w in f32).
* ``interior_edges`` follows crates/vistio-mesh/src/topology.rs:58-120 and
  :158-168, except that interior edges come out sorted by (v0, v1): the
  reference iterates a randomly seeded HashMap, so its order is not
  reproducible. Within an edge, tri_a < tri_b (triangles are pushed in
  index order, topology.rs:85-91) and the wings are found per triangle.
"""
from __future__ import annotations

import numpy as np


def quad_grid(cols: int, rows: int, width: float, height: float):
    """Return (px, py, pz, indices) as float32 / uint32 arrays."""
    f = np.float32
    vx, vy = cols + 1, rows + 1
    i = np.arange(vx, dtype=np.int64)
    j = np.arange(vy, dtype=np.int64)
    u = (i.astype(f) / f(cols)).astype(f)
    v = (j.astype(f) / f(rows)).astype(f)
    half_w, half_h = f(f(width) / f(2.0)), f(f(height) / f(2.0))
    xs = (-half_w + (u * f(width)).astype(f)).astype(f)
    ys = (half_h - (v * f(height)).astype(f)).astype(f)
    px = np.tile(xs, vy)
    py = np.repeat(ys, vx)
    pz = np.zeros(vx * vy, dtype=f)
    jj, ii = np.meshgrid(np.arange(rows), np.arange(cols), indexing="ij")
    tl = (jj * vx + ii).ravel().astype(np.int64)
    tr, bl = tl + 1, tl + vx
    br = bl + 1
    even = ((ii + jj) % 2 == 0).ravel()
    tri = np.empty((tl.size, 2, 3), dtype=np.int64)
    tri[even, 0] = np.stack([tl, bl, tr], 1)[even]
    tri[even, 1] = np.stack([tr, bl, br], 1)[even]
    tri[~even, 0] = np.stack([tl, bl, br], 1)[~even]
    tri[~even, 1] = np.stack([tl, br, tr], 1)[~even]
    return px, py, pz, tri.reshape(-1).astype(np.uint32)


def interior_edges(indices: np.ndarray, n_vertices: int) -> np.ndarray:
    """Interior edges as an (H, 6) uint32 array: v0, v1, wing_a, wing_b, tri_a, tri_b."""
    tri = np.asarray(indices, dtype=np.int64).reshape(-1, 3)
    t = tri.shape[0]
    a, b, c = tri[:, 0], tri[:, 1], tri[:, 2]
    e0 = np.concatenate([a, b, c])
    e1 = np.concatenate([b, c, a])
    wing = np.concatenate([c, a, b])
    tid = np.concatenate([np.arange(t)] * 3)
    lo, hi = np.minimum(e0, e1), np.maximum(e0, e1)
    key = lo * n_vertices + hi
    order = np.lexsort((tid, key))
    key, lo, hi, wing, tid = key[order], lo[order], hi[order], wing[order], tid[order]
    starts = np.flatnonzero(np.r_[True, key[1:] != key[:-1]])
    counts = np.diff(np.r_[starts, key.size])
    s = starts[counts == 2]
    return np.stack([lo[s], hi[s], wing[s], wing[s + 1], tid[s], tid[s + 1]], 1).astype(np.uint32)


def hinges4(indices: np.ndarray, n_vertices: int) -> np.ndarray:
    """(H, 4) array of v0, v1, wing_a, wing_b in the canonical order."""
    return np.ascontiguousarray(interior_edges(indices, n_vertices)[:, :4])


def pinned_top_row(cols: int, rows: int) -> np.ndarray:
    """Pins the first cols+1 vertices (row 0), as crates/vistio-bench/src/scenarios.rs:97-100."""
    p = np.zeros((cols + 1) * (rows + 1), dtype=np.uint8)
    p[: cols + 1] = 1
    return p
