#!/usr/bin/env python3
"""Writes the golden fixtures of this directory from the oracle (oracle/).

Fixtures (all at the scale of the 20 x 20 sheet, small, numpy .npz):

  sheet20_steps.npz  the 20 x 20 hanging sheet (z = 0 curtain) and the
                     reference's rotated pose (XZ plane at y = 1): inputs
                     (positions, triangles, hinges, pins) and the positions
                     after steps 1, 10 and 120 (rotated: 1, 10, 60) in FP64
                     and in the reference's precision split (f32 state, f64
                     solve).
  sheet20_local.npz  the local step at one perturbed state of the rotated
                     sheet (seed 7): per-triangle and per-hinge targets, the
                     RHS and the residual rhs - A x.
  folded_pairs.npz   the candidate triangle pairs of a folded 8 x 8 patch
                     (16 x 8 strip folded in half, 2 cm gap) at margins 0,
                     0.01 and 0.03.
  ccd_table.npz      vertex-triangle CCD of 64 seeded swept configurations
                     (about half of them hits): the TOI per case (2 = none)
                     and the whole-scene step fraction of each case as a
                     two-triangle mesh.

Run from the repository root:  python tests/golden/make_golden.py
The fixtures pin the oracle against drift (tests/test_golden.py) and let the
GPU tests compare against stored values.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

DT = 1.0 / 60.0


def sheet(c, rotate_xz):
    from oracle import mesh
    px, py, pz, idx = mesh.quad_grid(c, c, 1.0, 1.0)
    if rotate_xz:  # scenarios.rs:91-96
        pz = py.copy()
        py = np.full_like(px, np.float32(1.0))
    return px, py, pz, idx, mesh.hinges4(idx, px.size), mesh.pinned_top_row(c, c)


def oracle(mode, px, py, pz, idx, hg, pin):
    from oracle import ref
    return ref.Oracle(mode, px, py, pz, idx, hg, pin, ref.params())


def steps_fixture():
    from oracle import ref
    out = {}
    for name, rot, marks in (("curtain", False, (1, 10, 120)), ("rotated", True, (1, 10, 60))):
        px, py, pz, idx, hg, pin = sheet(20, rot)
        out[f"{name}_pos"] = np.stack([px, py, pz])
        out[f"{name}_tri"] = idx
        out[f"{name}_hinges"] = hg
        out[f"{name}_pinned"] = pin
        out[f"{name}_marks"] = np.array(marks)
        for tag, mode in (("f64", ref.F64), ("f32ref", ref.F32_REF)):
            o = oracle(mode, px, py, pz, idx, hg, pin)
            done = 0
            for m in marks:
                o.step(DT, m - done)
                done = m
                out[f"{name}_{tag}_x{m}"] = o.positions()
    np.savez_compressed(os.path.join(HERE, "sheet20_steps.npz"), **out)


def perturbed_state(px, py, pz, seed, amp=2e-3):
    rng = np.random.default_rng(seed)
    x = np.stack([px, py, pz]).astype(np.float64)
    x = x + amp * rng.standard_normal(x.shape) * np.array([[1.0], [1.0], [0.5]])
    pred = x + 1e-3 * rng.standard_normal(x.shape)
    return x, pred


def local_fixture():
    from oracle import ref
    px, py, pz, idx, hg, pin = sheet(20, True)
    o = oracle(ref.F64, px, py, pz, idx, hg, pin)
    x, pred = perturbed_state(px, py, pz, 7)
    o.set(0, x)
    o.set(3, pred)
    rhs, tt, ht = o.local_rhs()
    A = o.matrix()
    resid = np.stack([rhs[k] - A @ x[k] for k in range(3)])
    np.savez_compressed(os.path.join(HERE, "sheet20_local.npz"), x=x, pred=pred, tri_targets=tt,
                        hinge_targets=ht, rhs=rhs, residual=resid)


def folded_patch(c=8, gap=0.02, seed=0):
    from oracle import mesh
    px, py, pz, idx = mesh.quad_grid(2 * c, c, 1.0, 0.5)
    x = px.astype(np.float64)
    y = np.zeros_like(x)
    z = py.astype(np.float64)
    up = x > 0.0
    x[up] = -x[up]
    y[up] = gap
    rng = np.random.default_rng(seed)
    y = y + rng.normal(size=y.size) * gap * 0.05
    return np.stack([x, y, z]).astype(np.float32), idx


def pairs_fixture():
    from oracle import ref
    pos, idx = folded_patch()
    out = dict(pos=pos, tri=idx, margins=np.array([0.0, 0.01, 0.03], dtype=np.float32))
    for k, m in enumerate(out["margins"]):
        out[f"pairs{k}"] = ref.tri_pairs(pos[0], pos[1], pos[2], idx, m, brute=True)
    np.savez_compressed(os.path.join(HERE, "folded_pairs.npz"), **out)


def ccd_fixture():
    from oracle import ref
    rng = np.random.default_rng(11)
    cases, toi, alpha = [], [], []
    tri = np.array([0, 1, 2, 3, 4, 5], dtype=np.uint32)
    q = ref.ipc_params(self_collision=True, thickness=0.5)
    for k in range(64):
        a0 = np.array([0.0, 0.0, 0.0]); b0 = np.array([1.0, 0.0, 0.0]); c0 = np.array([0.0, 1.0, 0.0])
        tmove = rng.normal(size=3) * 0.05
        p0 = np.array([rng.uniform(0.05, 0.6), rng.uniform(0.05, 0.6), rng.uniform(0.05, 0.3)])
        # half the cases cross the plane inside the triangle, half pass beside or stop short
        p1 = p0 + np.array([rng.normal() * 0.1, rng.normal() * 0.1, -rng.uniform(0.0, 0.6)])
        if k % 2:
            p1[:2] += rng.uniform(0.6, 1.2, size=2)
        P0 = np.stack([a0, b0, c0, p0, p0 + [0.3, 0.0, 0.2], p0 + [0.0, 0.3, 0.2]])
        P1 = np.stack([a0 + tmove, b0 + tmove, c0 + tmove, p1, p1 + [0.3, 0.0, 0.2], p1 + [0.0, 0.3, 0.2]])
        t = ref.vt_ccd(P0[3], P1[3], P0[0], P1[0], P0[1], P1[1], P0[2], P1[2])
        cases.append(np.stack([P0, P1]))
        toi.append(2.0 if t is None else t)
        alpha.append(ref.ccd_step(P0.T, P1.T, tri, q))
    np.savez_compressed(os.path.join(HERE, "ccd_table.npz"), cases=np.array(cases), toi=np.array(toi),
                        alpha=np.array(alpha), tri=tri)


def main():
    from oracle import ref
    ref.lib()
    steps_fixture()
    local_fixture()
    pairs_fixture()
    ccd_fixture()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")


if __name__ == "__main__":
    main()
