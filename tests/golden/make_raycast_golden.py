"""Writes tests/golden/raycast_small.npz from the reference's own code.

The expected outputs come from RISE's TriangleMeshGeometryIndexed BVH4 path
(float filter + double certifier), compiled from the reference sources by
oracle/ref/Makefile into oracle/_ref/libriseref_ieee.so.  Run in a checkout
that has the reference tree and the built oracle/_ref:

    python tests/golden/make_raycast_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from rise_amd import scenes  # noqa: E402

OUT = Path(__file__).resolve().parent / "raycast_small.npz"


def make_inputs():
    tris = scenes.soup_triangles(1500, seed=101, sigma=0.05)
    a = scenes.pinhole_rays(40, 30)
    b = scenes.incoherent_rays(1500, seed=102)
    # Axis-aligned directions (zero components) from inside the cube.
    g = np.random.Generator(np.random.PCG64(103))
    n = 96
    o = g.random((n, 3))
    d = np.zeros((n, 3))
    d[np.arange(n), np.arange(n) % 3] = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    c = scenes.rays_from_rows(o, d)
    rays = np.ascontiguousarray(np.concatenate([a, b, c], 1), dtype=np.float32)
    return tris, rays


def main() -> None:
    tris, rays = make_inputs()
    ref = oracle.RefMesh(tris, variant="ieee")
    pid, t, t64 = ref.closest(rays, threads=1)
    hit = pid >= 0
    # Shadow limits: just short of the hit for half of the hit rays, beyond it
    # for the other half, 0.25 for misses.
    half = (np.arange(rays.shape[1]) % 2) == 0
    tmax = np.where(hit, np.where(half, t * 0.999, t * 1.001), 0.25).astype(np.float32)
    occ = ref.any_hit(rays, tmax, threads=1)
    np.savez_compressed(OUT, tris=tris, rays=rays, prim_id=pid, t64=t64, tmax=tmax, occluded=occ)
    print(f"wrote {OUT}: {rays.shape[1]} rays, {hit.mean():.3f} hit, {occ.mean():.3f} occluded")


if __name__ == "__main__":
    main()
