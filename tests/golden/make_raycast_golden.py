"""Writes tests/golden/raycast_small.npz and raycast_filter.npz from the
reference's own code.

The expected outputs come from RISE's TriangleMeshGeometryIndexed BVH4 path
(float filter + double certifier), compiled from the reference sources by
oracle/ref/Makefile into oracle/_ref/libriseref_ieee.so (every operation
rounded) and libriseref.so (the reference's Linux flags, which contract
products into fused multiply-adds).  raycast_filter.npz holds rays aimed
within 3e-5 (barycentric) of sliver-triangle edges, where the float filter
decides hits the double certifier alone would accept, with the outputs of
both builds.  Run in a checkout that has the reference tree and the built
oracle/_ref:

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
OUT_FILTER = Path(__file__).resolve().parent / "raycast_filter.npz"


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


def make_filter_inputs():
    tris = scenes.soup_triangles(1000, seed=7, sigma=0.003)
    rays = np.concatenate([scenes.near_edge_rays(tris, 12000, 8), scenes.incoherent_rays(2000, 9)], 1)
    return tris, np.ascontiguousarray(rays, dtype=np.float32)


def main_filter() -> None:
    tris, rays = make_filter_inputs()
    out = {"tris": tris, "rays": rays}
    for v in ("ieee", "fast"):
        ref = oracle.RefMesh(tris, variant=v)
        pid, t, t64 = ref.closest(rays, threads=1)
        hit = pid >= 0
        half = (np.arange(rays.shape[1]) % 2) == 0
        tmax = np.where(hit, np.where(half, t * 0.999, t * 1.001), 0.25).astype(np.float32)
        out[f"prim_id_{v}"] = pid
        out[f"t64_{v}"] = t64
        out[f"tmax_{v}"] = tmax
        out[f"occluded_{v}"] = ref.any_hit(rays, tmax, threads=1)
        ref.close()
    np.savez_compressed(OUT_FILTER, **out)
    print(f"wrote {OUT_FILTER}: {rays.shape[1]} rays; ieee and fast builds differ on "
          f"{int((out['prim_id_ieee'] != out['prim_id_fast']).sum())} ids")


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
    main_filter()
