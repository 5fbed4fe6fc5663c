"""The kernels' traversal and certifier code (rise_amd/csrc/trace_core.h),
compiled for the host and run through rise_hip_debug_trace_host, against the
oracle: bit-exact ids and distances, on the edge cases the reference's tests
cover (tests/BVHBuilderTest.cpp: random workloads, deep clustered trees;
tests/RayBoxIntersectionTest.cpp: axis-parallel and grazing rays).  This
checks the 8-wide build, the outward quantization and the conservative slab
test before any GPU run."""
from pathlib import Path

import numpy as np
import pytest

import oracle
from rise_amd import raycast, scenes

GOLDEN_FILTER = Path(__file__).resolve().parent / "golden" / "raycast_filter.npz"


def _check(tris, rays, tmax=None):
    pid, t, t64 = oracle.closest(tris, rays)
    hp, ht, cnt, st = raycast.host_trace(tris, rays)
    bad = np.nonzero(hp != pid)[0]
    assert bad.size == 0, f"{bad.size} id mismatches, first {bad[:5]}: host {hp[bad[:5]]} oracle {pid[bad[:5]]}"
    hit = pid >= 0
    assert np.array_equal(ht[hit], t[hit])
    assert np.all(np.isinf(ht[~hit]))
    if tmax is None:
        tmax = np.where(hit, np.where(np.arange(len(pid)) % 2 == 0, t * 0.999, t * 1.001), 0.5).astype(np.float32)
    occ = oracle.any_hit(tris, rays, tmax)
    hocc, _, _ = raycast.host_trace(tris, rays, 1, tmax)
    assert np.array_equal(occ, hocc)
    return st, cnt


@pytest.mark.parametrize("seed,n,sigma", [(1, 30000, 0.003), (2, 10000, 0.03), (3, 3000, 0.2)])
def test_random_soup(seed, n, sigma):
    tris = scenes.soup_triangles(n, seed=seed, sigma=sigma)
    rays = np.concatenate([scenes.pinhole_rays(80, 60), scenes.incoherent_rays(5000, seed + 50)], 1)
    st, cnt = _check(tris, rays)
    assert st["n_tris"] == n and st["depth"] >= 2


def test_empty_scene_and_no_rays():
    tris = np.zeros((0, 9), np.float32)
    rays = scenes.incoherent_rays(100, 1)
    hp, ht, _, st = raycast.host_trace(tris, rays)
    assert np.all(hp == -1) and np.all(np.isinf(ht))
    hp, ht, _, _ = raycast.host_trace(scenes.soup_triangles(10, 1), np.zeros((6, 0), np.float32))
    assert hp.size == 0


def test_single_triangle_root_leaf():
    tris = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], np.float32)
    g = np.random.Generator(np.random.PCG64(4))
    o = np.column_stack([g.uniform(-0.2, 1.2, 2000), g.uniform(-0.2, 1.2, 2000), np.full(2000, 1.0)])
    d = np.tile([0, 0, -1.0], (2000, 1))
    _check(tris, scenes.rays_from_rows(o, d))


def test_axis_parallel_rays_and_flat_geometry():
    # A closed box of axis-aligned quads (two triangles each, shared edges)
    # and rays with zero direction components, some starting on a face plane.
    q = []
    for axis in range(3):
        for c in (0.0, 1.0):
            u, v = [a for a in range(3) if a != axis]
            p = np.zeros((4, 3))
            p[:, axis] = c
            p[:, u] = [0, 1, 1, 0]
            p[:, v] = [0, 0, 1, 1]
            q.append(np.concatenate([p[0], p[1], p[2]]))
            q.append(np.concatenate([p[0], p[2], p[3]]))
    tris = np.array(q, np.float32)
    g = np.random.Generator(np.random.PCG64(5))
    n = 3000
    o = g.random((n, 3))
    o[::7, 0] = 0.0  # on the x = 0 plane
    d = np.zeros((n, 3))
    d[np.arange(n), np.arange(n) % 3] = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    d[1::5] = g.normal(size=(len(d[1::5]), 3))
    # Rays through the shared diagonals (exact-t ties between two triangles).
    o[2::11, 0] = o[2::11, 1]
    d[2::11] = [0, 0, 1.0]
    _check(tris, scenes.rays_from_rows(o, d))


def test_duplicate_triangles_tie_to_lowest_id():
    base = scenes.soup_triangles(500, seed=8, sigma=0.05)
    tris = np.concatenate([base, base, base])  # ids i, i+500, i+1000 coincide
    rays = scenes.incoherent_rays(4000, 9)
    pid, _, _ = oracle.closest(tris, rays)
    assert np.all(pid[pid >= 0] < 500)
    _check(tris, rays)


def test_deep_clustered_tree():
    # 4096 tiny triangles clustered at geometric scales (a deep, skewed tree;
    # cf. BVHBuilderTest.cpp:308-379).
    g = np.random.Generator(np.random.PCG64(10))
    n = 4096
    scale = 0.5 ** (np.arange(n) % 40)
    c = 0.5 + (g.random((n, 3)) - 0.5) * scale[:, None]
    e1 = g.normal(0, 1e-3, (n, 3)) * scale[:, None]
    e2 = g.normal(0, 1e-3, (n, 3)) * scale[:, None]
    v0 = c - (e1 + e2) / 3
    tris = np.concatenate([v0, v0 + e1, v0 + e2], 1).astype(np.float32)
    o = np.full((3000, 3), 0.5) + g.normal(0, 0.3, (3000, 3))
    d = 0.5 - o + g.normal(0, 1e-4, (3000, 3))
    st, _ = _check(tris, scenes.rays_from_rows(o, d))
    assert st["depth"] > 6


def test_large_coordinates_and_tiny_triangles():
    g = np.random.Generator(np.random.PCG64(12))
    n = 5000
    c = 1000.0 + g.random((n, 3)) * 10
    e1 = g.normal(0, 1e-3, (n, 3))
    e2 = g.normal(0, 1e-3, (n, 3))
    tris = np.concatenate([c, c + e1, c + e2], 1).astype(np.float32)
    o = 1000.0 + g.random((4000, 3)) * 10
    d = g.normal(size=(4000, 3))
    _check(tris, scenes.rays_from_rows(o, d))


def test_degenerate_triangles_never_hit():
    tris = scenes.soup_triangles(1000, seed=13, sigma=0.05)
    tris[::3, 3:6] = tris[::3, 0:3]  # zero-area: v1 == v0
    rays = scenes.incoherent_rays(3000, 14)
    pid, _, _ = oracle.closest(tris, rays)
    assert not np.any(pid[pid >= 0] % 3 == 0)
    _check(tris, rays)


def test_unnormalized_directions_and_shadow_scaling():
    tris = scenes.soup_triangles(3000, seed=15, sigma=0.05)
    rays = scenes.incoherent_rays(3000, 16)
    rays[3:6] *= np.linspace(0.01, 50.0, 3000, dtype=np.float32)
    _check(tris, rays)


def test_nonfinite_vertices_rejected():
    tris = scenes.soup_triangles(10, 1)
    tris[3, 4] = np.nan
    with pytest.raises(Exception, match="non-finite"):
        raycast.host_trace(tris, scenes.incoherent_rays(4, 1))


@pytest.mark.parametrize("arith", ["ieee", "fast"])
def test_float_filter_fixture(arith):
    """Rays within 3e-5 of sliver-triangle edges: the host-compiled traversal
    gives the ids, distances and occlusion of the reference build with the
    same arithmetic (tests/golden/raycast_filter.npz)."""
    z = np.load(GOLDEN_FILTER)
    a = raycast.ARITH_X86_FMA if arith == "fast" else raycast.ARITH_IEEE
    hp, ht, _, _ = raycast.host_trace(z["tris"], z["rays"], arith=a)
    assert np.array_equal(hp, z[f"prim_id_{arith}"])
    hit = hp >= 0
    assert np.array_equal(ht[hit], z[f"t64_{arith}"][hit].astype(np.float32))
    occ, _, _ = raycast.host_trace(z["tris"], z["rays"], 1, z[f"tmax_{arith}"], arith=a)
    assert np.array_equal(occ, z[f"occluded_{arith}"])
