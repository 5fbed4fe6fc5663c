"""GPU parity for the ray-cast path, through the C ABI (ctypes).

* The golden fixture produced by the reference's own BVH4 path.
* The oracle (brute force, FP64) on seeded soups and edge cases.
* The host-compiled traversal: identical ids, distances and node/triangle
  counts, which also pins the GPU's double-precision divide and square root.
* Rays within 3e-5 of sliver-triangle edges, where the reference's float
  pre-filter decides the hit (tests/golden/raycast_filter.npz, written by
  the reference's two builds), in both arithmetic models.
* The full ray-cast workload (1M triangles, 6,073,600 rays): every ray
  against the reference's BVH4 (oracle/_ref) when built, in both arithmetic
  models, a brute-force sample, and size-independent shadow-ray properties.
Bar: ids bit-exact; t bit-exact to the oracle's float32 world distance
(north-star tolerance 1e-5 relative is asserted as well).
"""
from pathlib import Path

import numpy as np
import pytest

import oracle
from rise_amd import raycast, scenes
from tests.conftest import log_case

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden" / "raycast_small.npz"
GOLDEN_FILTER = Path(__file__).resolve().parent / "golden" / "raycast_filter.npz"
ARITH = {"ieee": raycast.ARITH_IEEE, "fast": raycast.ARITH_X86_FMA}


def _gpu(cuda, tris, rays, tmax=None, counted=False, arith="ieee"):
    import torch

    om = raycast.ObjectManager(ARITH[arith])
    om.add_mesh(tris)
    om.prepare_for_rendering()
    r = torch.from_numpy(np.ascontiguousarray(rays, np.float32)).to(cuda)
    cnt = torch.zeros(2, dtype=torch.int64, device=cuda) if counted else None
    pid, t = om.intersect_ray(r, counters=cnt)
    out = {"pid": pid.cpu().numpy(), "t": t.cpu().numpy(), "stats": om.stats()}
    if counted:
        out["counters"] = cnt.cpu().numpy().astype(np.uint64)
    if tmax is not None:
        tm = torch.from_numpy(np.ascontiguousarray(tmax, np.float32)).to(cuda)
        out["occ"] = om.intersect_shadow_ray(r, tm).cpu().numpy()
    torch.cuda.synchronize()
    om.release()
    return out


def _shadow_limits(pid, t):
    hit = pid >= 0
    return np.where(hit, np.where(np.arange(len(pid)) % 2 == 0, t * 0.999, t * 1.001), 0.5).astype(np.float32)


def test_gpu_golden_fixture(cuda):
    log_case("golden fixture")
    z = np.load(GOLDEN)
    g = _gpu(cuda, z["tris"], z["rays"], z["tmax"])
    assert np.array_equal(g["pid"], z["prim_id"])
    hit = z["prim_id"] >= 0
    assert np.array_equal(g["t"][hit], z["t64"][hit].astype(np.float32))
    assert np.array_equal(g["occ"], z["occluded"])


@pytest.mark.parametrize("arith", ["ieee", "fast"])
def test_gpu_float_filter_fixture(cuda, arith):
    log_case(f"float-filter fixture, {arith} arithmetic")
    z = np.load(GOLDEN_FILTER)
    g = _gpu(cuda, z["tris"], z["rays"], z[f"tmax_{arith}"], arith=arith)
    assert np.array_equal(g["pid"], z[f"prim_id_{arith}"])
    hit = g["pid"] >= 0
    assert np.array_equal(g["t"][hit], z[f"t64_{arith}"][hit].astype(np.float32))
    assert np.array_equal(g["occ"], z[f"occluded_{arith}"])


@pytest.mark.parametrize("seed,n,sigma", [(1, 30000, 0.003), (2, 10000, 0.03), (3, 3000, 0.2)])
def test_gpu_matches_oracle_and_host(cuda, seed, n, sigma):
    log_case(f"soup seed={seed} n={n}")
    tris = scenes.soup_triangles(n, seed=seed, sigma=sigma)
    rays = np.concatenate([scenes.pinhole_rays(80, 60), scenes.incoherent_rays(5000, seed + 50)], 1)
    pid, t, _ = oracle.closest(tris, rays)
    tmax = _shadow_limits(pid, t)
    g = _gpu(cuda, tris, rays, tmax, counted=True)
    assert np.array_equal(g["pid"], pid)
    hit = pid >= 0
    assert np.array_equal(g["t"][hit], t[hit])
    assert np.all(np.isinf(g["t"][~hit]))
    assert np.array_equal(g["occ"], oracle.any_hit(tris, rays, tmax))
    hp, ht, hcnt, _ = raycast.host_trace(tris, rays)
    assert np.array_equal(g["counters"], hcnt)


def test_gpu_edge_cases(cuda):
    log_case("edge cases: box, axis rays, ties, duplicates")
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
    box = np.array(q, np.float32)
    dup = scenes.soup_triangles(300, seed=31, sigma=0.05)
    tris = np.concatenate([box, dup, dup], 0)
    g = np.random.Generator(np.random.PCG64(32))
    n = 4000
    o = g.random((n, 3))
    o[::7, 0] = 0.0
    d = np.zeros((n, 3))
    d[np.arange(n), np.arange(n) % 3] = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    d[1::5] = g.normal(size=(len(d[1::5]), 3))
    o[2::11, 0] = o[2::11, 1]
    d[2::11] = [0, 0, 1.0]
    rays = scenes.rays_from_rows(o, d)
    pid, t, _ = oracle.closest(tris, rays)
    tmax = _shadow_limits(pid, t)
    gg = _gpu(cuda, tris, rays, tmax)
    assert np.array_equal(gg["pid"], pid)
    assert np.array_equal(gg["t"][pid >= 0], t[pid >= 0])
    assert np.array_equal(gg["occ"], oracle.any_hit(tris, rays, tmax))


def test_gpu_empty_inputs(cuda):
    import torch

    log_case("empty scene / zero rays")
    g = _gpu(cuda, np.zeros((0, 9), np.float32), scenes.incoherent_rays(1000, 3))
    assert np.all(g["pid"] == -1) and np.all(np.isinf(g["t"]))
    om = raycast.ObjectManager()
    om.add_mesh(scenes.soup_triangles(100, 1))
    om.prepare_for_rendering()
    pid, t = om.intersect_ray(torch.zeros((6, 0), dtype=torch.float32, device=cuda))
    assert pid.numel() == 0
    om.release()


def test_gpu_full_ray_cast_workload(cuda):
    """1M triangles, 2,073,600 coherent + 4,000,000 incoherent rays."""
    log_case("full ray-cast workload 1M tris / 6,073,600 rays")
    tris = scenes.soup_triangles()
    rays = scenes.cast_rays()
    n = rays.shape[1]
    assert n == 6_073_600
    g = _gpu(cuda, tris, rays)
    pid, t = g["pid"], g["t"]
    assert g["stats"]["n_tris"] == 1_000_000
    hit = pid >= 0
    assert np.all((pid[hit] >= 0) & (pid[hit] < 1_000_000))
    assert np.all(np.isfinite(t[hit])) and np.all(np.isinf(t[~hit]))
    # Brute-force sample (every 20000th ray) against the oracle.
    idx = np.arange(0, n, 20000)
    sp, st, _ = oracle.closest(tris, rays[:, idx])
    assert np.array_equal(pid[idx], sp)
    assert np.array_equal(t[idx][sp >= 0], st[sp >= 0])
    # Every ray against the reference's own BVH4 path when it is built, in
    # both of its builds' arithmetic: ids bit-exact, distances within 1e-5.
    if oracle.ref_available("fast"):
        for arith in ("ieee", "fast"):
            log_case(f"full workload vs reference BVH4 (oracle/_ref), {arith} arithmetic")
            ga = g if arith == "ieee" else _gpu(cuda, tris, rays, arith=arith)
            ref = oracle.RefMesh(tris, arith)
            rp, rt, rt64 = ref.closest(rays, threads=16)
            ref.close()
            bad = np.nonzero(rp != ga["pid"])[0]
            assert bad.size == 0, f"{bad.size} id mismatches vs the reference ({arith}), first {bad[:5]}"
            ok = rp >= 0
            rel = np.abs(ga["t"][ok].astype(np.float64) - rt64[ok]) / rt64[ok]
            assert rel.max() <= 1e-5
    # Size-independent shadow-ray property: nothing lies closer than the
    # closest hit, and the closest hit occludes a limit just beyond it.
    import torch

    log_case("full workload shadow-ray properties")
    sub = np.arange(0, n, 7)
    r2 = rays[:, sub]
    h2 = pid[sub] >= 0
    t2 = t[sub]
    om = raycast.ObjectManager()
    om.add_mesh(tris)
    om.prepare_for_rendering()
    rr = torch.from_numpy(np.ascontiguousarray(r2)).to(cuda)
    before = om.intersect_shadow_ray(rr, torch.from_numpy(np.where(h2, t2 * 0.9999, 1e30).astype(np.float32)).to(cuda))
    after = om.intersect_shadow_ray(rr, torch.from_numpy(np.where(h2, t2 * 1.0001, 1e30).astype(np.float32)).to(cuda))
    before = before.cpu().numpy()
    after = after.cpu().numpy()
    om.release()
    assert not np.any(before[h2])
    assert np.all(after[h2] == 1)
    assert not np.any(after[~h2])
