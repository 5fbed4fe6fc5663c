"""GPU ray queries over objects with transforms (rise_hip_scene_upload_objects)
through the C ABI, against the oracle (oracle/rt_oracle.c, itself pinned
against the reference's Object::IntersectRay path in tests/test_objects.py).

* A scene of 12 objects on 3 shared meshes (two identity objects, one
  mirrored), rays of random length, both arithmetic models, every face
  selection: hit records (object, id, world range, barycentrics, world
  point and normal), the plain closest-hit ids and distances, the
  allocation-free entry point, shadow rays with and without limits.
* A larger scene (48 objects on 4 meshes of 40,000 triangles, 1,000,000 rays)
  against the library's per-ray code run on the host on a sample, and
  identity objects against the world-space scene on every ray.
Bar: bit-exact (float fields as the oracle's doubles rounded to float)."""
import numpy as np
import pytest

import oracle
from rise_amd import _lib, raycast, scenes
from tests.conftest import log_case

pytestmark = pytest.mark.gpu

ARITH = {"ieee": raycast.ARITH_IEEE, "fast": raycast.ARITH_X86_FMA}


def _manager(meshes, objs, arith):
    om = raycast.ObjectManager(arith)
    for m in meshes:
        om.add_mesh(m)
    for ob in objs:
        om.add_object(int(ob["mesh"]), ob["to_world"], ob["to_object"])
    om.prepare_for_rendering()
    return om


def _tmax(n, seed):
    t = np.random.default_rng(seed).uniform(0.5, 12.0, n).astype(np.float32)
    t[::3] = np.float32(3.4e38)
    return t


@pytest.mark.parametrize("arith", ["ieee", "fast"])
def test_gpu_objects_match_oracle(cuda, arith):
    import torch

    meshes, objs = scenes.instanced_scene(tris_per_mesh=2000)
    rays = scenes.object_rays(meshes, objs, 8000, seed=11)
    tmax = _tmax(rays.shape[1], 12)
    om = _manager(meshes, objs, ARITH[arith])
    r = torch.from_numpy(rays).to(cuda)
    tm = torch.from_numpy(tmax).to(cuda)
    for faces in (raycast.FACES_BOTH, raycast.FACES_FRONT, raycast.FACES_BACK, raycast.FACES_NONE):
        log_case(f"objects: 12 objects x 2000 tris, {rays.shape[1]} rays, arith {arith}, faces {faces}")
        g = raycast.hit_fields(om.intersect_ray_ex(r, faces))
        o = oracle.objects_closest(meshes, objs, rays, faces=faces, arith=arith)
        hit = o["prim_id"] >= 0
        assert np.array_equal(g["prim_id"], o["prim_id"]) and np.array_equal(g["mesh"], o["object"])
        assert np.array_equal(g["range"][hit], o["range"][hit])
        assert np.array_equal(g["alpha"][hit], o["alpha"][hit]) and np.array_equal(g["beta"][hit], o["beta"][hit])
        assert np.array_equal(g["point"][hit], o["point"][hit].astype(np.float32))
        assert np.array_equal(g["geom_normal"][hit], o["normal"][hit].astype(np.float32))
        assert np.isinf(g["range"][~hit]).all()
        occ = om.intersect_shadow_ray_ex(r, faces, how_far=tm).cpu().numpy()
        assert np.array_equal(occ, oracle.objects_any(meshes, objs, rays, tmax, faces=faces, arith=arith))
        if faces == raycast.FACES_BOTH:
            assert set(np.unique(o["object"][hit])) == set(range(len(objs)))
            pid, t = om.intersect_ray(r)
            assert np.array_equal(pid.cpu().numpy(), g["prim_id"])
            assert np.array_equal(t.cpu().numpy(), g["range"].astype(np.float32))
            log_case("objects: allocation-free closest hit")
            n = rays.shape[1]
            pid2 = torch.full((n,), -7, dtype=torch.int32, device=cuda)
            t2 = torch.full((n,), -7.0, dtype=torch.float32, device=cuda)
            _lib.check(_lib.lib().rise_hip_intersect_closest_noalloc(om.handle, r.data_ptr(), n, pid2.data_ptr(),
                                                                     t2.data_ptr(), None))
            assert np.array_equal(pid2.cpu().numpy(), g["prim_id"])
            assert np.array_equal(t2.cpu().numpy(), g["range"].astype(np.float32))
            occ_all = om.intersect_shadow_ray(r).cpu().numpy()
            assert np.array_equal(occ_all, oracle.objects_any(meshes, objs, rays, arith=arith))
            assert np.array_equal(om.intersect_shadow_ray(r, how_far=tm).cpu().numpy(),
                                  oracle.objects_any(meshes, objs, rays, tmax, arith=arith))
    om.release()


def test_gpu_objects_unsupported_calls(cuda):
    import torch

    meshes, objs = scenes.instanced_scene(tris_per_mesh=200)
    om = _manager(meshes, objs, raycast.ARITH_IEEE)
    r = torch.from_numpy(scenes.object_rays(meshes, objs, 100, seed=1)).to(cuda)
    counters = torch.zeros(2, dtype=torch.int64, device=cuda)
    with pytest.raises(_lib.RiseHipError, match="transformed objects"):
        om.intersect_ray(r, counters=counters)
    with pytest.raises(_lib.RiseHipError, match="transformed objects"):
        om.save_cache("unused.rhbvh")
    om.release()


def test_gpu_objects_large_scene(cuda):
    import torch

    meshes, objs = scenes.instanced_scene(n_meshes=4, n_objects=48, tris_per_mesh=40000, seed=41, n_identity=4)
    rays = scenes.object_rays(meshes, objs, 200000, seed=43)
    rays = np.ascontiguousarray(np.concatenate([rays] * 5, 1))
    n = rays.shape[1]
    log_case(f"objects: 48 objects x 40000 tris, {n} rays")
    om = _manager(meshes, objs, raycast.ARITH_IEEE)
    r = torch.from_numpy(rays).to(cuda)
    g = raycast.hit_fields(om.intersect_ray_ex(r))
    occ = om.intersect_shadow_ray(r, how_far=torch.from_numpy(_tmax(n, 44)).to(cuda)).cpu().numpy()
    # the five copies of the ray set agree with each other
    k = 200000
    for c in range(1, 5):
        assert np.array_equal(g[c * k:(c + 1) * k], g[:k])
    sample = np.arange(0, k, 40)
    h = raycast.host_trace_objects(meshes, objs, np.ascontiguousarray(rays[:, sample]))
    assert np.array_equal(g[sample], h)
    ho = raycast.host_trace_objects(meshes, objs, np.ascontiguousarray(rays[:, sample]), mode=1,
                                    tmax=_tmax(n, 44)[sample])
    assert np.array_equal(occ[sample], ho)
    assert 0.3 < (g["prim_id"] >= 0).mean() < 0.95
    om.release()
    log_case("objects: identity objects against the world-space scene")
    ident = raycast.make_objects([(m, *raycast.object_matrices()) for m in range(4)])
    om = _manager(meshes, ident, raycast.ARITH_IEEE)
    w = raycast.ObjectManager()
    w.add_mesh(np.concatenate(meshes))
    w.prepare_for_rendering()
    a = raycast.hit_fields(om.intersect_ray_ex(r))
    b = raycast.hit_fields(w.intersect_ray_ex(r))
    hit = b["prim_id"] >= 0
    sizes = np.array([m.shape[0] for m in meshes])
    obj = np.searchsorted(np.cumsum(sizes), b["prim_id"][hit], side="right")
    assert np.array_equal(a["prim_id"] >= 0, hit) and np.array_equal(a["mesh"][hit], obj)
    for f in ("range", "alpha", "beta", "point", "geom_normal"):
        assert np.array_equal(a[f][hit], b[f][hit])
    om.release()
    w.release()
