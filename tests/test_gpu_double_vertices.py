"""GPU ray queries on meshes given as double vertices
(rise_hip_mesh_desc.vertices_f64) through the C ABI, against the oracle on the
same doubles (itself pinned against the reference's mesh built from them in
tests/test_double_vertices.py): a 30,000-triangle soup off the float grid,
both arithmetic models, every face selection, the record, plain, shadow and
allocation-free entry points; and a scene of transformed objects on such
meshes.  Bar: bit-exact (float fields as the oracle's doubles rounded)."""
import numpy as np
import pytest

import oracle
from rise_amd import _lib, raycast, scenes
from tests.conftest import log_case
from tests.test_double_vertices import _double_soup

pytestmark = pytest.mark.gpu

ARITH = {"ieee": raycast.ARITH_IEEE, "fast": raycast.ARITH_X86_FMA}


@pytest.mark.parametrize("arith", ["ieee", "fast"])
def test_gpu_double_vertices_match_oracle(cuda, arith):
    import torch

    tris = _double_soup(30000, 21)
    rays = np.ascontiguousarray(np.concatenate([scenes.pinhole_rays(100, 80), scenes.incoherent_rays(12000, 22),
                                                scenes.near_edge_rays(tris, 12000, 23)], 1), np.float32)
    n = rays.shape[1]
    tmax = np.random.default_rng(24).uniform(0.05, 1.5, n).astype(np.float32)
    om = raycast.ObjectManager(ARITH[arith])
    om.add_mesh(tris)
    om.prepare_for_rendering()
    r = torch.from_numpy(rays).to(cuda)
    tm = torch.from_numpy(tmax).to(cuda)
    for faces in (raycast.FACES_BOTH, raycast.FACES_FRONT, raycast.FACES_BACK):
        log_case(f"double vertices: 30000 tris, {n} rays, arith {arith}, faces {faces}")
        g = raycast.hit_fields(om.intersect_ray_ex(r, faces))
        o = oracle.closest_attr(tris, rays, faces=faces, arith=arith)
        hit = o["prim_id"] >= 0
        assert np.array_equal(g["prim_id"], o["prim_id"])
        assert np.array_equal(g["range"][hit], o["range"][hit])
        assert np.array_equal(g["alpha"][hit], o["alpha"][hit]) and np.array_equal(g["beta"][hit], o["beta"][hit])
        assert np.array_equal(g["point"][hit], o["point"][hit].astype(np.float32))
        assert np.array_equal(g["geom_normal"][hit], o["normal"][hit].astype(np.float32))
        occ = om.intersect_shadow_ray_ex(r, faces, how_far=tm).cpu().numpy()
        assert np.array_equal(occ, oracle.any_hit(tris, rays, tmax, faces=faces, arith=arith))
        if faces == raycast.FACES_BOTH:
            pid, t = om.intersect_ray(r)
            assert np.array_equal(pid.cpu().numpy(), o["prim_id"])
            assert np.array_equal(t.cpu().numpy()[hit], o["range"][hit].astype(np.float32))
            pid2 = torch.full((n,), -7, dtype=torch.int32, device=cuda)
            t2 = torch.full((n,), -7.0, dtype=torch.float32, device=cuda)
            _lib.check(_lib.lib().rise_hip_intersect_closest_noalloc(om.handle, r.data_ptr(), n, pid2.data_ptr(),
                                                                     t2.data_ptr(), None))
            assert np.array_equal(pid2.cpu().numpy(), o["prim_id"])
            assert np.array_equal(om.intersect_shadow_ray(r, how_far=tm).cpu().numpy(),
                                  oracle.any_hit(tris, rays, tmax, arith=arith))
    with pytest.raises(_lib.RiseHipError, match="double-precision"):
        om.save_cache("unused.rhbvh")
    om.release()


def test_gpu_transformed_double_meshes(cuda):
    import torch

    meshes, objs = scenes.instanced_scene(tris_per_mesh=1500, seed=33)
    g = np.random.default_rng(34)
    meshes = [np.ascontiguousarray(m.astype(np.float64) * (1.0 + g.uniform(-1e-9, 1e-9, m.shape))) for m in meshes]
    rays = scenes.object_rays(meshes, objs, 8000, seed=35)
    log_case("double vertices: 12 objects x 1500 tris, 8000 rays")
    om = raycast.ObjectManager()
    for m in meshes:
        om.add_mesh(m)
    for ob in objs:
        om.add_object(int(ob["mesh"]), ob["to_world"], ob["to_object"])
    om.prepare_for_rendering()
    r = torch.from_numpy(rays).to(cuda)
    h = raycast.hit_fields(om.intersect_ray_ex(r))
    o = oracle.objects_closest(meshes, objs, rays)
    hit = o["prim_id"] >= 0
    assert np.array_equal(h["prim_id"], o["prim_id"]) and np.array_equal(h["mesh"], o["object"])
    assert np.array_equal(h["range"][hit], o["range"][hit])
    assert np.array_equal(h["point"][hit], o["point"][hit].astype(np.float32))
    assert np.array_equal(h["geom_normal"][hit], o["normal"][hit].astype(np.float32))
    assert np.array_equal(om.intersect_shadow_ray(r).cpu().numpy(), oracle.objects_any(meshes, objs, rays))
    om.release()


def test_gpu_double_sided(cuda):
    """Double-sided meshes (RISE_HIP_MESH_DOUBLE_SIDED): face selection does
    not apply to them and the record's normal faces the ray; a world-space
    scene of one such mesh, and objects mixing both kinds."""
    import torch

    tris = _double_soup(20000, 41)
    rays = np.ascontiguousarray(np.concatenate([scenes.pinhole_rays(80, 60), scenes.incoherent_rays(8000, 42)], 1),
                                np.float32)
    om = raycast.ObjectManager()
    om.add_mesh(tris, double_sided=True)
    om.prepare_for_rendering()
    r = torch.from_numpy(rays).to(cuda)
    for faces in (raycast.FACES_BOTH, raycast.FACES_FRONT, raycast.FACES_BACK):
        log_case(f"double-sided mesh: 20000 tris, faces {faces}")
        g = raycast.hit_fields(om.intersect_ray_ex(r, faces))
        o = oracle.closest_attr(tris, rays, faces=faces, double_sided=True)
        hit = o["prim_id"] >= 0
        assert np.array_equal(g["prim_id"], o["prim_id"]) and np.array_equal(g["range"][hit], o["range"][hit])
        assert np.array_equal(g["geom_normal"][hit], o["normal"][hit].astype(np.float32))
        assert np.array_equal(om.intersect_shadow_ray_ex(r, faces).cpu().numpy(),
                              oracle.any_hit(tris, rays, faces=faces, double_sided=True))
    om.release()
    meshes, objs = scenes.instanced_scene(tris_per_mesh=1500, seed=43)
    sides = [True, False, True]
    orays = scenes.object_rays(meshes, objs, 8000, seed=44)
    om = raycast.ObjectManager()
    for m, sd in zip(meshes, sides):
        om.add_mesh(m, double_sided=sd)
    for ob in objs:
        om.add_object(int(ob["mesh"]), ob["to_world"], ob["to_object"])
    om.prepare_for_rendering()
    r = torch.from_numpy(orays).to(cuda)
    for faces in (raycast.FACES_FRONT, raycast.FACES_BACK):
        log_case(f"double-sided objects: faces {faces}")
        g = raycast.hit_fields(om.intersect_ray_ex(r, faces))
        o = oracle.objects_closest(meshes, objs, orays, faces=faces, double_sided=sides)
        hit = o["prim_id"] >= 0
        assert np.array_equal(g["prim_id"], o["prim_id"]) and np.array_equal(g["mesh"], o["object"])
        assert np.array_equal(g["geom_normal"][hit], o["normal"][hit].astype(np.float32))
        assert np.array_equal(om.intersect_shadow_ray_ex(r, faces).cpu().numpy(),
                              oracle.objects_any(meshes, objs, orays, faces=faces, double_sided=sides))
    om.release()
