"""GPU intersection records and face selection (rise_hip_intersect_closest_ex,
rise_hip_intersect_any_ex) through the C ABI, against the oracle
(oracle/rt_oracle.c, itself pinned against the reference's mesh path in
tests/test_raycast_attrs.py), in both arithmetic models and for every face
selection, on a scene of three meshes (the record's mesh index).
Bar: ids, world range, barycentrics bit-exact; hit point and normal equal
to the oracle's doubles rounded to float."""
import numpy as np
import pytest

import oracle
from rise_amd import raycast, scenes
from tests.conftest import log_case

pytestmark = pytest.mark.gpu

ARITH = {"ieee": raycast.ARITH_IEEE, "fast": raycast.ARITH_X86_FMA}


@pytest.mark.parametrize("arith", ["ieee", "fast"])
def test_gpu_hit_records_and_faces(cuda, arith):
    import torch

    tris = scenes.soup_triangles(30000, seed=31, sigma=0.01)
    rays = np.ascontiguousarray(np.concatenate([scenes.pinhole_rays(120, 90), scenes.incoherent_rays(20000, 32)], 1),
                                np.float32)
    cuts = [0, 9000, 21000, 30000]
    om = raycast.ObjectManager(ARITH[arith])
    for a, b in zip(cuts[:-1], cuts[1:]):
        om.add_mesh(tris[a:b])
    om.prepare_for_rendering()
    r = torch.from_numpy(rays).to(cuda)
    mesh_of = np.repeat(np.arange(3), np.diff(cuts))
    for faces in (raycast.FACES_BOTH, raycast.FACES_FRONT, raycast.FACES_BACK, raycast.FACES_NONE):
        log_case(f"hit records: 30000 tris in 3 meshes, {rays.shape[1]} rays, arith {arith}, faces {faces}")
        g = raycast.hit_fields(om.intersect_ray_ex(r, faces))
        o = oracle.closest_attr(tris, rays, faces=faces, arith=arith)
        hit = o["prim_id"] >= 0
        assert np.array_equal(g["prim_id"], o["prim_id"])
        assert np.array_equal(g["range"][hit], o["range"][hit])
        assert np.array_equal(g["alpha"][hit], o["alpha"][hit]) and np.array_equal(g["beta"][hit], o["beta"][hit])
        assert np.array_equal(g["point"][hit], o["point"][hit].astype(np.float32))
        assert np.array_equal(g["geom_normal"][hit], o["normal"][hit].astype(np.float32))
        assert np.array_equal(g["mesh"][hit], mesh_of[o["prim_id"][hit]])
        assert (g["mesh"][~hit] == -1).all() and np.isinf(g["range"][~hit]).all()
        occ = om.intersect_shadow_ray_ex(r, faces).cpu().numpy()
        assert np.array_equal(occ, oracle.any_hit(tris, rays, faces=faces, arith=arith))
        if faces == raycast.FACES_BOTH:
            pid, t = om.intersect_ray(r)
            assert np.array_equal(pid.cpu().numpy(), g["prim_id"])
            assert np.array_equal(t.cpu().numpy(), g["range"].astype(np.float32))
        if faces == raycast.FACES_NONE:
            assert not hit.any()
    om.release()


def test_gpu_ex_rejects_bad_arguments(cuda):
    from rise_amd import _lib

    L = _lib.lib()
    om = raycast.ObjectManager()
    om.add_mesh(scenes.soup_triangles(100, seed=2))
    om.prepare_for_rendering()
    assert L.rise_hip_intersect_closest_ex(om.handle, None, 4, 5, None, None) != _lib.OK
    assert b"faces" in L.rise_hip_last_error()
    assert L.rise_hip_intersect_any_ex(om.handle, None, None, 4, -1, None, None) != _lib.OK
    assert L.rise_hip_intersect_closest_ex(om.handle, None, 0, 3, None, None) == _lib.OK
    om.release()
