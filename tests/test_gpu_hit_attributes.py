"""GPU hit records with shading normals and texture coordinates
(rise_hip_mesh_desc.normals / .coords) through the C ABI, against the oracle
(itself pinned against the reference's meshes in tests/test_hit_attributes.py):
a world-space mesh with per-corner normals and coordinates, and transformed
objects mixing meshes with and without them, single- and double-sided, both
arithmetic models.  Bar: bit-exact (float fields as the oracle's doubles
rounded)."""
import numpy as np
import pytest

import oracle
from rise_amd import raycast, scenes
from tests.conftest import log_case
from tests.test_hit_attributes import _scene

pytestmark = pytest.mark.gpu

ARITH = {"ieee": raycast.ARITH_IEEE, "fast": raycast.ARITH_X86_FMA}


def _check(g, o):
    hit = o["prim_id"] >= 0
    assert np.array_equal(g["prim_id"], o["prim_id"])
    assert np.array_equal(g["range"][hit], o["range"][hit])
    assert np.array_equal(g["geom_normal"][hit], o["normal"][hit].astype(np.float32))
    assert np.array_equal(g["shading_normal"][hit], o["shading"][hit].astype(np.float32))
    assert np.array_equal(g["uv"][hit], o["uv"][hit].astype(np.float32))
    assert np.array_equal(g["point"][hit], o["point"][hit].astype(np.float32))


@pytest.mark.parametrize("arith", ["ieee", "fast"])
def test_gpu_hit_attributes(cuda, arith):
    import torch

    meshes, objs, normals, _, coords, rays = _scene(seed=71, n_rays=8000)
    r = torch.from_numpy(rays).to(cuda)
    log_case(f"hit attributes: world-space mesh, arith {arith}")
    om = raycast.ObjectManager(ARITH[arith])
    om.add_mesh(meshes[0], normals=normals[0], coords=coords[0], double_sided=True)
    om.prepare_for_rendering()
    g = raycast.hit_fields(om.intersect_ray_ex(r))
    ident = raycast.make_objects([(0, *raycast.object_matrices())])
    o = oracle.objects_closest(meshes[:1], ident, rays, arith=arith, double_sided=[True], normals=normals[:1],
                               coords=coords[:1])
    _check(g, o)
    assert (g["prim_id"] >= 0).sum() > 500
    om.release()
    log_case(f"hit attributes: 12 objects, arith {arith}")
    sides = [False, True, False]
    om = raycast.ObjectManager(ARITH[arith])
    for m, n, c, sd in zip(meshes, normals, coords, sides):
        om.add_mesh(m, normals=n, coords=c, double_sided=sd)
    for ob in objs:
        om.add_object(int(ob["mesh"]), ob["to_world"], ob["to_object"])
    om.prepare_for_rendering()
    g = raycast.hit_fields(om.intersect_ray_ex(r))
    o = oracle.objects_closest(meshes, objs, rays, arith=arith, double_sided=sides, normals=normals, coords=coords)
    _check(g, o)
    assert np.array_equal(g["mesh"], o["object"])
    om.release()
