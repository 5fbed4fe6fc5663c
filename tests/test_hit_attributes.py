"""Shading normals and texture coordinates of hit records
(rise_hip_mesh_desc.normals / .coords; rise_hip_hit.shading_normal / .uv), on
the CPU.

The reference interpolates the corners' vertex normals and texture
coordinates at the certifier's barycentrics, turns the normalized face normal
to the shading normal's side (TriangleMeshGeometryIndexedSpecializations.h:
191-216), turns both normals toward the ray on a double-sided mesh
(TriangleMeshGeometryIndexed.cpp:170-188), and Object::IntersectRay takes them
by the inverse transpose and normalizes them (Object.cpp:531-533).
* The oracle against the reference's meshes built with per-corner normals and
  coordinates (oracle/_ref; the u coordinate carries the triangle id, v is
  random), on transformed and identity objects, single- and double-sided, in
  both builds: shading normal, geometric normal and v bit for bit.
* The library's per-ray code on the host against the oracle.
Bar: bit-exact (float fields as the oracle's doubles rounded)."""
import numpy as np
import pytest

import oracle
from rise_amd import raycast, scenes

ARITH = {"ieee": raycast.ARITH_IEEE, "fast": raycast.ARITH_X86_FMA}
needs_ref = pytest.mark.skipif(not (oracle.ref_available("ieee") and oracle.ref_available("fast")),
                               reason="reference mesh path (oracle/_ref) not built")


def _scene(seed=51, n_rays=5000):
    meshes, objs = scenes.instanced_scene(tris_per_mesh=400, seed=seed)
    g = np.random.default_rng(seed + 1)
    meshes = [np.ascontiguousarray(m.astype(np.float64) * (1 + g.uniform(-1e-9, 1e-9, m.shape))) for m in meshes]
    normals = [g.normal(size=(m.shape[0], 9)) if k != 1 else None for k, m in enumerate(meshes)]
    vco = [g.uniform(-2, 2, size=(m.shape[0], 3)) for m in meshes]
    coords = [np.ascontiguousarray(np.stack([np.repeat(np.arange(m.shape[0], dtype=np.float64)[:, None], 3, 1), v],
                                            2).reshape(-1, 6)) for m, v in zip(meshes, vco)]
    rays = scenes.object_rays(meshes, objs, n_rays, seed=seed + 2)
    d = rays[3:6].astype(np.float64)
    rays[3:6] = (d / np.linalg.norm(d, axis=0)).astype(np.float32)
    return meshes, objs, normals, vco, coords, np.ascontiguousarray(rays)


@needs_ref
@pytest.mark.parametrize("arith", ["ieee", "fast"])
@pytest.mark.parametrize("sides", [(False, False, False), (True, False, True)])
def test_oracle_attributes_match_reference(arith, sides):
    meshes, objs, normals, vco, coords, rays = _scene()
    ref = [oracle.RefMesh(m, variant=arith, normals=n, vcoords=v, double_sided=s)
           for m, n, v, s in zip(meshes, normals, vco, sides)]
    r = oracle.ref_objects_closest(ref, objs, rays)
    o = oracle.objects_closest(meshes, objs, rays, arith=arith, double_sided=sides, normals=normals, coords=coords)
    hit = o["prim_id"] >= 0
    assert np.array_equal(o["prim_id"], r["prim_id"]) and np.array_equal(o["object"], r["object"])
    for k in ("normal", "shading", "range", "point"):
        assert np.array_equal(o[k][hit], r[k][hit]), k
    assert np.array_equal(o["uv"][hit, 1], r["vcoord"][hit])
    assert np.array_equal(o["uv"][hit, 0], o["prim_id"][hit].astype(np.float64))
    # the vertex normals do shade: shading and geometric normals differ
    with_normals = hit & (o["object"] % 3 != 1)
    assert (np.abs(o["shading"][with_normals] - o["normal"][with_normals]).max(axis=1) > 1e-3).mean() > 0.9
    # the geometric normal lies on the shading normal's side (in the mesh's
    # frame; shown on the identity objects, whose frame is the world's, of
    # single-sided meshes, whose normals are not turned one by one)
    ident = hit & (o["object"] < 2)
    assert ident.sum() > 100
    if not any(sides):
        assert (np.sum(o["shading"][ident] * o["normal"][ident], axis=1) >= -1e-12).all()


@pytest.mark.parametrize("arith", ["ieee", "fast"])
def test_host_attributes_match_oracle(arith):
    meshes, objs, normals, _, coords, rays = _scene(seed=61)
    sides = [True, False, False]
    h = raycast.host_trace_objects(meshes, objs, rays, arith=ARITH[arith], double_sided=sides, normals=normals,
                                   coords=coords)
    o = oracle.objects_closest(meshes, objs, rays, arith=arith, double_sided=sides, normals=normals, coords=coords)
    hit = o["prim_id"] >= 0
    assert np.array_equal(h["prim_id"], o["prim_id"]) and np.array_equal(h["mesh"], o["object"])
    assert np.array_equal(h["geom_normal"][hit], o["normal"][hit].astype(np.float32))
    assert np.array_equal(h["shading_normal"][hit], o["shading"][hit].astype(np.float32))
    assert np.array_equal(h["uv"][hit], o["uv"][hit].astype(np.float32))
    assert (h["uv"][~hit] == 0).all() and (h["shading_normal"][~hit] == 0).all()
    # without attributes the shading normal is the geometric one
    h0 = raycast.host_trace_objects(meshes, objs, rays, arith=ARITH[arith])
    assert np.array_equal(h0["shading_normal"][hit], h0["geom_normal"][hit]) and (h0["uv"] == 0).all()
