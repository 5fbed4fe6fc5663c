"""Meshes whose vertices are doubles that floats cannot hold (RISE's meshes
keep Point3 vertices, Polygon.h:142; rise_hip_mesh_desc.vertices_f64), on the
CPU.

The reference's float pre-filter takes the vertices' float casts
(GetFloatTriangleVertices, TriangleMeshGeometryIndexedSpecializations.h:32-47)
and its certifier and face test the doubles (:141-142).
* The oracle on double vertices against the reference's mesh built from the
  same doubles (oracle/_ref), both builds, every face selection: ids, world
  range, point, normal, barycentrics, shadow rays.
* The same vertices rounded to float change the results (so the test would
  catch a certifier that used the floats).
* The library's per-ray code run on the host (world tree of an identity
  object, and transformed objects) against the oracle.
Bar: bit-exact."""
import numpy as np
import pytest

import oracle
from rise_amd import raycast, scenes

ARITH = {"ieee": raycast.ARITH_IEEE, "fast": raycast.ARITH_X86_FMA}
FACES = [raycast.FACES_BOTH, raycast.FACES_FRONT, raycast.FACES_BACK]
needs_ref = pytest.mark.skipif(not (oracle.ref_available("ieee") and oracle.ref_available("fast")),
                               reason="reference mesh path (oracle/_ref) not built")


def _double_soup(n, seed):
    """A soup whose vertices are doubles off the float grid (a relative
    perturbation of about 1e-9, far below a float step)."""
    t = scenes.soup_triangles(n, seed=seed, sigma=0.03).astype(np.float64)
    g = np.random.default_rng(seed + 100)
    t = t * (1.0 + g.uniform(-1e-9, 1e-9, t.shape)) + g.uniform(-1e-10, 1e-10, t.shape)
    assert (t != t.astype(np.float32).astype(np.float64)).mean() > 0.99
    return np.ascontiguousarray(t)


def _rays(tris, seed):
    r = np.concatenate([scenes.pinhole_rays(40, 30), scenes.incoherent_rays(3000, seed),
                        scenes.near_edge_rays(tris, 3000, seed + 1)], 1)
    return np.ascontiguousarray(r, np.float32)


@needs_ref
@pytest.mark.parametrize("arith", ["ieee", "fast"])
def test_oracle_double_vertices_match_reference(arith):
    tris = _double_soup(3000, 7)
    rays = _rays(tris, 8)
    plain = oracle.RefMesh(tris, variant=arith)
    bary = oracle.RefMesh(tris, variant=arith, bary=True)
    for faces in FACES:
        o = oracle.closest_attr(tris, rays, faces=faces, arith=arith)
        a = plain.closest_ex(rays, faces)
        b = bary.closest_ex(rays, faces)
        hit = o["prim_id"] >= 0
        assert np.array_equal(o["prim_id"], a["prim_id"]) and np.array_equal(a["prim_id"], b["prim_id"])
        for k in ("range", "point", "normal"):
            assert np.array_equal(o[k][hit], a[k][hit]), k
        assert np.array_equal(o["alpha"][hit], b["alpha"][hit]) and np.array_equal(o["beta"][hit], b["beta"][hit])
        assert np.array_equal(oracle.any_hit(tris, rays, faces=faces, arith=arith), plain.any_ex(rays, faces=faces))
    # rounding the vertices to float changes the answer
    o = oracle.closest_attr(tris, rays, arith=arith)
    f = oracle.closest_attr(tris.astype(np.float32), rays, arith=arith)
    both = (o["prim_id"] >= 0) & (o["prim_id"] == f["prim_id"])
    assert (o["range"][both] != f["range"][both]).mean() > 0.9
    assert (o["prim_id"] != f["prim_id"]).sum() > 0


@pytest.mark.parametrize("arith", ["ieee", "fast"])
def test_host_double_vertices_match_oracle(arith):
    tris = _double_soup(3000, 11)
    rays = _rays(tris, 12)
    ident = raycast.make_objects([(0, *raycast.object_matrices())])
    tmax = np.random.default_rng(3).uniform(0.1, 2.0, rays.shape[1]).astype(np.float32)
    for faces in FACES:
        h = raycast.host_trace_objects([tris], ident, rays, faces=faces, arith=ARITH[arith])
        o = oracle.closest_attr(tris, rays, faces=faces, arith=arith)
        hit = o["prim_id"] >= 0
        assert np.array_equal(h["prim_id"], o["prim_id"])
        assert np.array_equal(h["range"][hit], o["range"][hit])
        assert np.array_equal(h["alpha"][hit], o["alpha"][hit]) and np.array_equal(h["beta"][hit], o["beta"][hit])
        assert np.array_equal(h["point"][hit], o["point"][hit].astype(np.float32))
        assert np.array_equal(h["geom_normal"][hit], o["normal"][hit].astype(np.float32))
        occ = raycast.host_trace_objects([tris], ident, rays, mode=1, faces=faces, tmax=tmax, arith=ARITH[arith])
        assert np.array_equal(occ, oracle.any_hit(tris, rays, tmax, faces=faces, arith=arith))


@needs_ref
@pytest.mark.parametrize("arith", ["ieee", "fast"])
def test_transformed_double_meshes(arith):
    meshes, objs = scenes.instanced_scene(tris_per_mesh=400)
    g = np.random.default_rng(5)
    meshes = [np.ascontiguousarray(m.astype(np.float64) * (1.0 + g.uniform(-1e-9, 1e-9, m.shape))) for m in meshes]
    rays = scenes.object_rays(meshes, objs, 5000, seed=6)
    d = rays[3:6].astype(np.float64)
    rays[3:6] = (d / np.linalg.norm(d, axis=0)).astype(np.float32)
    o = oracle.objects_closest(meshes, objs, rays, arith=arith)
    r = oracle.ref_objects_closest([oracle.RefMesh(m, variant=arith) for m in meshes], objs, rays)
    h = raycast.host_trace_objects(meshes, objs, rays, arith=ARITH[arith])
    hit = o["prim_id"] >= 0
    assert np.array_equal(o["prim_id"], r["prim_id"]) and np.array_equal(o["object"], r["object"])
    for k in ("range", "point", "normal"):
        assert np.array_equal(o[k][hit], r[k][hit]), k
    assert np.array_equal(h["prim_id"], o["prim_id"]) and np.array_equal(h["mesh"], o["object"])
    assert np.array_equal(h["range"][hit], o["range"][hit])
    assert np.array_equal(h["point"][hit], o["point"][hit].astype(np.float32))


@needs_ref
@pytest.mark.parametrize("arith", ["ieee", "fast"])
def test_double_sided_mesh_matches_reference(arith):
    """A double-sided mesh (TriangleMeshGeometryIndexed.cpp:151-205): face
    flags do not apply, and the record's normal faces the ray."""
    tris = _double_soup(3000, 17)
    rays = _rays(tris, 18)
    plain = oracle.RefMesh(tris, variant=arith, double_sided=True)
    bary = oracle.RefMesh(tris, variant=arith, bary=True, double_sided=True)
    one = oracle.RefMesh(tris, variant=arith)
    ident = raycast.make_objects([(0, *raycast.object_matrices())])
    for faces in FACES:
        o = oracle.closest_attr(tris, rays, faces=faces, arith=arith, double_sided=True)
        a = plain.closest_ex(rays, faces)
        b = bary.closest_ex(rays, faces)
        hit = o["prim_id"] >= 0
        assert np.array_equal(o["prim_id"], a["prim_id"]) and np.array_equal(a["prim_id"], b["prim_id"])
        for k in ("range", "point", "normal"):
            assert np.array_equal(o[k][hit], a[k][hit]), k
        # the bary mesh's interpolated shading normal (1 - a - b, a, b) is
        # turned toward the ray as well, so it holds +-a
        assert np.array_equal(np.abs(o["alpha"][hit]), np.abs(b["alpha"][hit]))
        d = rays[3:6].T.astype(np.float64)
        assert (np.sum(o["normal"][hit] * d[hit], axis=1) <= 0).all()
        assert np.array_equal(oracle.any_hit(tris, rays, faces=faces, arith=arith, double_sided=True),
                              plain.any_ex(rays, faces=faces))
        h = raycast.host_trace_objects([tris], ident, rays, faces=faces, arith=ARITH[arith], double_sided=[True])
        assert np.array_equal(h["prim_id"], o["prim_id"]) and np.array_equal(h["range"][hit], o["range"][hit])
        assert np.array_equal(h["geom_normal"][hit], o["normal"][hit].astype(np.float32))
    # single-sided, the face flags cut hits
    assert (one.closest_ex(rays, raycast.FACES_FRONT)["prim_id"] >= 0).sum() < hit.sum()


@needs_ref
def test_mixed_sidedness_objects():
    meshes, objs = scenes.instanced_scene(tris_per_mesh=400, seed=19)
    rays = scenes.object_rays(meshes, objs, 5000, seed=20)
    d = rays[3:6].astype(np.float64)
    rays[3:6] = (d / np.linalg.norm(d, axis=0)).astype(np.float32)
    sides = [True, False, True]
    ref = [oracle.RefMesh(m, variant="ieee", double_sided=sd) for m, sd in zip(meshes, sides)]
    for faces in FACES:
        o = oracle.objects_closest(meshes, objs, rays, faces=faces, double_sided=sides)
        r = oracle.ref_objects_closest(ref, objs, rays, faces=faces)
        h = raycast.host_trace_objects(meshes, objs, rays, faces=faces, double_sided=sides)
        hit = o["prim_id"] >= 0
        assert np.array_equal(o["prim_id"], r["prim_id"]) and np.array_equal(o["object"], r["object"])
        assert np.array_equal(o["normal"][hit], r["normal"][hit]) and np.array_equal(o["range"][hit], r["range"][hit])
        assert np.array_equal(h["prim_id"], o["prim_id"]) and np.array_equal(h["mesh"], o["object"])
        assert np.array_equal(h["geom_normal"][hit], o["normal"][hit].astype(np.float32))
        assert np.array_equal(oracle.objects_any(meshes, objs, rays, faces=faces, double_sided=sides),
                              oracle.ref_objects_any(ref, objs, rays, faces=faces))
