"""The intersection record and the face selection of the ray queries
(rise_hip_intersect_closest_ex / _any_ex), on the CPU.

* The oracle (oracle/rt_oracle.c: brute force, FP64) against the reference's
  own mesh path compiled from its sources (oracle/_ref), in both arithmetic
  models and for every face selection: ids, world range, hit point and
  geometric normal from IntersectRay on a plain mesh, and the certifier's
  barycentrics read back through vertex normals (1,0,0), (0,1,0), (0,0,1),
  which make the reference's interpolated shading normal (1 - a - b, a, b);
  shadow-ray occlusion with the same face flags.
* The kernels' traversal code run on the host against the oracle.
Bar: everything bit-exact (doubles compared as doubles)."""
import numpy as np
import pytest

import oracle
from rise_amd import raycast, scenes

ARITH = {"ieee": raycast.ARITH_IEEE, "fast": raycast.ARITH_X86_FMA}
FACES = [raycast.FACES_BOTH, raycast.FACES_FRONT, raycast.FACES_BACK, raycast.FACES_NONE]


def _case(seed):
    tris = scenes.soup_triangles(3000, seed=seed, sigma=0.03)
    rays = np.concatenate([scenes.pinhole_rays(40, 30), scenes.incoherent_rays(3000, seed + 1)], 1)
    return tris, np.ascontiguousarray(rays, np.float32)


def _same(a, b):
    return np.array_equal(np.asarray(a), np.asarray(b))


@pytest.mark.skipif(not (oracle.ref_available("ieee") and oracle.ref_available("fast")),
                    reason="reference mesh path (oracle/_ref) not built")
@pytest.mark.parametrize("arith", ["ieee", "fast"])
def test_oracle_record_matches_reference(arith):
    tris, rays = _case(5)
    plain = oracle.RefMesh(tris, variant=arith)
    bary = oracle.RefMesh(tris, variant=arith, bary=True)
    n_front = n_back = 0
    for faces in FACES:
        o = oracle.closest_attr(tris, rays, faces=faces, arith=arith)
        a = plain.closest_ex(rays, faces)
        b = bary.closest_ex(rays, faces)
        hit = o["prim_id"] >= 0
        assert _same(o["prim_id"], a["prim_id"]) and _same(a["prim_id"], b["prim_id"])
        assert _same(o["range"][hit], a["range"][hit])
        assert _same(o["point"][hit], a["point"][hit])
        assert _same(o["normal"][hit], a["normal"][hit])
        assert _same(o["alpha"][hit], b["alpha"][hit]) and _same(o["beta"][hit], b["beta"][hit])
        assert _same(oracle.any_hit(tris, rays, faces=faces, arith=arith), plain.any_ex(rays, faces=faces))
        if faces == raycast.FACES_FRONT:
            n_front = int(hit.sum())
        if faces == raycast.FACES_BACK:
            n_back = int(hit.sum())
        if faces == raycast.FACES_NONE:
            assert not hit.any()
    # both selections hit something, and the face test is really applied
    full = oracle.closest_attr(tris, rays, arith=arith)["prim_id"] >= 0
    assert n_front > 100 and n_back > 100 and n_front + n_back > full.sum()


def test_oracle_record_closed_forms():
    """One triangle, rays aimed at known barycentric points from both sides:
    alpha, beta, the unit normal (0, 0, +-1 by winding), the hit point and
    which faces are hit."""
    tri = np.array([[0, 0, 0, 2, 0, 0, 0, 2, 0]], np.float32)   # normal +z
    pts = [(0.5, 0.25), (0.125, 0.75), (0.25, 0.25)]
    o = np.array([[2 * a, 2 * b, 3.0] for a, b in pts] + [[2 * a, 2 * b, -3.0] for a, b in pts], np.float32)
    d = np.array([[0, 0, -1.0]] * 3 + [[0, 0, 1.0]] * 3, np.float32)
    rays = np.ascontiguousarray(np.concatenate([o.T, d.T], 0))
    r = oracle.closest_attr(tri, rays)
    assert (r["prim_id"] == 0).all()
    ab = np.array(pts * 2)
    assert np.allclose(r["alpha"], ab[:, 0], atol=1e-15) and np.allclose(r["beta"], ab[:, 1], atol=1e-15)
    assert np.allclose(r["normal"], [[0, 0, 1.0]] * 6, atol=0)
    assert np.allclose(r["range"], 3.0 - 1e-12, rtol=1e-15)
    assert np.allclose(r["point"][:, :2], o[:, :2], atol=0) and np.allclose(np.abs(r["point"][:, 2]), 1e-12, rtol=1e-3)
    # from above (d = -z) the +z normal faces the ray: front; from below: back
    front = oracle.closest_attr(tri, rays, faces=raycast.FACES_FRONT)["prim_id"]
    back = oracle.closest_attr(tri, rays, faces=raycast.FACES_BACK)["prim_id"]
    assert list(front) == [0, 0, 0, -1, -1, -1] and list(back) == [-1, -1, -1, 0, 0, 0]


@pytest.mark.parametrize("arith", ["ieee", "fast"])
def test_host_traversal_record_matches_oracle(arith):
    tris, rays = _case(9)
    for faces in FACES[:3]:
        o = oracle.closest_attr(tris, rays, faces=faces, arith=arith)
        h = raycast.host_trace_ex(tris, rays, 0, faces, arith=ARITH[arith])
        hit = o["prim_id"] >= 0
        assert _same(h["prim_id"], o["prim_id"])
        assert _same(h["range"][hit], o["range"][hit])
        assert _same(h["alpha"][hit], o["alpha"][hit]) and _same(h["beta"][hit], o["beta"][hit])
        assert _same(h["point"][hit], o["point"][hit].astype(np.float32))
        assert _same(h["geom_normal"][hit], o["normal"][hit].astype(np.float32))
        assert (h["mesh"][hit] == 0).all() and (h["mesh"][~hit] == -1).all()
        occ = raycast.host_trace_ex(tris, rays, 1, faces, arith=ARITH[arith])
        assert _same(occ, oracle.any_hit(tris, rays, faces=faces, arith=arith))
    # with both faces selected the record's ids are the plain query's
    pid, _, _, _ = raycast.host_trace(tris, rays, 0, arith=ARITH[arith])
    assert _same(raycast.host_trace_ex(tris, rays, 0, raycast.FACES_BOTH, arith=ARITH[arith])["prim_id"], pid)
