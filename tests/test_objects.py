"""Objects with transforms (rise_hip_scene_upload_objects), on the CPU.

The reference intersects each object in its own frame: Object::IntersectRay /
_IntersectionOnly (Objects/Object.cpp:392-712) inside ObjectManager's loop
(Managers/ObjectManager.cpp:270-373).

* The oracle (oracle/rt_oracle.c, brute force) against the reference's mesh
  path compiled from its sources, with Object::IntersectRay's steps performed
  around it by the reference's own Math3D (oracle/ref/ref_harness.cpp), in
  both arithmetic models and for every face selection: object, triangle id,
  world range, world hit point, world normal, barycentrics; shadow rays.
* The reference's box pre-test (Object.cpp:472-527) scales the running best
  by |M^-1 d|, which for a direction shorter than 1 is short of the best's
  object-space image, and can drop a closer object.  The oracle models it on
  request; with it the oracle equals the reference on such rays, without it
  every difference is a ray with |d| < 1 where the reference's hit is
  farther.  The library, like the oracle's default, does not reproduce that
  cull (RISE's cameras and integrators trace unit directions).
* The library's per-ray code (object_trace.h, run on the host) against the
  oracle; identity objects against the plain world-space path.
* Closed forms: a quad placed by a known transform.
* The slab test of object-space rays against exact rational slabs.
Bar: bit-exact (doubles compared as doubles; the library's float fields as
the oracle's doubles rounded to float)."""
import ctypes
import struct
from fractions import Fraction as F

import numpy as np
import pytest

import oracle
from rise_amd import _lib, raycast, scenes

ARITH = {"ieee": raycast.ARITH_IEEE, "fast": raycast.ARITH_X86_FMA}
FACES = [raycast.FACES_BOTH, raycast.FACES_FRONT, raycast.FACES_BACK]
needs_ref = pytest.mark.skipif(not (oracle.ref_available("ieee") and oracle.ref_available("fast")),
                               reason="reference mesh path (oracle/_ref) not built")


def _scene(tris=400, n_rays=6000, seed=5, unit=True):
    meshes, objs = scenes.instanced_scene(tris_per_mesh=tris)
    rays = scenes.object_rays(meshes, objs, n_rays, seed=seed)
    if unit:
        d = rays[3:6].astype(np.float64)
        rays[3:6] = (d / np.linalg.norm(d, axis=0)).astype(np.float32)
    return meshes, objs, np.ascontiguousarray(rays)


def _tmax(n, seed=1):
    t = np.random.default_rng(seed).uniform(0.5, 12.0, n).astype(np.float32)
    t[::3] = np.float32(3.4e38)
    return t


def _same(a, b):
    return np.array_equal(np.asarray(a), np.asarray(b))


@needs_ref
@pytest.mark.parametrize("arith", ["ieee", "fast"])
def test_oracle_objects_match_reference(arith):
    meshes, objs, rays = _scene()
    plain = [oracle.RefMesh(m, variant=arith) for m in meshes]
    bary = [oracle.RefMesh(m, variant=arith, bary=True) for m in meshes]
    tmax = _tmax(rays.shape[1])
    for faces in FACES:
        o = oracle.objects_closest(meshes, objs, rays, faces=faces, arith=arith)
        a = oracle.ref_objects_closest(plain, objs, rays, faces=faces)
        b = oracle.ref_objects_closest(bary, objs, rays, faces=faces, bary=True)
        hit = o["prim_id"] >= 0
        assert _same(o["prim_id"], a["prim_id"]) and _same(o["object"], a["object"])
        assert _same(a["prim_id"], b["prim_id"]) and _same(a["object"], b["object"])
        assert _same(o["range"][hit], a["range"][hit])
        assert _same(o["point"][hit], a["point"][hit])
        assert _same(o["normal"][hit], a["normal"][hit])
        assert _same(o["alpha"][hit], b["alpha"][hit]) and _same(o["beta"][hit], b["beta"][hit])
        occ = oracle.objects_any(meshes, objs, rays, tmax, faces=faces, arith=arith)
        assert _same(occ, oracle.ref_objects_any(plain, objs, rays, tmax, faces=faces))
        # every object is hit, transformed and identity ones alike, and both
        # outcomes of the shadow test occur
        assert set(np.unique(o["object"][hit])) == set(range(len(objs)))
        assert 0.2 < occ.mean() < 0.9


@needs_ref
def test_box_pretest_explains_short_directions():
    meshes, objs, rays = _scene(unit=False, n_rays=8000)
    ref = [oracle.RefMesh(m, variant="ieee") for m in meshes]
    r = oracle.ref_objects_closest(ref, objs, rays)
    with_pre = oracle.objects_closest(meshes, objs, rays, prehit=True)
    plain = oracle.objects_closest(meshes, objs, rays)
    assert _same(with_pre["prim_id"], r["prim_id"]) and _same(with_pre["object"], r["object"])
    assert _same(with_pre["range"], r["range"])
    diff = (plain["prim_id"] != r["prim_id"]) | (plain["object"] != r["object"])
    length = np.linalg.norm(rays[3:6].astype(np.float64), axis=0)
    assert diff.sum() > 10  # the workload exercises the cull
    assert (length[diff] < 1.0).all() and (r["range"][diff] > plain["range"][diff]).all()
    tmax = _tmax(rays.shape[1])
    occ_ref = oracle.ref_objects_any(ref, objs, rays, tmax)
    assert _same(oracle.objects_any(meshes, objs, rays, tmax, prehit=True), occ_ref)
    assert _same(oracle.objects_any(meshes, objs, rays, tmax), occ_ref)


@pytest.mark.parametrize("arith", ["ieee", "fast"])
@pytest.mark.parametrize("unit", [True, False])
def test_host_objects_match_oracle(arith, unit):
    meshes, objs, rays = _scene(unit=unit, seed=7)
    tmax = _tmax(rays.shape[1], seed=3)
    for faces in FACES + [raycast.FACES_NONE]:
        o = oracle.objects_closest(meshes, objs, rays, faces=faces, arith=arith)
        h = raycast.host_trace_objects(meshes, objs, rays, faces=faces, arith=ARITH[arith])
        hit = o["prim_id"] >= 0
        assert _same(h["prim_id"], o["prim_id"]) and _same(h["mesh"], o["object"])
        assert _same(h["range"][hit], o["range"][hit])
        assert _same(h["alpha"][hit], o["alpha"][hit]) and _same(h["beta"][hit], o["beta"][hit])
        assert _same(h["point"][hit], o["point"][hit].astype(np.float32))
        assert _same(h["geom_normal"][hit], o["normal"][hit].astype(np.float32))
        assert np.isinf(h["range"][~hit]).all() and (h["mesh"][~hit] == -1).all()
        occ = raycast.host_trace_objects(meshes, objs, rays, mode=1, faces=faces, tmax=tmax, arith=ARITH[arith])
        assert _same(occ, oracle.objects_any(meshes, objs, rays, tmax, faces=faces, arith=arith))
        if faces == raycast.FACES_NONE:
            assert not hit.any() and not occ.any()


def test_coincident_objects_keep_the_first():
    """Two objects with the same mesh and the same matrices give equal world
    ranges; the reference keeps the first in its order (strictly closer
    wins), whatever order the top-level tree visits them in."""
    meshes, objs, rays = _scene(seed=13, n_rays=3000)
    dup = np.concatenate([objs[[5]], objs[2:5], objs[[5]], objs[[9]], objs[[9]]])
    for arith in ("ieee", "fast"):
        o = oracle.objects_closest(meshes, dup, rays, arith=arith)
        h = raycast.host_trace_objects(meshes, dup, rays, arith=ARITH[arith])
        assert _same(h["prim_id"], o["prim_id"]) and _same(h["mesh"], o["object"])
        assert (o["object"] == 0).sum() > 100 and (o["object"] == 5).sum() > 100
        assert not ((o["object"] == 4) | (o["object"] == 6)).any()


def test_identity_objects_equal_world_space_path():
    meshes, _, rays = _scene(seed=9)
    objs = raycast.make_objects([(k, *raycast.object_matrices()) for k in range(len(meshes))])
    h = raycast.host_trace_objects(meshes, objs, rays)
    w = raycast.host_trace_ex(np.concatenate(meshes), rays)
    sizes = np.array([m.shape[0] for m in meshes])
    start = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    hit = w["prim_id"] >= 0
    assert _same(h["prim_id"] >= 0, hit)
    obj = np.searchsorted(np.cumsum(sizes), w["prim_id"][hit], side="right")
    assert _same(h["mesh"][hit], obj)
    assert _same(h["prim_id"][hit], w["prim_id"][hit] - start[obj])
    for k in ("range", "alpha", "beta", "point", "geom_normal"):
        assert _same(h[k][hit], w[k][hit])


def test_transformed_quad_closed_form():
    """A unit quad z = 0, [0,1]^2, scaled by (2, 3, 4), rotated 90 degrees
    about z and translated by (5, 0, -1): its world image is the rectangle
    x in [2, 5], y in [0, 2], z = -1, with normal +-z.  Rays from z = 3 straight
    down hit at world distance 4 exactly, at barycentrics fixed by the
    object-space point; a ray past the quad misses."""
    quad = np.array([[0, 0, 0, 1, 0, 0, 1, 1, 0], [0, 0, 0, 1, 1, 0, 0, 1, 0]], np.float32)
    rot = np.array([[0.0, 1.0, 0.0], [-1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])  # row vectors: x -> y, y -> -x
    lin = np.diag([2.0, 3.0, 4.0]) @ rot
    tw, to = raycast.object_matrices(lin, (5.0, 0.0, -1.0))
    objs = raycast.make_objects([(0, tw, to)])
    pts = np.array([[3.5, 1.0], [2.75, 0.5], [4.25, 1.5], [6.0, 1.0]])
    o = np.column_stack([pts, np.full(len(pts), 3.0)])
    d = np.tile([0.0, 0.0, -1.0], (len(pts), 1))
    rays = scenes.rays_from_rows(o, d)
    for arith in (raycast.ARITH_IEEE, raycast.ARITH_X86_FMA):
        h = raycast.host_trace_objects([quad], objs, rays, arith=arith)
        assert (h["prim_id"][:3] >= 0).all() and h["prim_id"][3] == -1
        assert np.allclose(h["range"][:3], 4.0, rtol=0, atol=1e-9)
        assert np.allclose(h["point"][:3], np.column_stack([pts[:3], np.full(3, -1.0)]), atol=1e-6)
        assert np.allclose(np.abs(h["geom_normal"][:3, 2]), 1.0, atol=1e-7)
        # object-space point: world (x, y) -> local (y / 2, (5 - x) / 3)
        local = np.column_stack([pts[:3, 1] / 2.0, (5.0 - pts[:3, 0]) / 3.0])
        for i in range(3):
            v = quad[h["prim_id"][i]].reshape(3, 3).astype(np.float64)
            p = v[0] + h["alpha"][i] * (v[1] - v[0]) + h["beta"][i] * (v[2] - v[0])
            assert np.allclose(p[:2], local[i], atol=1e-9) and abs(p[2]) < 1e-12
        occ = raycast.host_trace_objects([quad], objs, rays, mode=1, tmax=np.array([4.5, 3.5, 4.5, 9], np.float32),
                                         arith=arith)
        assert occ.tolist() == [1, 0, 1, 0]


def test_object_arguments_are_validated():
    meshes, objs, rays = _scene(tris=50, n_rays=10)
    bad = objs.copy()
    bad[3]["mesh"] = len(meshes)
    with pytest.raises(_lib.RiseHipError, match="mesh index"):
        raycast.host_trace_objects(meshes, bad, rays)
    bad = objs.copy()
    bad[2]["reserved"] = 1
    with pytest.raises(_lib.RiseHipError, match="reserved"):
        raycast.host_trace_objects(meshes, bad, rays)
    bad = objs.copy()
    bad[4]["to_object"][5] = np.nan
    with pytest.raises(_lib.RiseHipError, match="non-finite"):
        raycast.host_trace_objects(meshes, bad, rays)
    empty = raycast.host_trace_objects(meshes, objs[:0], rays)
    assert (empty["prim_id"] == -1).all()


# ---------------------------------------------------------------------------
# The slab test of object-space rays.

def _obj_ray(ray, to_object):
    """o' and dn' exactly as the library and the reference compute them
    (Point3Ops::Transform, Vector3Ops::Transform, Normalize; IEEE double,
    left to right)."""
    m = [float(x) for x in to_object]
    o = [float(x) for x in ray[:3]]
    d = [float(x) for x in ray[3:]]
    op = [m[j] * o[0] + m[4 + j] * o[1] + m[8 + j] * o[2] + m[12 + j] for j in range(3)]
    dl = [m[j] * d[0] + m[4 + j] * d[1] + m[8 + j] * d[2] for j in range(3)]
    mag = (dl[0] * dl[0] + dl[1] * dl[1] + dl[2] * dl[2]) ** 0.5
    ov = 1.0 / mag
    return op, [x * ov for x in dl]


def _accept_obj(node, ray, to_object, best):
    r = (ctypes.c_float * 6)(*[float(x) for x in ray])
    m = (ctypes.c_double * 16)(*[float(x) for x in to_object])
    out = ctypes.c_uint32()
    _lib.check(_lib.lib().rise_hip_debug_node_accept_obj_host(node, r, m, ctypes.c_double(best), ctypes.byref(out)))
    return out.value


def _exact_hit(op, dn, lo, hi, best):
    tn, tf = F(0), None
    for a in range(3):
        oa, da = F(op[a]), F(dn[a])
        if da == 0:
            if not (lo[a] <= oa <= hi[a]):
                return False
            continue
        t1, t2 = (lo[a] - oa) / da, (hi[a] - oa) / da
        if t1 > t2:
            t1, t2 = t2, t1
        tn = max(tn, t1)
        tf = t2 if tf is None else min(tf, t2)
    if tf is not None and tn > tf:
        return False
    return best >= 1e308 or tn <= F(best)


@pytest.mark.parametrize("seed", range(3))
def test_object_slab_test_is_conservative(seed):
    """The FP32 slab test of the float object-space ray, widened by the
    origin's rounding: boxes the exact object-space ray (o', dn') meets within
    `best` are taken; boxes missed by more than 1e-5 of the coordinates and
    the origin's size, or entered beyond best (1 + 1e-5), are not."""
    g = np.random.default_rng(500 + seed)
    must = taken = far = wrongly = 0
    for _ in range(150):
        scale = 10.0 ** g.uniform(-2, 3)
        p = np.float32(g.uniform(-1, 1, 3) * scale)
        e = np.clip(np.round(np.log2(scale)) - 8 + g.integers(-3, 2, 3), -126, 40).astype(int)
        qlo = [[255] * 8 for _ in range(3)]
        qhi = [[0] * 8 for _ in range(3)]
        meta = [0] * 8
        for s in range(8):
            if g.random() < 0.15:
                continue
            meta[s] = 1 << 5
            for a in range(3):
                lo = int(g.integers(0, 255))
                qlo[a][s], qhi[a][s] = lo, int(g.integers(lo, 256))
        w0 = struct.pack("<3f4b", *[float(x) for x in p], *[int(x) for x in e], 0)
        w1 = struct.pack("<II8B", 0, 0, *meta)
        planes = b"".join(bytes(bytearray(qlo[a])) for a in range(3)) + b"".join(bytes(bytearray(qhi[a])) for a in range(3))
        node = w0 + w1 + planes
        occupied = [s for s in range(8) if meta[s]]
        if not occupied:
            continue
        boxes = {s: ([F(float(p[a])) + qlo[a][s] * F(2) ** int(e[a]) for a in range(3)],
                     [F(float(p[a])) + qhi[a][s] * F(2) ** int(e[a]) for a in range(3)]) for s in occupied}
        for _ in range(6):
            tw, to = scenes.random_object_matrix(g, mirror=bool(g.integers(0, 2)), spread=scale)
            s0 = occupied[int(g.integers(0, len(occupied)))]
            lo, hi = boxes[s0]
            target_obj = np.array([float(lo[a] + (hi[a] - lo[a]) * F(int(g.integers(0, 3)), 2)) for a in range(3)])
            m = tw.reshape(4, 4)
            target = target_obj @ m[:3, :3] + m[3, :3]
            o = np.float32(target + g.normal(0, 1, 3) * scale * 10.0 ** g.uniform(-1, 2))
            d = np.float32(target - o.astype(np.float64))
            if not np.any(d):
                continue
            ray = np.concatenate([o, d])
            op, dn = _obj_ray(ray, to)
            best = 1.8e308 if g.random() < 0.5 else float(np.linalg.norm(np.array(op) - target_obj)) * (
                1 + (g.random() - 0.5) * 1e-6)
            got = _accept_obj(node, ray, to, best)
            for s in occupied:
                bl, bh = boxes[s]
                if _exact_hit(op, dn, bl, bh, best):
                    must += 1
                    taken += (got >> s) & 1
                else:
                    gpad = F(1, 10 ** 5) * (max(abs(x) for x in bl + bh) + max(abs(F(x)) for x in op)) + F(1, 10 ** 30)
                    if not _exact_hit(op, dn, [x - gpad for x in bl], [x + gpad for x in bh],
                                      best * (1 + 1e-5) if best < 1e308 else 1.8e308):
                        far += 1
                        wrongly += (got >> s) & 1
            assert not (got & ~sum(1 << s for s in occupied))
    assert must > 150 and far > 150, (must, far)
    assert taken == must, f"{must - taken} of {must} boxes the exact ray meets were rejected"
    assert wrongly == 0, f"{wrongly} of {far} clear misses were taken"
