"""The allocation-free closest-hit query (rise_hip_intersect_closest_noalloc),
through the C ABI: its ids and distances equal rise_hip_intersect_closest's,
bit for bit, on two calls in a row (the second with fewer rays, so a stale
fetch counter or leftover results from the first call would show), on a
60,000-triangle and a 500-triangle soup and on a scene of transformed objects,
in both arithmetic models.  Bar: ids and distances bit-exact (integer and index work).
"""
import numpy as np
import pytest

from rise_amd import _lib, raycast, scenes
from tests.conftest import log_case

pytestmark = pytest.mark.gpu


def _closest(fn, om, rays, cuda):
    import torch

    n = rays.shape[1]
    pid = torch.full((n,), -7, dtype=torch.int32, device=cuda)
    t = torch.full((n,), -7.0, dtype=torch.float32, device=cuda)
    _lib.check(fn(om.handle, rays.data_ptr(), n, pid.data_ptr(), t.data_ptr(), None))
    torch.cuda.synchronize()
    return pid.cpu().numpy(), t.cpu().numpy()


@pytest.mark.parametrize("arith", [raycast.ARITH_IEEE, raycast.ARITH_X86_FMA])
@pytest.mark.parametrize("n_tris,sigma", [(60000, 0.003), (500, 0.05)])
def test_noalloc_equals_closest_on_two_calls(cuda, n_tris, sigma, arith):
    import torch

    L = _lib.lib()
    tris = scenes.soup_triangles(n_tris, seed=21, sigma=sigma)
    om = raycast.ObjectManager(arith)
    om.add_mesh(tris)
    om.prepare_for_rendering()
    big = np.concatenate([scenes.pinhole_rays(160, 120), scenes.incoherent_rays(60000, 22)], 1)
    small = np.ascontiguousarray(scenes.incoherent_rays(7001, 23))
    for k, rays in enumerate((big, small)):
        log_case(f"no-alloc closest hit: {n_tris} tris, arith {arith}, call {k + 1}, {rays.shape[1]} rays")
        r = torch.from_numpy(np.ascontiguousarray(rays, np.float32)).to(cuda)
        want_p, want_t = _closest(L.rise_hip_intersect_closest, om, r, cuda)
        got_p, got_t = _closest(L.rise_hip_intersect_closest_noalloc, om, r, cuda)
        assert np.array_equal(got_p, want_p)
        assert np.array_equal(got_t.view(np.uint32), want_t.view(np.uint32))
        assert (want_p >= 0).any()
    om.release()


@pytest.mark.parametrize("arith", [raycast.ARITH_IEEE, raycast.ARITH_X86_FMA])
def test_noalloc_equals_closest_on_objects(cuda, arith):
    """A scene of transformed objects: the object kernels take their fetch
    counter from the same scratch array."""
    import torch

    L = _lib.lib()
    meshes, objs = scenes.instanced_scene(tris_per_mesh=3000, seed=25)
    om = raycast.ObjectManager(arith)
    for m in meshes:
        om.add_mesh(m)
    for ob in objs:
        om.add_object(int(ob["mesh"]), ob["to_world"], ob["to_object"])
    om.prepare_for_rendering()
    for k, n in enumerate((30000, 4001)):
        log_case(f"no-alloc closest hit on objects: arith {arith}, call {k + 1}, {n} rays")
        rays = scenes.object_rays(meshes, objs, n, seed=26 + k)
        r = torch.from_numpy(np.ascontiguousarray(rays, np.float32)).to(cuda)
        want_p, want_t = _closest(L.rise_hip_intersect_closest, om, r, cuda)
        got_p, got_t = _closest(L.rise_hip_intersect_closest_noalloc, om, r, cuda)
        assert np.array_equal(got_p, want_p)
        assert np.array_equal(got_t.view(np.uint32), want_t.view(np.uint32))
        assert (want_p >= 0).any()
    om.release()


def test_noalloc_rejects_bad_arguments(cuda):
    L = _lib.lib()
    st = L.rise_hip_intersect_closest_noalloc(None, None, 1, None, None, None)
    assert st != _lib.OK
    assert b"NULL" in L.rise_hip_last_error()
    tris = scenes.soup_triangles(100, seed=2)
    om = raycast.ObjectManager()
    om.add_mesh(tris)
    om.prepare_for_rendering()
    assert L.rise_hip_intersect_closest_noalloc(om.handle, None, -1, None, None, None) != _lib.OK
    assert L.rise_hip_intersect_closest_noalloc(om.handle, None, 0, None, None, None) == _lib.OK
    om.release()
