""".risemesh files on the GPU, through the C ABI.

* A file the reference wrote (tests/golden/ref_soup_300.risemesh) becomes a
  ray-cast scene whose 8-wide tree is collapsed from the file's BVH2 cache
  (stats bvh_source = 1); its closest and any hits equal the brute-force
  oracle's, ids and t bit-exact.
* The Cornell box whose boxes come from the reference-written
  unit_cube.risemesh (their subtrees from its cache, the rest built) renders
  bit-identically to the same scene from RAW2 meshes: the tree changes the
  work, never the hits.
"""
from pathlib import Path

import numpy as np
import pytest

import oracle
from rise_amd import scenes
from rise_amd.risemesh import RiseMesh
from rise_amd.scene_file import SceneFile
from tests.conftest import log_case

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


def test_gpu_ray_cast_from_reference_cache(cuda):
    import torch

    log_case("risemesh ref_soup_300 ray cast")
    m = RiseMesh(GOLDEN / "ref_soup_300.risemesh")
    tris = m.triangles()
    rays = np.ascontiguousarray(np.concatenate([scenes.pinhole_rays(64, 48), scenes.incoherent_rays(4000, 9)], 1),
                                dtype=np.float32)
    om = m.object_manager()
    assert om.stats()["bvh_source"] == 1
    r = torch.from_numpy(rays).to(cuda)
    pid, t = om.intersect_ray(r)
    torch.cuda.synchronize()
    bid, bt = oracle.closest(tris, rays)[:2]
    np.testing.assert_array_equal(pid.cpu().numpy(), bid)
    np.testing.assert_array_equal(t.cpu().numpy(), bt)
    tmax = np.where(np.isfinite(bt), bt * 0.999, 10.0).astype(np.float32)
    occ = om.intersect_shadow_ray(r, torch.from_numpy(tmax).to(cuda))
    np.testing.assert_array_equal(occ.cpu().numpy(), oracle.any_hit(tris, rays, tmax))
    om.release()


def test_gpu_scene_with_risemesh_renders_like_raw2(cuda):
    log_case("scene file cornell_risemesh vs cornell")
    films = []
    for name in ("cornell.RISEscene", "cornell_risemesh.RISEscene"):
        f = SceneFile(GOLDEN / "scenes" / name)
        r = f.settings()
        r.width, r.height, r.spp = 96, 96, 4
        pt = f.rasterizer()
        film, _, st = pt.rasterize_scene(r)
        films.append((film.cpu().numpy(), st))
        pt.release()
    assert films[0][1] == films[1][1]
    np.testing.assert_array_equal(films[0][0], films[1][0])


def test_gpu_scene_cache_round_trip(cuda, tmp_path):
    """A ray-cast scene saved as its built 8-wide tree loads with no build and
    answers bit-identically to the scene it was saved from and to the oracle."""
    import torch

    from rise_amd.raycast import ObjectManager

    log_case("scene cache round trip")
    tris = scenes.soup_triangles(30000, seed=8, sigma=0.01)
    rays = np.ascontiguousarray(np.concatenate([scenes.pinhole_rays(64, 48), scenes.incoherent_rays(4000, 10)], 1),
                                dtype=np.float32)
    om = ObjectManager()
    om.add_mesh(tris)
    om.prepare_for_rendering()
    om.save_cache(tmp_path / "s.rhbvh")
    om2 = ObjectManager.load_cache(tmp_path / "s.rhbvh")
    st, st2 = om.stats(), om2.stats()
    assert st2["bvh_source"] == 2 and (st2["n_nodes"], st2["n_tris"], st2["depth"]) == (
        st["n_nodes"], st["n_tris"], st["depth"])
    r = torch.from_numpy(rays).to(cuda)
    a, ta = om.intersect_ray(r)
    b, tb = om2.intersect_ray(r)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(a.cpu().numpy(), b.cpu().numpy())
    np.testing.assert_array_equal(ta.cpu().numpy(), tb.cpu().numpy())
    np.testing.assert_array_equal(b.cpu().numpy(), oracle.closest(tris, rays)[0])
    om.release()
    om2.release()
