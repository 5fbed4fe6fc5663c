"""GPU tests of the continuous-batching render server (rise_hip_server_*),
through the C ABI.

* Alone vs in the mix: requests submitted while others are rendering (their
  passes already in flight) produce films bit-identical to rise_hip_render of
  each request alone, and the server's ray counts are the sums of the single
  renders' counts.
* Admission by HBM budget: with a budget that holds one request's film
  accumulators at a time, requests wait, are admitted as memory frees up, and
  still finish with identical films; the memory in use never exceeds the
  budget while more than one request is admitted.
"""
import numpy as np
import pytest

from rise_amd import pt_scene
from rise_amd.render import PathTracingSpectralRasterizer, RenderServer
from tests.conftest import log_case

pytestmark = pytest.mark.gpu


def _requests():
    cb, r0 = pt_scene.cornell_box()
    cl, rc = pt_scene.cloud_cornell(seed=11, n_extra=30000)
    ms, rm = pt_scene.cornell_materials()
    scenes = [PathTracingSpectralRasterizer(s) for s in (cb, cl, ms)]
    reqs = []
    r = pt_scene.RenderSettings(**{**r0.__dict__})
    r.width, r.height, r.spp = 256, 256, 32          # 2M paths: two chunks
    reqs.append((scenes[0], r))
    r = pt_scene.RenderSettings(**{**rc.__dict__})
    r.width, r.height, r.spp, r.seed = 128, 96, 8, 21
    reqs.append((scenes[1], r))
    r = pt_scene.RenderSettings(**{**rm.__dict__})
    r.width, r.height, r.spp = 64, 64, 16
    reqs.append((scenes[2], r))
    r = pt_scene.RenderSettings(**{**r0.__dict__})
    r.width, r.height, r.spp, r.seed = 40, 30, 4, 8
    reqs.append((scenes[0], r))
    return scenes, reqs


def _alone(rast, settings):
    film, _, st = rast.rasterize_scene(settings)
    return film.cpu().numpy(), st


def test_server_alone_equals_mix(cuda):
    import torch

    log_case("server: 4 requests submitted while others render")
    scenes, reqs = _requests()
    alone = [_alone(*it) for it in reqs]
    torch.cuda.synchronize()
    sv = RenderServer()
    films = {}
    t, f, _ = sv.submit(*reqs[0])
    films[t] = (0, f)
    done = []
    for _ in range(3):
        done += sv.poll()          # passes of the first request go in flight
    for k in (1, 2, 3):
        t, f, _ = sv.submit(*reqs[k])
        films[t] = (k, f)
        done += sv.poll()
    done += sv.drain()
    torch.cuda.synchronize()
    assert sorted(t for t, _ in done) == sorted(films)
    assert all(lat > 0 for _, lat in done)
    for t, (k, f) in films.items():
        assert np.array_equal(f.cpu().numpy(), alone[k][0]), k
    st = sv.stats()
    assert st["completed"] == 4 and st["waiting"] == 0 and st["admitted"] == 0
    for key in ("extension_rays", "shadow_rays", "paths"):
        assert st[key] == sum(a[1][key] for a in alone), key
    sv.release()
    for s in scenes:
        s.release()


def test_server_admission_by_budget(cuda):
    import torch

    log_case("server: HBM budget admits one request at a time")
    scenes, reqs = _requests()
    alone = [_alone(*it) for it in reqs]
    biggest = max(12 * 4 * r.width * r.height for _, r in reqs)
    sv = RenderServer(hbm_budget_bytes=biggest)
    films = {}
    for k, it in enumerate(reqs):
        t, f, _ = sv.submit(*it)
        films[t] = (k, f)
    done = []
    seen_waiting = False
    for _ in range(10000):
        done += sv.poll()
        st = sv.stats()
        seen_waiting = seen_waiting or st["waiting"] > 0
        if st["admitted"] > 1:
            assert st["hbm_in_use"] <= biggest
        if len(done) == len(films):
            break
    done += sv.drain()
    torch.cuda.synchronize()
    assert seen_waiting
    assert sorted(t for t, _ in done) == sorted(films)
    for t, (k, f) in films.items():
        assert np.array_equal(f.cpu().numpy(), alone[k][0]), k
    sv.release()
    for s in scenes:
        s.release()
