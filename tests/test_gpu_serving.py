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
* Cancellation and timeouts: a waiting request and one whose passes are in
  flight end as cancelled, one with an elapsed timeout as timed out, and the
  others still finish with films identical to rendering them alone.
* Failed admissions, forced from the host (the next film allocations report
  out-of-memory): the request is re-admitted at a later call and finishes
  with the same film; one that keeps failing while nothing else runs is
  reported failed and the server goes on with the next.
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


def test_server_cancel_and_timeout(cuda):
    import torch

    from rise_amd import render

    log_case("server: cancel a waiting and a rendering request, time one out")
    scenes, reqs = _requests()
    big = pt_scene.RenderSettings(**{**reqs[0][1].__dict__})
    big.width, big.height, big.spp = 512, 512, 32      # 8M paths: 8 chunks, one per pass below
    alone = {k: _alone(*reqs[k])[0] for k in (1, 2, 3)}
    torch.cuda.synchronize()
    sv = RenderServer(tuning={"max_paths": 1 << 20})
    t_big, _, _ = sv.submit(scenes[0], big)
    t_wait, _, _ = sv.submit(*reqs[0])
    sv.cancel(t_wait)                                   # never admitted
    res = []
    for _ in range(3):
        res += sv.poll_results()                        # the big request's passes go in flight
    sv.cancel(t_big)
    kept = {}
    for k in (1, 2, 3):
        t, f, _ = sv.submit(*reqs[k])
        kept[t] = (k, f)
    t_late, _, _ = sv.submit(*reqs[1])
    sv.set_timeout(t_late, 1e-9)                        # already past at the next poll
    res += sv.drain_results()
    torch.cuda.synchronize()
    status = {t: st for t, _, st in res}
    assert status[t_wait] == render.REQUEST_CANCELLED
    assert status[t_big] == render.REQUEST_CANCELLED
    assert status[t_late] == render.REQUEST_TIMED_OUT
    for t, (k, f) in kept.items():
        assert status[t] == render.REQUEST_DONE
        assert np.array_equal(f.cpu().numpy(), alone[k]), k
    st = sv.stats()
    assert (st["cancelled"], st["timed_out"], st["completed"]) == (2, 1, 6)
    assert st["admitted"] == 0 and st["waiting"] == 0 and st["hbm_in_use"] == 0
    with pytest.raises(Exception):
        sv.cancel(t_big)                                # finished: unknown ticket now
    sv.release()
    for sc in scenes:
        sc.release()


def test_server_readmits_after_failed_allocations(cuda):
    import torch

    from rise_amd import render

    log_case("server: film allocations forced to fail from the host")
    scenes, reqs = _requests()
    alone = {k: _alone(*reqs[k])[0] for k in (1, 3)}
    torch.cuda.synchronize()
    sv = RenderServer()
    sv.fail_next_allocations(2)                         # the first request is tried three times
    films = {}
    for k in (1, 3):
        t, f, _ = sv.submit(*reqs[k])
        films[t] = (k, f)
    res = sv.drain_results()
    torch.cuda.synchronize()
    assert sorted(t for t, _, _ in res) == sorted(films)
    for t, _, st in res:
        assert st == render.REQUEST_DONE
        k, f = films[t]
        assert np.array_equal(f.cpu().numpy(), alone[k]), k
    assert sv.stats()["alloc_retries"] == 2
    # a request that fails three times while nothing else runs is reported failed;
    # the next one is admitted (after the remaining forced failure) and completes
    sv.fail_next_allocations(4)
    t_fail, _, _ = sv.submit(*reqs[3])
    t_ok, f_ok, _ = sv.submit(*reqs[1])
    res = {t: st for t, _, st in sv.drain_results()}
    torch.cuda.synchronize()
    assert res[t_fail] == render.REQUEST_FAILED and res[t_ok] == render.REQUEST_DONE
    assert np.array_equal(f_ok.cpu().numpy(), alone[1])
    st = sv.stats()
    assert st["failed"] == 1 and st["alloc_retries"] == 6 and st["hbm_in_use"] == 0
    sv.release()
    for sc in scenes:
        sc.release()
