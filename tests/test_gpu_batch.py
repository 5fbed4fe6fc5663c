"""GPU tests of batched rendering (rise_hip_batch_*), through the C ABI.

* Alone vs in the mix: every request of a batch (different scenes, image
  sizes, sample counts, seeds, bounce caps, progressive sample ranges, one
  request large enough to need several passes) writes a film bit-identical
  to rise_hip_render of that request alone, and the batch's ray counts are
  the sums of the single renders' counts.
* Graph capture: the batch launch captured in a HIP graph and replayed
  several times, with other GPU work between replays, reproduces the eager
  films bit for bit on every replay.
* A batched film also matches the FP64 oracle on the same random streams.
"""
import numpy as np
import pytest

import oracle
from rise_amd import pt_scene
from rise_amd.render import PathTracingSpectralRasterizer, RenderBatch
from tests.conftest import log_case

pytestmark = pytest.mark.gpu


def _mixed_requests():
    cb, r0 = pt_scene.cornell_box()
    fb, rf = pt_scene.furnace_box(albedo=0.5, radiance=1.0, max_bounce=16)
    cl, rc = pt_scene.cloud_cornell(seed=7, n_extra=20000)
    scenes = [PathTracingSpectralRasterizer(s) for s in (cb, fb, cl)]
    reqs = []
    r = pt_scene.RenderSettings(**{**r0.__dict__})
    r.width, r.height, r.spp = 64, 48, 16
    reqs.append((scenes[0], r))
    r = pt_scene.RenderSettings(**{**rf.__dict__})
    r.spp = 32
    reqs.append((scenes[1], r))
    r = pt_scene.RenderSettings(**{**rc.__dict__})
    r.width, r.height, r.spp, r.seed = 96, 80, 4, 99
    reqs.append((scenes[2], r))
    r = pt_scene.RenderSettings(**{**r0.__dict__})
    r.width, r.height, r.spp, r.seed, r.max_diffuse_bounce = 512, 512, 24, 5, 3   # 6.3M paths: several passes
    reqs.append((scenes[0], r))
    r = pt_scene.RenderSettings(**{**rc.__dict__})
    r.width, r.height, r.spp = 33, 17, 8
    reqs.append((scenes[2], r, 4, 4, 8))   # second half of a progressive 8-spp render
    return scenes, reqs


def _alone(item):
    rast, settings = item[0], item[1]
    spp, spp_offset, spp_total = (list(item[2:]) + [None, 0, None][len(item) - 2:])[:3]
    film, _, st = rast.rasterize_scene(settings, spp=spp, spp_offset=spp_offset, spp_total=spp_total)
    return film.cpu().numpy(), st


def test_batch_alone_equals_mix(cuda):
    import torch

    log_case("batch of 5 mixed requests vs each alone")
    scenes, reqs = _mixed_requests()
    alone = [_alone(it) for it in reqs]
    b = RenderBatch(reqs)
    b.launch()
    st = b.stats()
    torch.cuda.synchronize()
    for k, (a, _) in enumerate(alone):
        assert np.array_equal(b.films[k].cpu().numpy(), a), k
    for key in ("extension_rays", "shadow_rays", "paths"):
        assert st[key] == sum(s[key] for _, s in alone), key
    # Launching again gives the same films.
    b.launch()
    torch.cuda.synchronize()
    for k, (a, _) in enumerate(alone):
        assert np.array_equal(b.films[k].cpu().numpy(), a), k
    b.release()
    for s in scenes:
        s.release()


def test_batch_graph_replays(cuda):
    import torch

    log_case("batch captured in a graph, 3 replays")
    scenes, reqs = _mixed_requests()
    reqs = reqs[:3] + reqs[4:]
    b = RenderBatch(reqs)
    b.launch()
    torch.cuda.synchronize()
    eager = [f.cpu().numpy().copy() for f in b.films]
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        b.launch(side)   # warm-up on the capture stream
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=side):
        b.launch(side)
    torch.cuda.synchronize()
    for rep in range(3):
        for f in b.films:
            f.fill_(-1.0)
        x = torch.randn(1 << 20, device="cuda")
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert float(x.abs().sum()) > 0
        for k, e in enumerate(eager):
            assert np.array_equal(b.films[k].cpu().numpy(), e), (rep, k)
    del g
    b.release()
    for s in scenes:
        s.release()


def test_batch_matches_oracle(cuda):
    import torch

    log_case("batched cornell 48x48 vs FP64 oracle, same streams")
    s, r = pt_scene.cornell_box()
    r.width = r.height = 48
    r.spp = 16
    fb, rf = pt_scene.furnace_box()
    a, b_ = PathTracingSpectralRasterizer(s), PathTracingSpectralRasterizer(fb)
    batch = RenderBatch([(b_, rf), (a, r)])
    batch.launch()
    torch.cuda.synchronize()
    g = batch.films[1].cpu().numpy().astype(np.float64)
    sm, _, _ = oracle.render(s, r)
    o = sm / r.spp
    floor = 0.01 * np.abs(o[..., 1]).mean()
    rel = np.abs(g - o) / (np.abs(o) + floor)
    assert float(np.mean(np.all(rel <= 1e-3, axis=-1))) >= 0.99
    assert abs(g[..., 1].mean() / o[..., 1].mean() - 1) < 1e-3
    batch.release()
    a.release()
    b_.release()


def test_batch_tuning_does_not_change_films(cuda):
    import torch

    log_case("batch under three launch tunings")
    scenes, reqs = _mixed_requests()
    ref = RenderBatch(reqs)
    ref.launch()
    torch.cuda.synchronize()
    films = [f.cpu().numpy() for f in ref.films]
    ref.release()
    for tn in ({"persistent": 2, "trace_block": 64, "max_paths": 1 << 20},
               {"persistent": 1, "trace_block": 256, "stack": -1},
               {"persistent": 1, "trace_block": 128, "stack": 16, "max_paths": 3 << 20}):
        b = RenderBatch(reqs, tuning=tn)
        got = b.tuning()
        assert got["trace_block"] == tn["trace_block"] and got["persistent"] == tn["persistent"]
        b.launch()
        torch.cuda.synchronize()
        for k, f in enumerate(films):
            assert np.array_equal(b.films[k].cpu().numpy(), f), (tn, k)
        b.release()
    for s in scenes:
        s.release()
