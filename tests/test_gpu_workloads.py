"""GPU tests of the terrain-and-spheres, batched and serving workloads at their
own sizes, through the C ABI.

* Terrain and spheres (2048x1024 value-noise heightfield, 4,188,162 +
  327,680 triangles, GGX and glass icospheres, spectral sky, 1920x1080,
  64 spp, depth 8): the film is finite and non-negative (the XYZ of physical
  spectra), bit-identical on a rerun; its two progressive halves (samples
  0-31 and 32-63) average to the whole render and their ray counts sum to
  its counts exactly.
* A reduced scene from the same generator (128x64 grid, 4 icospheres)
  against the FP64 oracle on the same random streams: at least 97% of pixels
  within 1e-2 (glossy and refracted paths turn FP32/FP64 differences into
  different hits now and then), image mean within 2e-3, ray counts within 1%.
* The batched workload (64 requests over 8 seeded scenes, 128^2-1024^2,
  1-64 spp, depth 6) in one batch launch: a seeded subset of 6 requests
  rendered alone gives bit-identical films; the batch captured in a HIP graph
  and replayed twice, with other GPU work between replays, reproduces the
  eager films.
* The serving workload (Poisson stream of 64 requests over 16 resident
  scenes of 50k-2M triangles, every other scene with a fog box) submitted at
  its arrival times: every request completes, and a seeded subset of 6
  requests rendered alone gives bit-identical films.
"""
import time

import numpy as np
import pytest

import oracle
from rise_amd import pt_scene
from rise_amd.render import PathTracingSpectralRasterizer, RenderBatch, RenderServer
from tests.conftest import log_case

pytestmark = pytest.mark.gpu


def test_terrain_full_size_properties(cuda):
    import torch

    log_case("terrain and spheres, full size: 4.5M tris, 1920x1080, 64 spp")
    scene, rs = pt_scene.terrain_spheres()
    assert scene.n_tris == 4_515_842
    pt = PathTracingSpectralRasterizer(scene)
    film, _, st = pt.rasterize_scene(rs)
    a = film.cpu().numpy()
    assert np.all(np.isfinite(a)) and a.min() >= 0.0
    assert a[..., 1].mean() > 0 and st["paths"] == 1920 * 1080 * 64
    again, _, _ = pt.rasterize_scene(rs)
    assert np.array_equal(again.cpu().numpy(), a)
    h0, _, s0 = pt.rasterize_scene(rs, spp=32, spp_offset=0, spp_total=64)
    h0 = h0.cpu().numpy()
    h1, _, s1 = pt.rasterize_scene(rs, spp=32, spp_offset=32, spp_total=64)
    h1 = h1.cpu().numpy()
    for key in ("extension_rays", "shadow_rays", "paths"):
        assert s0[key] + s1[key] == st[key], key
    assert np.allclose(0.5 * (h0 + h1), a, rtol=1e-5, atol=1e-7)
    torch.cuda.synchronize()
    pt.release()


def test_reduced_terrain_matches_oracle(cuda):
    log_case("reduced terrain (128x64 grid, 4 icospheres) 96x54 8 spp vs FP64 oracle, same streams")
    s, r = pt_scene.terrain_spheres(128, 64, 2, n_side=2)
    r.width, r.height, r.spp = 96, 54, 8
    pt = PathTracingSpectralRasterizer(s)
    film, _, st = pt.rasterize_scene(r)
    g = film.cpu().numpy().astype(np.float64)
    pt.release()
    sm, _, cnt = oracle.render(s, r)
    o = sm / r.spp
    floor = 0.01 * np.abs(o[..., 1]).mean()
    rel = np.abs(g - o) / (np.abs(o) + floor)
    frac = float(np.mean(np.all(rel <= 1e-2, axis=-1)))
    assert frac >= 0.97, frac
    assert abs(g[..., 1].mean() / o[..., 1].mean() - 1) < 2e-3
    assert abs(st["extension_rays"] / cnt[0] - 1) < 1e-2
    assert st["paths"] == cnt[2]


def _subset(n, k, seed):
    return sorted(np.random.default_rng(seed).choice(n, size=k, replace=False).tolist())


def test_batch_workload_alone_equals_mix_and_graph(cuda):
    import torch

    log_case("batched workload: 64 requests over 8 scenes in one launch")
    scenes, reqs = pt_scene.batch_workload()
    rasts = [PathTracingSpectralRasterizer(s) for s in scenes]
    items = [(rasts[k], r) for k, r in reqs]
    b = RenderBatch(items)
    b.launch()
    torch.cuda.synchronize()
    films = [f.cpu().numpy().copy() for f in b.films]
    assert all(np.all(np.isfinite(f)) for f in films)
    for i in _subset(len(items), 6, 41):
        log_case(f"batched workload: request {i} alone ({reqs[i][1].width}^2, {reqs[i][1].spp} spp)")
        alone, _, _ = items[i][0].rasterize_scene(items[i][1])
        assert np.array_equal(alone.cpu().numpy(), films[i]), i
    log_case("batched workload: graph capture and two replays")
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        b.launch(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=side):
        b.launch(side)
    torch.cuda.synchronize()
    for rep in range(2):
        for f in b.films:
            f.fill_(-1.0)
        x = torch.randn(1 << 20, device="cuda")
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert float(x.abs().sum()) > 0
        for k, e in enumerate(films):
            assert np.array_equal(b.films[k].cpu().numpy(), e), (rep, k)
    del g
    b.release()
    for r in rasts:
        r.release()


def test_serving_workload_stream(cuda):
    import torch

    log_case("serving workload: Poisson stream of 64 requests over 16 scenes, half with fog")
    scenes, reqs = pt_scene.serving_workload()
    assert sum(1 for s in scenes if s.media) == 8
    rasts = [PathTracingSpectralRasterizer(s) for s in scenes]
    sv = RenderServer()
    films, lat = {}, {}
    pending = list(enumerate(reqs))
    t0 = time.perf_counter()
    while len(lat) < len(reqs):
        now = time.perf_counter() - t0
        while pending and pending[0][1][0] <= now:
            i, (_, k, r) = pending.pop(0)
            tk, f, _ = sv.submit(rasts[k], r)
            films[tk] = (i, f)
        for tk, l in sv.poll():
            lat[tk] = l
        assert time.perf_counter() - t0 < 300.0
    torch.cuda.synchronize()
    st = sv.stats()
    assert st["completed"] == len(reqs) and st["waiting"] == 0
    by_req = {i: f.cpu().numpy() for i, f in films.values()}
    sv.release()
    for i in _subset(len(reqs), 6, 43):
        _, k, r = reqs[i]
        log_case(f"serving workload: request {i} alone (scene {k}, {r.width}^2, {r.spp} spp)")
        alone, _, _ = rasts[k].rasterize_scene(r)
        assert np.array_equal(alone.cpu().numpy(), by_req[i]), i
    for r in rasts:
        r.release()
