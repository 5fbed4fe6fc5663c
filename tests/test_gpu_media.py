"""GPU parity for participating media (homogeneous medium, Henyey-Greenstein
phase, shadow-ray transmittance walk, non-shadow-casting shells), through the
C ABI.

* Closed form: a purely absorbing cube inside a uniformly emitting box shows
  exp(-sigma_t * chord) (the image mean against the pixel-footprint average).
* Same random streams as the FP64 oracle: the scattering cube in the emitting
  box and the Cornell box with a floating fog box (sigma_s 0.5, sigma_a 0.05,
  g 0.3).  Free-flight distances and phase directions computed in FP32
  against FP64 send a few paths elsewhere, so the bar is the image mean within
  3e-3 and 95% of pixels within 2e-2 (relative, floored at 1% of the mean).
* Statistics: independent seeds, per-pixel z-test pass rate >= 99%.
* Batching: a fog scene rendered alone and inside a batch with scenes
  without media gives bit-identical films.
"""
import math

import numpy as np
import pytest

import oracle
from rise_amd import pt_scene
from rise_amd.render import PathTracingSpectralRasterizer, RenderBatch
from tests.conftest import log_case

pytestmark = pytest.mark.gpu


def _render(scene, settings, **kw):
    pt = PathTracingSpectralRasterizer(scene)
    film, sq, st = pt.rasterize_scene(settings, want_sq=True, **kw)
    out = film.cpu().numpy().astype(np.float64), sq.cpu().numpy().astype(np.float64), st
    pt.release()
    return out


def _slab_chord(o, d, lo, hi):
    t0, t1 = -np.inf, np.inf
    for k in range(3):
        if abs(d[k]) < 1e-15:
            if o[k] < lo[k] or o[k] > hi[k]:
                return 0.0
            continue
        a, b = (lo[k] - o[k]) / d[k], (hi[k] - o[k]) / d[k]
        t0, t1 = max(t0, min(a, b)), min(t1, max(a, b))
    return max(0.0, t1 - max(t0, 0.0))


def test_gpu_absorbing_cube_closed_form(cuda):
    log_case("absorbing cube in emitting box 12x12 512spp")
    s, r = pt_scene.furnace_medium(sigma_a=(0.7, 0.7, 0.7), sigma_s=(0.0, 0.0, 0.0), half=0.5)
    r.spp = 512
    g, _, st = _render(s, r)
    c = s.camera
    cp = np.array(c.position)
    f = np.array(c.look_at) - cp
    f /= np.linalg.norm(f)
    rt = np.cross(f, c.up)
    rt /= np.linalg.norm(rt)
    up = np.cross(rt, f)
    hh = 2 * math.tan(math.radians(c.fov_deg) / 2)
    sig = 0.7 * (0.2126 + 0.7152 + 0.0722)
    vals = []
    for py in range(r.height):
        for px in range(r.width):
            for a in range(6):
                for b in range(6):
                    sx, sy = px + (a + 0.5) / 6, py + (b + 0.5) / 6
                    d = f + rt * ((sx - 0.5 * r.width) * hh / r.height) + up * (-(sy - 0.5 * r.height) * hh / r.height)
                    d /= np.linalg.norm(d)
                    vals.append(math.exp(-sig * _slab_chord(cp, d, (-0.5,) * 3, (0.5,) * 3)))
    expect = float(np.mean(vals))
    se = 0.5 / math.sqrt(r.spp * r.width * r.height)
    assert abs(g[..., 1].mean() - expect) < 4 * se + 2e-3, (g[..., 1].mean(), expect)
    assert st["paths"] == r.width * r.height * r.spp


@pytest.mark.parametrize("case", ["scatter_cube", "cornell_fog"])
def test_gpu_media_same_streams_match_oracle(cuda, case):
    if case == "scatter_cube":
        s, r = pt_scene.furnace_medium(sigma_a=(0.0, 0.0, 0.0), sigma_s=(1.5, 1.5, 1.5), g=0.3, half=0.6)
        r.spp = 64
    else:
        s, r = pt_scene.cornell_fog()
        r.width = r.height = 48
        r.spp = 16
    log_case(f"{case} {r.width}x{r.height} {r.spp}spp same streams")
    g, _, st = _render(s, r)
    sm, _, cnt = oracle.render(s, r)
    o = sm / r.spp
    assert abs(g[..., 1].mean() / o[..., 1].mean() - 1) < 3e-3, (g[..., 1].mean(), o[..., 1].mean())
    floor = 0.01 * np.abs(o[..., 1]).mean()
    rel = np.abs(g - o) / (np.abs(o) + floor)
    frac = float(np.mean(np.all(rel <= 2e-2, axis=-1)))
    assert frac >= 0.95, frac
    assert abs(st["extension_rays"] / cnt[0] - 1) < 1e-2
    assert st["paths"] == cnt[2]


def test_gpu_cornell_fog_statistical_parity(cuda):
    log_case("cornell fog 48x48 64spp z-test")
    s, r = pt_scene.cornell_fog()
    r.width = r.height = 48
    r.spp = 64
    r.seed = 3
    g, gq, _ = _render(s, r)
    r.seed = 29
    sm, sq, _ = oracle.render(s, r)
    n = r.spp
    o = sm / n
    vo = np.maximum(sq / n - o * o, 0) / n
    vg = np.maximum(gq - g * g, 0) / n
    z = (g - o) / np.sqrt(vg + vo + 1e-30)
    frac = float(np.mean(np.abs(z[..., 1]) < 3.29))
    assert frac >= 0.99, frac


def test_gpu_media_batch_alone_equals_mix(cuda):
    import torch

    log_case("cornell fog alone vs in a batch with cornell and materials")
    fog, rf = pt_scene.cornell_fog()
    rf.width = rf.height = 40
    rf.spp = 8
    plain, rp = pt_scene.cornell_box()
    rp.width = rp.height = 32
    rp.spp = 4
    mats, rm = pt_scene.cornell_materials()
    rm.width = rm.height = 24
    rm.spp = 4
    alone, _, _ = _render(fog, rf)
    pts = [PathTracingSpectralRasterizer(x) for x in (plain, fog, mats)]
    b = RenderBatch([(pts[0], rp), (pts[1], rf), (pts[2], rm)])
    b.launch()
    torch.cuda.synchronize()
    assert np.array_equal(b.films[1].cpu().numpy().astype(np.float64), alone)
    b.release()
    for p in pts:
        p.release()
