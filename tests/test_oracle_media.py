"""The FP64 oracle's participating media (HomogeneousMedium,
HenyeyGreensteinPhaseFunction, the shadow-ray transmittance walk and the
medium transport of the single-wavelength loop), pinned by closed forms:

* HG: the pdf integrates to 1 with mean cosine g; sampled directions follow
  the closed-form CDF of cos(theta) (Kolmogorov-Smirnov), their azimuth is
  uniform, and the azimuth convention is the one the reference's Perturb
  matrices give for an axis-aligned direction (derived by hand below).
* Walk: exp(-sigma_t * chord) through one box, from inside a box, through
  nested boxes, the luminance mapping of RGB coefficients, and the
  16-boundary cap of EvalShadowTransmittanceNM.
* Transport: a purely absorbing cube inside a uniformly emitting box shows
  exp(-sigma_t * chord) per pixel (free flight, survival weight, the IOR
  stack through the index-matched shell).  A purely scattering cube would
  leave the uniform field at radiance 1, but the reference's estimator adds
  more: its in-scattering NEE is weighted by the power heuristic against the
  phase pdf, while the phase-sampled path reaches the light through the delta
  shell, which resets the pdf to 0 and gives the emission full weight.  The
  oracle is checked against an independent vectorized estimate of exactly
  that estimator (analytic cube geometry instead of triangles), which pins
  the scatter weight, phase sampling, the NEE with its MIS weight and walk
  transmittance together.
"""
import math

import numpy as np
import pytest

import oracle
from rise_amd import pt_scene


@pytest.mark.parametrize("g", [-0.7, 0.0, 0.3, 0.9])
def test_hg_normalization_and_mean_cosine(g):
    from scipy import integrate

    total = 2 * math.pi * integrate.quad(lambda m: oracle.hg_eval(m, g), -1, 1, points=[0.9, 0.99], limit=200)[0]
    mean = 2 * math.pi * integrate.quad(lambda m: m * oracle.hg_eval(m, g), -1, 1, points=[0.9, 0.99], limit=200)[0]
    assert abs(total - 1.0) < 1e-9
    assert abs(mean - g) < 1e-9


def _hg_cdf(mu, g):
    # P(cos <= mu) for the HG phase function, integrated in closed form.
    if abs(g) < 1e-10:
        return (mu + 1.0) / 2.0
    return (1 - g * g) / (2 * g) * (1.0 / np.sqrt(1 + g * g - 2 * g * mu) - 1.0 / (1 + g))


@pytest.mark.parametrize("g", [-0.5, 0.0, 0.3, 0.8])
@pytest.mark.parametrize("wo", [(0.0, 0.0, 1.0), (1.0, 0.0, 0.0), (0.0, -1.0, 0.0), (0.48, -0.6, 0.64)])
def test_hg_sampling_distribution(g, wo):
    from scipy import stats

    rng = np.random.default_rng(11)
    wo = np.array(wo) / np.linalg.norm(wo)
    n = 4000
    u = rng.random((n, 2))
    ws = np.array([oracle.hg_sample(wo, g, a, b) for a, b in u])
    assert np.allclose(np.linalg.norm(ws, axis=1), 1.0, atol=1e-12)
    mu = ws @ wo
    ks = stats.kstest(mu, lambda x: _hg_cdf(np.clip(x, -1, 1), g))
    assert ks.pvalue > 1e-3, ks
    # azimuth about wo is uniform
    ref = np.array([1.0, 0.0, 0.0]) if abs(wo[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    e1 = ref - wo * (ref @ wo)
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(wo, e1)
    phi = np.arctan2(ws @ e2, ws @ e1)
    ks = stats.kstest(phi, stats.uniform(-math.pi, 2 * math.pi).cdf)
    assert ks.pvalue > 1e-3, ks


def test_hg_azimuth_convention():
    # Perturb((0,0,1), theta, phi): CreateFromU gives U = (0,0,1), V =
    # Perpendicular(U) = (1,0,0), W = U x V = (0,1,0); the Y then X rotations
    # of MatricesOps.h (row vectors) take (1,0,0) to (cos t, sin t sin p,
    # -sin t cos p), i.e. (sin t sin p, -sin t cos p, cos t) in world space.
    g, u1 = 0.0, 0.3   # g = 0: cos t = 1 - 2 u1
    ct = 1 - 2 * u1
    st = math.sqrt(1 - ct * ct)
    for u2 in (0.0, 0.25, 0.6):
        p = 2 * math.pi * u2
        w = oracle.hg_sample((0.0, 0.0, 1.0), g, u1, u2)
        assert np.allclose(w, [st * math.sin(p), -st * math.cos(p), ct], atol=1e-12), (u2, w)


def _slab_chord(o, d, lo, hi):
    t0, t1 = -np.inf, np.inf
    for k in range(3):
        if abs(d[k]) < 1e-15:
            if o[k] < lo[k] or o[k] > hi[k]:
                return 0.0
            continue
        a, b = (lo[k] - o[k]) / d[k], (hi[k] - o[k]) / d[k]
        t0, t1 = max(t0, min(a, b)), min(t1, max(a, b))
    t0 = max(t0, 0.0)
    return max(0.0, t1 - t0)


def _sigma_t(a, s):
    # the scene's coefficients travel as float32; Y(sigma_a + sigma_s)
    t = [float(np.float32(x)) + float(np.float32(y)) for x, y in zip(a, s)]
    return 0.2126 * t[0] + 0.7152 * t[1] + 0.0722 * t[2]


def test_shadow_walk_closed_forms():
    s, _ = pt_scene.furnace_medium(sigma_a=(0.3, 0.3, 0.3), sigma_s=(0.2, 0.2, 0.2), half=0.5)
    shell = len(s.objects) - 1
    st = _sigma_t((0.3,) * 3, (0.2,) * 3)
    assert abs(st - 0.5) < 1e-7
    # through the cube along z: chord 1
    tr = oracle.shadow_transmittance(s, (0.1, 0.05, 1.5), (0, 0, -1), 3.0)
    assert abs(tr - math.exp(-st)) < 1e-9
    # oblique
    d = np.array([0.3, -0.2, -1.0])
    d /= np.linalg.norm(d)
    o = np.array([-0.2, 0.15, 1.5])
    chord = _slab_chord(o, d, (-0.5,) * 3, (0.5,) * 3)
    assert chord > 0.5
    assert abs(oracle.shadow_transmittance(s, o, d, 3.0) - math.exp(-st * chord)) < 1e-9
    # from inside the medium to beyond it, and ending inside it
    assert abs(oracle.shadow_transmittance(s, (0, 0, 0), (0, 0, 1), 1.0, shell) - math.exp(-st * 0.5)) < 1e-9
    assert abs(oracle.shadow_transmittance(s, (0, 0, 0), (0, 0, 1), 0.3, shell) - math.exp(-st * 0.3)) < 1e-9
    # missing the cube
    assert oracle.shadow_transmittance(s, (0.9, 0.9, 1.5), (0, 0, -1), 3.0) == 1.0
    # RGB coefficients: the luminance of sigma_a + sigma_s
    s2, _ = pt_scene.furnace_medium(sigma_a=(0.1, 0.2, 0.3), sigma_s=(0.0, 0.0, 0.0))
    st = _sigma_t((0.1, 0.2, 0.3), (0, 0, 0))
    assert abs(st - (0.2126 * 0.1 + 0.7152 * 0.2 + 0.0722 * 0.3)) < 1e-8
    assert abs(oracle.shadow_transmittance(s2, (0, 0, 1.5), (0, 0, -1), 3.0) - math.exp(-st)) < 1e-9


def test_shadow_walk_nested_and_step_cap():
    s, _ = pt_scene.furnace_medium(sigma_a=(0.5, 0.5, 0.5), sigma_s=(0.0, 0.0, 0.0), half=0.5)
    s.media.append(pt_scene.Medium((2.0, 2.0, 2.0), (0.0, 0.0, 0.0), 0.0))
    s.objects.append(pt_scene.medium_shell((-0.25,) * 3, (0.25,) * 3, 1, 1))
    tr = oracle.shadow_transmittance(s, (0, 0, 1.5), (0, 0, -1), 3.0)
    assert abs(tr - math.exp(-_sigma_t((0.5,) * 3, (0,) * 3) * 0.5 - _sigma_t((2.0,) * 3, (0,) * 3) * 0.5)) < 1e-9
    # ten thin slabs of sigma_t 1 and width 0.1 in a row: 20 boundaries, the
    # walk stops after 16 (MAX_WALK_STEPS), so only 8 slabs attenuate.
    s3, _ = pt_scene.furnace_medium(sigma_a=(1.0, 1.0, 1.0), sigma_s=(0.0, 0.0, 0.0))
    s3.objects = s3.objects[:6]
    for k in range(10):
        z0 = 1.0 - 0.2 * k
        s3.objects.append(pt_scene.medium_shell((-0.5, -0.5, z0 - 0.1), (0.5, 0.5, z0), 0, 1))
    tr = oracle.shadow_transmittance(s3, (0, 0, 1.5), (0, 0, -1), 3.4)
    assert abs(tr - math.exp(-_sigma_t((1.0,) * 3, (0,) * 3) * 0.8)) < 1e-6   # float32 slab faces


def _pixel_dirs(scene, r, sub=8):
    c = scene.camera
    cp = np.array(c.position)
    f = np.array(c.look_at) - cp
    f /= np.linalg.norm(f)
    rt = np.cross(f, c.up)
    rt /= np.linalg.norm(rt)
    up = np.cross(rt, f)
    hh = 2 * math.tan(math.radians(c.fov_deg) / 2)
    out = np.zeros((r.height, r.width, sub * sub, 3))
    for py in range(r.height):
        for px in range(r.width):
            k = 0
            for a in range(sub):
                for b in range(sub):
                    sx, sy = px + (a + 0.5) / sub, py + (b + 0.5) / sub
                    cx = (sx - 0.5 * r.width) * hh / r.height
                    cy = -(sy - 0.5 * r.height) * hh / r.height
                    d = f + rt * cx + up * cy
                    out[py, px, k] = d / np.linalg.norm(d)
                    k += 1
    return cp, out


def test_absorbing_cube_transmittance():
    s, r = pt_scene.furnace_medium(sigma_a=(0.7, 0.7, 0.7), sigma_s=(0.0, 0.0, 0.0), half=0.5)
    r.spp = 256
    sm, _, cnt = oracle.render(s, r)
    y = sm[..., 1] / r.spp
    cp, dirs = _pixel_dirs(s, r)
    st = _sigma_t((0.7,) * 3, (0,) * 3)
    expect = np.zeros((r.height, r.width))
    for py in range(r.height):
        for px in range(r.width):
            expect[py, px] = np.mean([math.exp(-st * _slab_chord(cp, d, (-0.5,) * 3, (0.5,) * 3))
                                      for d in dirs[py, px]])
    # each wavelength lane survives or not: per-sample Y variance <= 1/4
    se = 0.5 / math.sqrt(r.spp * r.width * r.height)
    assert abs(y.mean() - expect.mean()) < 4 * se, (y.mean(), expect.mean())
    assert np.all(np.abs(y - expect) < 0.2)
    assert expect.min() < 0.6   # the cube is in view
    assert cnt[2] == r.width * r.height * r.spp


def _exit_dist(x, d, half):
    """Distance from points x (inside the cube |x| <= half) along unit d to its boundary."""
    with np.errstate(divide="ignore", invalid="ignore"):
        t = np.where(d > 0, (half - x) / d, np.where(d < 0, (-half - x) / d, np.inf))
    return t.min(axis=1)


def _cube_entry(o, d, half):
    """Entry and exit distances of rays o + t d into the cube (nan where missed)."""
    with np.errstate(divide="ignore", invalid="ignore"):
        a = (-half - o) / d
        b = (half - o) / d
    t0 = np.minimum(a, b).max(axis=1)
    t1 = np.maximum(a, b).min(axis=1)
    hit = (t1 > np.maximum(t0, 0))
    return np.where(hit, t0, np.nan), np.where(hit, t1, np.nan)


def _hg_dirs(wo, g, rng):
    n = len(wo)
    u1, u2 = rng.random(n), rng.random(n)
    if abs(g) < 1e-10:
        ct = 1 - 2 * u1
    else:
        inner = (1 - g * g) / (1 - g + 2 * g * u1)
        ct = (1 + g * g - inner * inner) / (2 * g)
    ct = np.clip(ct, -1, 1)
    st = np.sqrt(1 - ct * ct)
    ph = 2 * math.pi * u2
    a = np.where(np.abs(wo[:, :1]) < 0.9, np.array([[1.0, 0, 0]]), np.array([[0, 1.0, 0]]))
    e1 = a - wo * np.sum(a * wo, axis=1, keepdims=True)
    e1 /= np.linalg.norm(e1, axis=1, keepdims=True)
    e2 = np.cross(wo, e1)
    return ct[:, None] * wo + (st * np.cos(ph))[:, None] * e1 + (st * np.sin(ph))[:, None] * e2


def _hg(c, g):
    d = 1 + g * g - 2 * g * c
    return (1 - g * g) / (4 * math.pi * d * np.sqrt(d))


def _furnace_estimate(scene, r, sigma_t, sigma_s, g, half, n, seed):
    """Independent vectorized estimate (analytic cube, no triangles) of the
    reference's estimator in furnace_medium: camera rays that miss the cube see
    radiance 1; inside, free flight, weight sigma_s / sigma_t, in-scattering
    NEE to the six emitting faces (pdf 1/6 each, uniform point, power
    heuristic against the HG pdf, transmittance to the cube boundary), HG
    continuation; a path leaving the cube crosses the delta shell and takes the
    wall's emission with full weight.  With sigma_a = 0 no Russian roulette
    fires (importance stays 1).  Returns per-path estimates."""
    rng = np.random.default_rng(seed)
    c = scene.camera
    cp = np.array(c.position, float)
    f = np.array(c.look_at) - cp
    f /= np.linalg.norm(f)
    rt = np.cross(f, c.up)
    rt /= np.linalg.norm(rt)
    up = np.cross(rt, f)
    hh = 2 * math.tan(math.radians(c.fov_deg) / 2)
    sx = rng.random(n) * r.width
    sy = rng.random(n) * r.height
    d = (f[None] + rt[None] * ((sx - 0.5 * r.width) * hh / r.height)[:, None]
         + up[None] * (-(sy - 0.5 * r.height) * hh / r.height)[:, None])
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = np.broadcast_to(cp, d.shape).copy()
    t0, _ = _cube_entry(o, d, half)
    total = np.ones(n)
    inside = ~np.isnan(t0)
    x = o[inside] + d[inside] * t0[inside, None]
    w = d[inside]
    idx = np.nonzero(inside)[0]
    nee = np.zeros(n)
    vb = np.zeros(len(idx), int)
    # the six emitting faces of the half-2 box, normals inward, area 16 each
    normals = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], float)
    while len(idx):
        te = _exit_dist(x, w, half)
        tt = -np.log(np.maximum(1 - rng.random(len(idx)), 1e-30)) / sigma_t
        sc = (tt < te) & (vb < 64)
        # leaving: radiance 1 (full weight after the delta shell); counted in `total` already
        idx, x, w, tt, vb = idx[sc], x[sc], w[sc], tt[sc], vb[sc]
        if not len(idx):
            break
        x = x + w * tt[:, None]
        # NEE: uniform face, uniform point on it
        fi = rng.integers(0, 6, len(idx))
        nrm = normals[fi]
        uv = rng.uniform(-2, 2, (len(idx), 2))
        pl = np.zeros((len(idx), 3))
        for k in range(3):
            ax = np.abs(nrm[:, k]) > 0
            pl[ax, k] = -2.0 * nrm[ax, k]
            others = [j for j in range(3) if j != k]
            pl[ax, others[0]] = uv[ax, 0]
            pl[ax, others[1]] = uv[ax, 1]
        to = pl - x
        dist = np.linalg.norm(to, axis=1)
        to /= dist[:, None]
        cos_l = np.sum(-to * nrm, axis=1)
        ph = _hg(np.sum(to * w, axis=1), g)
        p_light = (1 / 6) * dist * dist / (16 * cos_l)
        wgt = p_light ** 2 / (p_light ** 2 + ph ** 2)
        tr = np.exp(-sigma_t * _exit_dist(x, to, half))
        contrib = 1.0 * (16 * cos_l / (dist * dist)) * ph * tr * wgt / (1 / 6)
        contrib *= sigma_s / sigma_t   # the scatter weight (1 for sigma_a = 0)
        np.add.at(nee, idx, contrib)
        w = _hg_dirs(w, g, rng)
        vb = vb + 1
    return total + nee


@pytest.mark.parametrize("g", [0.3, -0.6])
def test_scattering_cube_matches_independent_estimate(g):
    # Purely scattering cube in the uniformly emitting box.  The physical
    # answer is radiance 1; the reference's estimator also adds the
    # in-scattering NEE share of light that the phase-sampled path reaches
    # through the delta shell with full weight.  The oracle must agree with an
    # independent estimate of exactly that estimator.
    sig = 1.5
    s, r = pt_scene.furnace_medium(sigma_a=(0.0, 0.0, 0.0), sigma_s=(sig,) * 3, g=g, half=0.6)
    r.spp = 128
    sm, sq, cnt = oracle.render(s, r)
    n = r.spp * r.width * r.height
    y = sm[..., 1].sum() / n
    se_o = math.sqrt(max(sq[..., 1].sum() / n - y * y, 1e-12) / n)
    st = _sigma_t((0,) * 3, (sig,) * 3)
    tot = _furnace_estimate(s, r, st, st, g, 0.6, 400000, 5)
    se_e = tot.std() / math.sqrt(len(tot))
    assert abs(y - tot.mean()) < 4 * math.sqrt(se_o ** 2 + se_e ** 2), (y, tot.mean(), se_o, se_e)
    assert tot.mean() > 1.05   # the double-counted share is visible
    assert cnt[1] > 0
