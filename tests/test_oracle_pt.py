"""Pins the path-tracing oracle (oracle/pt_oracle.c).

* Colour: the CIE Y integral over 380-780 nm is the value the reference
  states for its table (~106.86, PixelBasedSpectralIntegratingRasterizer.cpp:58-62),
  XYZFromNM reproduces the table at its nodes and interpolates linearly.
* Random numbers: uniform in [0, 1) with 24 bits, distinct per path and dim;
  the Owen-scrambled Sobol sampler's pixel blocks.
* White furnace (closed form L = Le / (1 - albedo)) with MIS, NEE, Russian
  roulette and the Y normalization all active.
* Directly viewed emitter: Y = Le (no MIS at the camera vertex).
* Reciprocity of the image under a seed change (statistical agreement of
  two independent renders).
* Finding: the reference's padded Sobol sampler, unshuffled, is biased on
  the spectral furnace; the shuffled default is not.
Standard errors of image means are taken from the spread of the per-pixel
means (each pixel has its own scramble, so pixels are independent), since
the per-sample variance misstates the error of quasi-random sampling.
"""
import math

import numpy as np
import pytest

import oracle
from rise_amd import pt_scene
from tests import pt_closed_forms


def test_cie_y_integral_and_table():
    assert abs(oracle.cie_y_integral(380, 780) - 106.86) < 0.01
    ok, xyz = oracle.xyz_from_nm(555.0)
    assert ok and abs(xyz[1] - 1.0) < 1e-12
    ok, a = oracle.xyz_from_nm(550.0)
    ok, b = oracle.xyz_from_nm(560.0)
    ok, m = oracle.xyz_from_nm(552.5)
    assert np.allclose(m, a + (oracle.xyz_from_nm(555.0)[1] - a) * 0.5)
    assert not oracle.xyz_from_nm(379.0)[0] and not oracle.xyz_from_nm(781.0)[0]
    # The trapezoid sum equals the integral of the interpolated table.
    xs = np.linspace(380, 720, 340 * 64 + 1)
    ys = np.array([oracle.xyz_from_nm(x)[1][1] for x in xs[::64]])
    assert abs(np.trapezoid(ys, xs[::64]) - oracle.cie_y_integral(380, 720)) < 1e-9


def test_rng_uniform_24_bit():
    u = np.array([oracle.rng(7, p, d) for p in range(200) for d in range(50)])
    assert u.min() >= 0.0 and u.max() < 1.0
    assert np.all(np.abs(u * 16777216.0 - np.round(u * 16777216.0)) == 0)
    assert abs(u.mean() - 0.5) < 0.01
    assert len(np.unique(u)) > 0.99 * u.size
    assert oracle.rng(1, 5, 3) != oracle.rng(2, 5, 3)


def _y_stats(sm, sq, n):
    """Image mean of Y and its standard error from the per-pixel means."""
    pm = sm[..., 1] / (n / sm[..., 1].size)
    return float(pm.mean()), float(pm.std(ddof=1) / math.sqrt(pm.size))


def test_white_furnace_closed_form():
    s, r = pt_scene.furnace_box(albedo=0.5, radiance=1.0, max_bounce=64)
    sm, sq, cnt = oracle.render(s, r)
    n = r.width * r.height * r.spp
    y, se = _y_stats(sm, sq, n)
    assert abs(y - 2.0) < 4 * se + 1e-3, (y, se)
    # Russian roulette keeps the estimate unbiased when it starts at depth 0.
    r.rr_min_depth = 0
    r.seed = 5
    sm, sq, cnt2 = oracle.render(s, r)
    y0, se0 = _y_stats(sm, sq, n)
    assert abs(y0 - 2.0) < 4 * se0 + 1e-3, (y0, se0)
    assert cnt2[0] < cnt[0]


def test_spectral_furnace_closed_form():
    """Step-spectrum albedo and emission (not flat): each companion wavelength
    must carry its own albedo and emission through NEE, emission MIS,
    scattering and Russian roulette.  The closed form is far (> 5%) from what
    an estimator carrying the hero wavelength's radiance on every companion
    would give, so that mistake fails here."""
    rho, le = pt_closed_forms.spectral_furnace_spectra()
    want = pt_closed_forms.furnace_y(rho, le)
    bug = pt_closed_forms.furnace_y_hero_only(rho, le)
    assert abs(bug / want - 1.0) > 0.05, (want, bug)
    s, r = pt_scene.furnace_box(albedo=rho, radiance=le, max_bounce=96)
    r.spp = 512
    sm, sq, _ = oracle.render(s, r)
    y, se = _y_stats(sm, sq, r.width * r.height * r.spp)
    assert abs(y - want) < 4 * se + 1e-3 * want, (y, want, se)
    assert abs(bug - want) > 12 * se


SKY_PLANE_CASES = [
    dict(create_sun=False),
    dict(create_sun=True),
    dict(create_sun=True, lights=[pt_scene.DirectionalLight((1.0, 2.0, -0.5), (0.2, 0.9, 0.4), 1.5),
                                  pt_scene.DirectionalLight((0.3, -1.0, 0.2), (1.0, 1.0, 1.0), 9.0)]),
    dict(create_sun=False, elevation_deg=8.0, azimuth_deg=250.0, turbidity=6.0),
]


@pytest.mark.parametrize("kw", SKY_PLANE_CASES)
def test_sky_plane_closed_form(kw):
    """A Lambertian ground under the sky: albedo / pi times the ground's
    irradiance from the sky (hemisphere quadrature of the pinned sky model),
    the sun (create_sun) and directional lights (one below the horizon, which
    must add nothing).  Checks the environment light sample, its MIS pair with
    an escaping BSDF ray (dropping the escape weight would double the sky's
    share) and the directional lights' cosine and luminance."""
    s, r = pt_scene.sky_plane(**kw)
    want = pt_closed_forms.sky_plane_y(s, n_mu=96, n_phi=192)
    r.spp = 256
    sm, sq, _ = oracle.render(s, r)
    y, se = _y_stats(sm, sq, r.width * r.height * r.spp)
    assert abs(y - want) < 4 * se + 1e-3 * want, (y, want, se)


def test_ior_stack_is_unbounded():
    """The index-of-refraction stack follows the reference's unbounded
    IORStack (Utilities/IORStack.h:95): four nested lossless glass spheres in
    an emitting box still show radiance 1, and the refractions into the 4th
    nested sphere (which the HIP renderer's 3-object stack reports as
    overflows) are counted; with three nested spheres there are none."""
    s, r = pt_scene.nested_glass(3)
    r.spp = 8
    oracle.render(s, r)
    assert oracle.last_ior_deep_pushes() == 0
    s, r = pt_scene.nested_glass(4)
    r.spp = 16
    sm, sq, _ = oracle.render(s, r)
    assert oracle.last_ior_deep_pushes() > 1000
    y, se = _y_stats(sm, sq, r.width * r.height * r.spp)
    assert abs(y - 1.0) < 4 * se + 2e-3, (y, se)


def test_unshuffled_sobol_is_biased_on_the_spectral_furnace():
    """The reference's padded Sobol draws every dimension pair at the same
    index (SobolSequence.h:180-200), so dimensions 0, 2, 4, ... are hash
    scrambles of one van der Corput point.  Over eight renders the unshuffled
    sampler misses the spectral furnace's closed form by about -1% (many
    standard errors of the eight-render mean); the default sampler, which
    shuffles the index per dimension pair, does not."""
    rho, le = pt_closed_forms.spectral_furnace_spectra()
    want = pt_closed_forms.furnace_y(rho, le)
    out = {}
    for sampler in (pt_scene.SAMPLER_ZSOBOL_UNSHUFFLED, pt_scene.SAMPLER_ZSOBOL):
        ys = []
        for seed in range(1, 9):
            s, r = pt_scene.furnace_box(albedo=rho, radiance=le, max_bounce=96)
            r.spp, r.seed, r.sampler = 64, seed, sampler
            sm, sq, _ = oracle.render(s, r)
            ys.append(_y_stats(sm, sq, r.width * r.height * r.spp)[0])
        ys = np.array(ys)
        out[sampler] = (ys.mean() / want - 1.0, (ys.mean() - want) / (ys.std(ddof=1) / math.sqrt(len(ys))))
    bias, z = out[pt_scene.SAMPLER_ZSOBOL_UNSHUFFLED]
    assert bias < -0.005 and z < -5, out
    bias, z = out[pt_scene.SAMPLER_ZSOBOL]
    assert abs(bias) < 0.003 and abs(z) < 4, out


def test_sobol_sampler_uses_each_pixels_block():
    """Z-order sampler: pixel (x, y) takes indices Morton(x, y) * 2^l2 + [0, spp);
    with the shuffle the samples of a pixel are a permutation of that block, so
    over all spp = 2^l2 samples the pixel's Sobol points are exactly those of the
    block (its (0, 2)-net in dimensions 0 and 1)."""
    # Dimensions 0 and 1 of the block's points form a (0, m, 2)-net: one point per
    # elementary interval of area 1/2^m, here checked on the 2^m x 1 and 1 x 2^m grids.
    m = 6
    seed = oracle.hash_combine(37, 0)
    x = np.array([oracle.sobol_u32(i, 0, seed) for i in range(5 << m, 6 << m)], np.uint64)
    y = np.array([oracle.sobol_u32(i, 1, seed) for i in range(5 << m, 6 << m)], np.uint64)
    assert len(np.unique(x >> (32 - m))) == 1 << m
    assert len(np.unique(y >> (32 - m))) == 1 << m
    assert len(np.unique(((x >> (32 - m // 2)) << (m // 2)) | (y >> (32 - m // 2)))) == 1 << m


def test_emitter_seen_directly():
    s, r = pt_scene.cornell_box()
    # A black emitter (no reflected light on top of Le), seen head-on.
    s.spds.append(pt_scene.Spd.flat(0.0))
    s.materials[3].reflectance = len(s.spds) - 1
    s.camera = pt_scene.Camera(position=(0.0, 1.0, 0.0), look_at=(0.0, 2.0, 0.0), up=(0.0, 0.0, 1.0), fov_deg=10.0)
    r.width = r.height = 8
    r.spp = 256
    sm, sq, _ = oracle.render(s, r)
    y, se = _y_stats(sm, sq, r.width * r.height * r.spp)
    assert abs(y - 15.0) < 4 * se + 1e-6, (y, se)


def test_two_seeds_agree_statistically():
    s, r = pt_scene.cornell_box()
    r.width = r.height = 24
    r.spp = 64
    a, aq, _ = oracle.render(s, r)
    r.seed = 99
    b, bq, _ = oracle.render(s, r)
    n = r.spp
    ma, mb = a / n, b / n
    va = np.maximum(aq / n - ma * ma, 0) / n
    vb = np.maximum(bq / n - mb * mb, 0) / n
    z = (ma - mb) / np.sqrt(va + vb + 1e-30)
    frac = np.mean(np.abs(z[..., 1]) < 3.0)
    assert frac >= 0.97, frac
    assert abs(ma[..., 1].mean() / mb[..., 1].mean() - 1) < 0.02


def test_oracle_dielectric_furnace():
    # A lossless glass sphere inside a uniformly emitting box (radiance 1, albedo 0): every
    # path ends on the box with weight R + T = 1 through the single-wavelength branch.
    s, r = pt_scene.furnace_with(pt_scene.DIELECTRIC)
    sm, _, cnt = oracle.render(s, r)
    y = sm[..., 1] / r.spp
    assert abs(y.mean() - 1.0) < 0.005, y.mean()
    assert cnt[0] > cnt[2]  # refracted and reflected continuations were traced


def test_oracle_ggx_furnace():
    # A GGX sphere with F ~ 1 in the same box: radiance = the lobe's albedo, ~1 with the
    # multiscatter compensation.
    s, r = pt_scene.furnace_with(pt_scene.GGX, alpha=0.5)
    sm, _, _ = oracle.render(s, r)
    y = sm[..., 1] / r.spp
    assert abs(y.mean() - 1.0) < 0.03, y.mean()


def _sky_scene(elev=45.0, az=0.0, turb=3.0, scale=1.0):
    s, r = pt_scene.cornell_box()
    s.sky = pt_scene.Sky(elev, az, turb, scale)
    return s, r


def _sky(s, d, nm):
    import ctypes

    L = oracle.oracle_lib()
    L.rise_oracle_sky_radiance.restype = ctypes.c_double
    # build the ro_scene the same way render does, through a 1x1 render's parameters
    rs = oracle._RoScene()
    rs.sky_enabled = 1
    rs.sky_elevation_deg, rs.sky_azimuth_deg = s.sky.elevation_deg, s.sky.azimuth_deg
    rs.sky_turbidity, rs.sky_scale = s.sky.turbidity, s.sky.scale
    v = (ctypes.c_double * 3)(*d)
    L.rise_oracle_sky_radiance.argtypes = [ctypes.POINTER(oracle._RoScene), ctypes.c_void_p, ctypes.c_double]
    return L.rise_oracle_sky_radiance(ctypes.byref(rs), ctypes.cast(v, ctypes.c_void_p), nm)


def test_oracle_sky_invariants():
    # tests/HosekWilkieReferenceTest.cpp: zenith and horizon positive, ground zero,
    # a clear sky bluer than a hazy one; plus outside-band zero, symmetry about the
    # sun's vertical plane, linearity in the scale, brightest toward the sun.
    s, _ = _sky_scene(85.0, 0.0, 2.0)
    zen = _sky(s, (0, 1, 0), 550)
    hor = _sky(s, (1, 0.01, 0), 550)
    assert zen > 0 and hor > 0 and math.isfinite(zen) and math.isfinite(hor)
    assert _sky(s, (0, -1, 0), 550) == 0.0
    assert _sky(s, (0, 1, 0), 379.0) == 0.0 and _sky(s, (0, 1, 0), 781.0) == 0.0
    clear, _ = _sky_scene(45.0, 0.0, 2.0)
    hazy, _ = _sky_scene(45.0, 0.0, 8.0)
    br = lambda sk: _sky(sk, (0, 1, 0), 450) / _sky(sk, (0, 1, 0), 650)  # noqa: E731
    assert br(clear) > br(hazy)
    s, _ = _sky_scene(30.0, 40.0, 3.0)
    sun = (math.cos(math.radians(30)) * math.sin(math.radians(40)), math.sin(math.radians(30)),
           math.cos(math.radians(30)) * math.cos(math.radians(40)))
    # mirror a direction across the vertical plane containing the sun
    n = np.cross(np.array(sun), np.array([0.0, 1.0, 0.0]))
    n /= np.linalg.norm(n)
    d = np.array([0.3, 0.5, -0.2])
    d /= np.linalg.norm(d)
    m = d - 2 * np.dot(d, n) * n
    assert _sky(s, tuple(d), 520) == pytest.approx(_sky(s, tuple(m), 520), rel=1e-12)
    assert _sky(s, sun, 520) > _sky(s, (-sun[0], sun[1], -sun[2]), 520)
    s2, _ = _sky_scene(30.0, 40.0, 3.0, 2.5)
    assert _sky(s2, tuple(d), 520) == pytest.approx(2.5 * _sky(s, tuple(d), 520), rel=1e-12)


def test_oracle_sky_seen_by_camera():
    # A camera ray that leaves the scene sees the sky at each of its wavelengths; with
    # the background off it sees black.
    s, r = pt_scene.sky_spheres()
    s.camera = pt_scene.Camera((0, 1, 0), (0, 2, 0.01), (0, 0, 1), 5.0)   # straight up, nothing above
    r.width = r.height = 4
    r.spp = 64
    sm, _, _ = oracle.render(s, r)
    assert sm[..., 1].mean() > 0
    r.sky_background = False
    sm0, _, _ = oracle.render(s, r)
    assert np.all(sm0 == 0)
