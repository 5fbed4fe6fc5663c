"""Pins the path-tracing oracle's building blocks value for value against the
reference's own code, compiled from its sources into oracle/_ref
(oracle/ref/ref_pt_pins.cpp; skipped when that build is absent):

* the sky radiance map (HosekWilkieSkyModel::SampleRadiance, Preetham 1999),
  over sun elevations, azimuths, turbidities, directions and wavelengths;
* conductor and dielectric Fresnel, refraction and reflection (Optics);
* the anisotropic GGX D, G1, G2, VNDF sampler and VNDF pdf (MicrofacetUtils.h);
* the energy tables' lookups, the multiscatter lobe's normalization, sampler
  and pdf, the hemispherical Fresnel average and F_ms (MicrofacetEnergyLUT.h);
* Russian roulette and the power heuristic (PathTransportUtilities.h);
* the HWSS wavelength sampler (SampledWavelengths.h);
* OrthonormalBasis3D::CreateFromW;
* the Owen-scrambled Sobol sample, HashCombine and the Morton code of the
  sampler (SobolSequence.h, MortonCode.h), bit for bit.

Tolerance: 1e-12 relative (both sides run in IEEE double without
contraction); Russian roulette, the power heuristic and the wavelengths are
compared for equality.
"""
import ctypes
import math

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.skipif(not oracle.ref_available("ieee"), reason="reference build (oracle/_ref) absent")

_D = ctypes.c_double
_DP = ctypes.POINTER(ctypes.c_double)
_SIG = {
    "rise_ref_sky_radiance": ([_D, _D, _D, _D, _DP, _D], _D),
    "rise_ref_sun_direction": ([_D, _D, _DP], None),
    "rise_ref_fresnel_conductor": ([_DP, _DP, _D, _D, _D], _D),
    "rise_ref_dielectric_reflectance": ([_DP, _DP, _DP, _D, _D], _D),
    "rise_ref_refract": ([_DP, _D, _D, _DP], ctypes.c_int),
    "rise_ref_reflect": ([_DP, _DP, _DP], None),
    "rise_ref_ggx_d": ([_D, _D, _DP], _D),
    "rise_ref_ggx_g1": ([_D, _D, _DP], _D),
    "rise_ref_ggx_g2": ([_D, _D, _DP, _DP], _D),
    "rise_ref_vndf_sample": ([_DP, _DP, _D, _D, _D, _D, _DP], None),
    "rise_ref_vndf_pdf": ([_DP, _DP, _DP, _D, _D], _D),
    "rise_ref_ggx_lookup_ess": ([_D, _D], _D),
    "rise_ref_ggx_lookup_eavg": ([_D], _D),
    "rise_ref_ggx_ms_z": ([_D], _D),
    "rise_ref_ggx_sample_ms_cos": ([_D, _D], _D),
    "rise_ref_ggx_ms_pdf": ([_D, _D, _D], _D),
    "rise_ref_fresnel_avg": ([_DP, _D, _D, _D], _D),
    "rise_ref_ggx_fms": ([_D, _D], _D),
    "rise_ref_russian_roulette": ([ctypes.c_uint, ctypes.c_uint, _D, _D, _D, _D, _DP], None),
    "rise_ref_power_heuristic": ([_D, _D], _D),
    "rise_ref_sample_equidistant": ([_D, _D, _D, _DP], None),
    "rise_ref_onb_from_w": ([_DP, _DP], None),
    "rise_ref_sobol_sample": ([ctypes.c_uint32] * 3, _D),
    "rise_ref_hash_combine": ([ctypes.c_uint32] * 2, ctypes.c_uint32),
    "rise_ref_morton2d": ([ctypes.c_uint32] * 2, ctypes.c_uint32),
}
_ORACLE_SIG = {
    "rise_oracle_russian_roulette": ([ctypes.c_int, ctypes.c_int, _D, _D, _D, _D, _DP], None),
    "rise_oracle_power_heuristic": ([_D, _D], _D),
    "rise_oracle_sample_wavelengths": ([_D, _D, _D, _DP], None),
    "rise_oracle_onb_from_w": ([_DP, _DP], None),
    "rise_oracle_reflect": ([_DP, _DP, _DP], None),
}

V = oracle.vec


@pytest.fixture(scope="module")
def ref():
    L = ctypes.CDLL(str(oracle.REF_DIR / "libriseref_ieee.so"))
    for name, (a, r) in _SIG.items():
        fn = getattr(L, name)
        fn.argtypes = a
        fn.restype = r
    return L


@pytest.fixture(scope="module")
def orc():
    L = oracle.bsdf_lib()
    for name, (a, r) in _ORACLE_SIG.items():
        fn = getattr(L, name)
        fn.argtypes = a
        fn.restype = r
    return L


def _close(a, b, rel=1e-12, abs_=1e-300):
    return abs(a - b) <= rel * max(abs(a), abs(b)) + abs_


def _unit(g, n):
    v = g.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def test_sky_radiance_matches_reference(ref):
    from tests.test_oracle_pt import _sky, _sky_scene

    g = np.random.Generator(np.random.PCG64(11))
    n_checked = 0
    for elev, az, turb, scale in [(45.0, 0.0, 3.0, 1.0), (85.0, 10.0, 2.0, 0.7), (12.0, 200.0, 6.5, 3.0),
                                  (3.0, 90.0, 9.5, 1.0), (60.0, -35.0, 1.2, 2.0)]:
        s, _ = _sky_scene(elev, az, turb, scale)
        dirs = _unit(g, 60)
        dirs[:5, 1] = np.abs(dirs[:5, 1]) * 1e-4  # grazing, just above the horizon
        for d in dirs:
            for nm in (380.0, 412.5, 455.0, 500.1, 555.0, 601.7, 650.0, 719.9, 780.0, 781.0):
                want = ref.rise_ref_sky_radiance(elev, az, turb, scale, V(*d), nm)
                got = _sky(s, tuple(d), nm)
                assert _close(got, want), (elev, az, turb, d, nm, got, want)
                n_checked += want > 0
    assert n_checked > 1000
    out = V(0, 0, 0)
    ref.rise_ref_sun_direction(30.0, 40.0, out)
    el, az = math.radians(30.0), math.radians(40.0)
    assert np.allclose(out[:], [math.cos(el) * math.sin(az), math.sin(el), math.cos(el) * math.cos(az)], 0, 1e-15)


def test_fresnel_and_refraction_match_reference(ref, orc):
    g = np.random.Generator(np.random.PCG64(12))
    n = V(0, 0, 1)
    for _ in range(300):
        ct = g.uniform(0.0, 1.0)
        st = math.sqrt(1 - ct * ct)
        ni, nt, kt = g.uniform(1.0, 1.6), g.uniform(0.1, 4.0), g.uniform(0.0, 5.0)
        want = ref.rise_ref_fresnel_conductor(V(st, 0, -ct), n, ni, nt, kt)
        assert _close(orc.rise_oracle_fresnel_conductor(ct, ni, nt, kt), want)
    for _ in range(300):
        vin = _unit(g, 1)[0]
        nn = _unit(g, 1)[0]
        ni, nt = (1.0, 1.5) if g.random() < 0.5 else (1.5, 1.0)
        a, b = V(*vin), V(*vin)
        ok_r = ref.rise_ref_refract(V(*nn), ni, nt, a)
        ok_o = orc.rise_oracle_refract(V(*nn), ni, nt, b)
        assert ok_r == ok_o
        if ok_r:
            assert all(_close(x, y) for x, y in zip(a[:], b[:]))
            want = ref.rise_ref_dielectric_reflectance(V(*vin), a, V(*nn), ni, nt)
            assert _close(orc.rise_oracle_dielectric_reflectance(V(*vin), a, V(*nn), ni, nt), want)
        ra, rb = V(0, 0, 0), V(0, 0, 0)
        ref.rise_ref_reflect(V(*vin), V(*nn), ra)
        orc.rise_oracle_reflect(V(*vin), V(*nn), rb)
        assert all(_close(x, y, abs_=1e-16) for x, y in zip(ra[:], rb[:]))


def test_ggx_microfacet_functions_match_reference(ref, orc):
    g = np.random.Generator(np.random.PCG64(13))
    for _ in range(400):
        ax, ay = g.uniform(1e-5, 1.0, 2)
        if g.random() < 0.5:
            ay = ax
        h = _unit(g, 1)[0]
        h[2] = abs(h[2])
        wi, wo = _unit(g, 2)
        wi[2], wo[2] = abs(wi[2]), abs(wo[2])
        assert _close(orc.rise_oracle_ggx_d(ax, ay, V(*h)), ref.rise_ref_ggx_d(ax, ay, V(*h)))
        assert _close(orc.rise_oracle_ggx_g1(ax, ay, V(*wi)), ref.rise_ref_ggx_g1(ax, ay, V(*wi)))
        assert _close(orc.rise_oracle_ggx_g2(ax, ay, V(*wi), V(*wo)), ref.rise_ref_ggx_g2(ax, ay, V(*wi), V(*wo)))
        w = _unit(g, 1)[0]
        onb_r, onb_o = V(*[0] * 9), V(*[0] * 9)
        ref.rise_ref_onb_from_w(V(*w), onb_r)
        orc.rise_oracle_onb_from_w(V(*w), onb_o)
        assert all(_close(x, y, abs_=1e-16) for x, y in zip(onb_r[:], onb_o[:]))
        wiw = _unit(g, 1)[0]
        if np.dot(wiw, w) < 0:
            wiw = -wiw
        u1, u2 = g.random(2)
        mr, mo = V(0, 0, 0), V(0, 0, 0)
        ref.rise_ref_vndf_sample(V(*wiw), onb_r, ax, ay, u1, u2, mr)
        orc.rise_oracle_vndf_sample(V(*wiw), onb_r, ax, ay, u1, u2, mo)
        assert all(_close(x, y, abs_=1e-15) for x, y in zip(mr[:], mo[:]))
        wow = _unit(g, 1)[0]
        assert _close(orc.rise_oracle_vndf_pdf(V(*wiw), V(*wow), onb_r, ax, ay),
                      ref.rise_ref_vndf_pdf(V(*wiw), V(*wow), onb_r, ax, ay))


def test_energy_tables_and_multiscatter_lobe_match_reference(ref, orc):
    g = np.random.Generator(np.random.PCG64(14))
    for _ in range(300):
        c, a = g.uniform(-0.1, 1.1), g.uniform(-0.1, 1.2)
        assert _close(orc.rise_oracle_ggx_lookup_ess(c, a), ref.rise_ref_ggx_lookup_ess(c, a))
        assert _close(orc.rise_oracle_ggx_lookup_eavg(a), ref.rise_ref_ggx_lookup_eavg(a))
    for a in np.concatenate([np.linspace(0.0, 1.0, 33), g.uniform(0, 1, 40)]):
        z = ref.rise_ref_ggx_ms_z(a)
        assert _close(orc.rise_oracle_ggx_ms_z(a), z, 1e-12)
        for u in g.random(8):
            assert _close(orc.rise_oracle_ggx_sample_ms_cos(a, u), ref.rise_ref_ggx_sample_ms_cos(a, u), 1e-12)
        for c in g.random(8):
            assert _close(orc.rise_oracle_ggx_ms_pdf(c, a, z), ref.rise_ref_ggx_ms_pdf(c, a, z))
    for _ in range(100):
        ni, nt, kt = g.uniform(1.0, 1.6), g.uniform(0.1, 4.0), g.uniform(0.0, 5.0)
        fa = ref.rise_ref_fresnel_avg(V(0, 0, 1), ni, nt, kt)
        assert _close(orc.rise_oracle_fresnel_avg(ni, nt, kt), fa)
        ea = g.random()
        assert _close(orc.rise_oracle_ggx_fms(fa, ea), ref.rise_ref_ggx_fms(fa, ea))


def test_russian_roulette_power_heuristic_wavelengths_match_reference(ref, orc):
    g = np.random.Generator(np.random.PCG64(15))
    for _ in range(2000):
        depth, mind = int(g.integers(0, 8)), int(g.integers(0, 5))
        thr = float(g.choice([0.0, 0.05, 0.2]))
        cur, prev = g.uniform(0, 2, 2) * (g.random(2) < 0.9)
        u = g.random()
        a, b = V(0, 0), V(0, 0)
        ref.rise_ref_russian_roulette(depth, mind, thr, cur, prev, u, a)
        orc.rise_oracle_russian_roulette(depth, mind, thr, cur, prev, u, b)
        assert a[:] == b[:], (depth, mind, thr, cur, prev, u, a[:], b[:])
        pa, pb = g.uniform(0, 10, 2)
        assert ref.rise_ref_power_heuristic(pa, pb) == orc.rise_oracle_power_heuristic(pa, pb)
    for a0, b0 in [(380.0, 720.0), (380.0, 780.0), (400.0, 700.0)]:
        for u in np.concatenate([g.random(200), [0.0, 0.25, 0.5, 0.75, 1 - 2 ** -24]]):
            x, y = V(0, 0, 0, 0), V(0, 0, 0, 0)
            ref.rise_ref_sample_equidistant(u, a0, b0, x)
            orc.rise_oracle_sample_wavelengths(u, a0, b0, y)
            assert x[:] == y[:]


def test_sobol_sampler_matches_reference(ref):
    g = np.random.Generator(np.random.PCG64(16))
    idx = np.concatenate([np.arange(70), g.integers(0, 2 ** 32, 300)])
    for i in idx:
        for dim in (0, 1, 2, 3, 16, 17, 31, 32 * 16 + 5, 32 * 40 + 30):
            seed = int(g.integers(0, 2 ** 32))
            assert oracle.sobol_u32(int(i), dim, seed) / 2 ** 32 == ref.rise_ref_sobol_sample(int(i), dim, seed)
    for a, b in g.integers(0, 2 ** 32, (200, 2)):
        assert oracle.hash_combine(int(a), int(b)) == ref.rise_ref_hash_combine(int(a), int(b))
    for x, y in g.integers(0, 2 ** 16, (200, 2)):
        want = ref.rise_ref_morton2d(int(x), int(y))
        got = 0
        for k in range(16):
            got |= ((int(x) >> k) & 1) << (2 * k) | ((int(y) >> k) & 1) << (2 * k + 1)
        assert got == want
