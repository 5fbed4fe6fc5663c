"""Pins the oracle's GGX conductor and dielectric math (oracle/pt_bsdf.c).

Checks come from the reference's own tests and from the mathematics:
* Fresnel: normal-incidence closed forms, Brewster's angle, total internal
  reflection, grazing behaviour (tests/OpticsTest.cpp cases).
* Snell's law and TIR for the refraction routine.
* GGX: NDF normalization, the VNDF pdf identity f cos / pdf = G2 / G1(wi),
  mixture pdf returned by the sampler equal to the pdf function at the
  sampled direction, G2 >= G1 G1 (tests/GGXWhiteFurnaceTest.cpp cases).
* The energy tables (data): E_ss against a Monte Carlo estimate of the
  directional albedo made with the oracle's VNDF sampler, E_avg against the
  cosine-weighted integral of E_ss, the multiscatter lobe's normalization,
  sampler and pdf against each other, and the white furnace: the GGX lobe
  plus its multiscatter term reflects ~all energy when F = 1.
* The 21-point Gauss-Legendre rule integrates polynomials of degree 41
  exactly and matches the reference's tabulated nodes.
"""
import math

import numpy as np
import pytest

import oracle

L = oracle.bsdf_lib()
V = oracle.vec


def _onb(n):
    n = np.asarray(n, np.float64)
    n = n / np.linalg.norm(n)
    a = np.array([1.0, 0, 0]) if abs(n[0]) < 0.9 else np.array([0, 1.0, 0])
    u = np.cross(n, a)
    u /= np.linalg.norm(u)
    v = np.cross(n, u)
    return u, v, n


def _ctx(d, n, rd=0.0, rs=1.0, alpha=0.3, ior=0.2, ext=3.0, amb=1.0, ng=None):
    u, v, w = _onb(n)
    ng = w if ng is None else ng
    return V(*d, *u, *v, *w, *ng, rd, rs, alpha, ior, ext, amb)


# ---- Fresnel, refraction ----------------------------------------------------

def test_conductor_normal_incidence_closed_form():
    n1, n2, k = 1.0, 0.2, 3.0
    r = L.rise_oracle_fresnel_conductor(1.0, n1, n2, k)
    assert r == pytest.approx(((n2 - n1) ** 2 + k * k) / ((n2 + n1) ** 2 + k * k), abs=1e-12)
    g = L.rise_oracle_fresnel_conductor(math.cos(math.radians(89.9)), n1, n2, k)
    assert 0.9 < g <= 1.0


def test_dielectric_reflectance_cases():
    # Normal incidence: ((n1 - n2) / (n1 + n2))^2.
    r = L.rise_oracle_dielectric_reflectance(V(0, -1, 0), V(0, -1, 0), V(0, 1, 0), 1.0, 1.5)
    assert r == pytest.approx(0.04, abs=1e-12)
    # Brewster's angle: Rp = 0, so R = Rs / 2.
    tb = math.atan(1.5)
    tt = math.asin(math.sin(tb) / 1.5)
    r = L.rise_oracle_dielectric_reflectance(V(math.sin(tb), -math.cos(tb), 0), V(math.sin(tt), -math.cos(tt), 0),
                                             V(0, 1, 0), 1.0, 1.5)
    rs = ((math.cos(tb) - 1.5 * math.cos(tt)) / (math.cos(tb) + 1.5 * math.cos(tt))) ** 2
    assert r == pytest.approx(0.5 * rs, rel=1e-10)
    # Equal indices: no reflection at any angle.
    for th in (0.0, 0.5, 1.2):
        d = V(math.sin(th), -math.cos(th), 0)
        assert L.rise_oracle_dielectric_reflectance(d, d, V(0, 1, 0), 1.3, 1.3) == pytest.approx(0.0, abs=1e-12)


def test_refraction_snell_and_tir():
    th = 0.6
    d = V(math.sin(th), -math.cos(th), 0)
    assert L.rise_oracle_refract(V(0, 1, 0), 1.0, 1.5, d) == 1
    t = np.array(d[:])
    assert np.linalg.norm(t) == pytest.approx(1.0, abs=1e-12)
    assert math.sin(th) == pytest.approx(1.5 * abs(t[0]), rel=1e-12)
    assert t[1] < 0
    # Past the critical angle going out of glass: total internal reflection.
    th = math.asin(1 / 1.5) + 0.05
    d = V(math.sin(th), math.cos(th), 0)
    assert L.rise_oracle_refract(V(0, 1, 0), 1.5, 1.0, d) == 0
    # Normal on the ray's side is flipped internally: same result either way.
    a, b = V(0.3, -0.95, 0.1), V(0.3, -0.95, 0.1)
    L.rise_oracle_refract(V(0, 1, 0), 1.0, 1.33, a)
    L.rise_oracle_refract(V(0, -1, 0), 1.0, 1.33, b)
    assert np.allclose(a[:], b[:], atol=1e-14)


# ---- GGX ----------------------------------------------------------------------

@pytest.mark.parametrize("alpha", [0.05, 0.2, 0.5, 0.9])
def test_ndf_normalization(alpha):
    # integral over the hemisphere of D(h) cos(theta_h) = 1 (u = cos^2 substitution)
    nt, npf = 4000, 8
    total = 0.0
    for i in range(nt):
        c = (i + 0.5) / nt
        h = V(math.sqrt(1 - c * c), 0, c)
        total += L.rise_oracle_ggx_d(alpha, alpha, h) * c * (1.0 / nt) * 2 * math.pi
    assert total == pytest.approx(1.0, rel=0.02)


def test_vndf_pdf_identity_and_g2_bound():
    rng = np.random.default_rng(5)
    onb = V(1, 0, 0, 0, 1, 0, 0, 0, 1)
    for _ in range(200):
        a = float(rng.uniform(0.05, 0.9))
        wi = rng.normal(size=3)
        wi[2] = abs(wi[2]) + 0.05
        wi /= np.linalg.norm(wi)
        m = V(0, 0, 0)
        L.rise_oracle_vndf_sample(V(*wi), onb, a, a, float(rng.random()), float(rng.random()), m)
        m = np.array(m[:])
        wo = 2 * np.dot(wi, m) * m - wi
        wo /= np.linalg.norm(wo)
        if wo[2] <= 1e-3 or np.dot(wi, m) <= 0:
            continue
        pdf = L.rise_oracle_vndf_pdf(V(*wi), V(*wo), onb, a, a)
        h = (wi + wo) / np.linalg.norm(wi + wo)
        D = L.rise_oracle_ggx_d(a, a, V(*h))
        G2 = L.rise_oracle_ggx_g2(a, a, V(*wi), V(*wo))
        G1 = L.rise_oracle_ggx_g1(a, a, V(*wi))
        f = D * G2 / (4 * wi[2] * wo[2])
        assert f * wo[2] / pdf == pytest.approx(G2 / G1, rel=1e-9)
        assert G2 >= G1 * L.rise_oracle_ggx_g1(a, a, V(*wo)) - 1e-12


def test_scatter_pdf_equals_pdf_function():
    rng = np.random.default_rng(9)
    n = np.array([0.2, 0.9, -0.3])
    n /= np.linalg.norm(n)
    checked = 0
    for _ in range(400):
        d = rng.normal(size=3)
        d -= (np.dot(d, n) + 0.2) * n * 2  # mostly against the normal
        d /= np.linalg.norm(d)
        ctx = _ctx(d, n, rd=float(rng.uniform(0, 0.6)), rs=float(rng.uniform(0.2, 1)),
                   alpha=float(rng.uniform(0.05, 0.8)), ior=float(rng.uniform(0.1, 2.5)),
                   ext=float(rng.uniform(0.5, 4)))
        out = V(*([0.0] * 7))
        L.rise_oracle_ggx_scatter(ctx, float(rng.random()), float(rng.random()), float(rng.random()), out)
        if out[0] == 0:
            continue
        wo = V(out[1], out[2], out[3])
        assert out[5] == pytest.approx(L.rise_oracle_ggx_pdf(ctx, wo), rel=1e-9)
        assert out[4] > 0
        checked += 1
    assert checked > 200


def test_ess_table_matches_vndf_estimate():
    # The reference fills E_ss with the mean of G1(wo) over VNDF-sampled reflections
    # (tools/GenerateMicrofacetEnergyLUT.cpp:100-130): the directional albedo of
    # GGX with the separable masking-shadowing G1(wi) G1(wo), F = 1.  (Its BRDF
    # uses the height-correlated G2; the two agree near normal incidence and
    # differ at grazing angles on rough surfaces, by 0.33 at alpha = 1.)
    rng = np.random.default_rng(3)
    onb = V(1, 0, 0, 0, 1, 0, 0, 0, 1)
    for ai, ci in [(3, 5), (10, 16), (20, 28), (31, 2), (15, 10), (31, 0)]:
        alpha = 0.01 + ai * (0.99 / 31)
        mu = (ci + 0.5) / 32
        wi = np.array([math.sqrt(1 - mu * mu), 0.0, mu])
        n = 20000
        sep = corr = 0.0
        G1 = L.rise_oracle_ggx_g1(alpha, alpha, V(*wi))
        for _ in range(n):
            m = V(0, 0, 0)
            L.rise_oracle_vndf_sample(V(*wi), onb, alpha, alpha, float(rng.random()), float(rng.random()), m)
            m = np.array(m[:])
            wo = 2 * np.dot(wi, m) * m - wi
            if wo[2] > 0:
                sep += L.rise_oracle_ggx_g1(alpha, alpha, V(*wo))
                corr += L.rise_oracle_ggx_g2(alpha, alpha, V(*wi), V(*wo)) / G1
        table = L.rise_oracle_ggx_lookup_ess(mu, alpha)
        assert table == pytest.approx(sep / n, abs=0.012), (ai, ci, table, sep / n)
        assert corr / n >= sep / n - 0.01  # the height-correlated albedo is never smaller


def test_eavg_is_cosine_weighted_ess():
    for ai in (0, 7, 15, 31):
        alpha = 0.01 + ai * (0.99 / 31)
        mus = (np.arange(2000) + 0.5) / 2000
        ess = np.array([L.rise_oracle_ggx_lookup_ess(m, alpha) for m in mus])
        integral = 2 * np.mean(ess * mus)
        assert L.rise_oracle_ggx_lookup_eavg(alpha) == pytest.approx(integral, abs=0.006)


def test_ms_lobe_normalization_sampler_and_pdf():
    for alpha in (0.08, 0.35, 0.8):
        z = L.rise_oracle_ggx_ms_z(alpha)
        assert z == pytest.approx(1 - L.rise_oracle_ggx_lookup_eavg(alpha), abs=0.006)
        mus = (np.arange(4000) + 0.5) / 4000
        pdf = np.array([L.rise_oracle_ggx_ms_pdf(m, alpha, z) for m in mus])
        assert 2 * math.pi * np.mean(pdf) == pytest.approx(1.0, abs=1e-3)  # solid-angle density integrates to 1
        us = (np.arange(20000) + 0.5) / 20000
        cs = np.array([L.rise_oracle_ggx_sample_ms_cos(alpha, u) for u in us])
        hist, edges = np.histogram(cs, bins=20, range=(0, 1))
        expect = []
        for a, b in zip(edges[:-1], edges[1:]):
            m = np.linspace(a, b, 51)
            expect.append(2 * math.pi * np.trapezoid([L.rise_oracle_ggx_ms_pdf(x, alpha, z) for x in m], m))
        assert np.allclose(hist / len(cs), expect, atol=2e-3)


def _header_table(name: str) -> np.ndarray:
    import re
    from pathlib import Path

    text = (Path(oracle.HERE).parent / "rise_amd" / "csrc" / "ggx_lut_data.h").read_text()
    i = text.index(f"#define {name}")
    body = text[i + len(name) + 8:].split("\n#define")[0]
    return np.array([float(v) for v in re.findall(r"[-+]?\d+\.\d+(?:[eE][-+]?\d+)?", body)])


def test_gauss_legendre_rule():
    """The hemispherical Fresnel average uses the reference's tabulated
    21-point rule (MicrofacetEnergyLUT.h:358-378), kept as data: its nodes are
    the exact rule's to 1e-15, its weights within 3.8e-8 (two of them are off
    by that much), summing to 1 + 7.4e-8."""
    x, w = np.polynomial.legendre.leggauss(21)
    nodes, weights = 0.5 * (x + 1), 0.5 * w
    assert np.sum(weights * nodes ** 41) == pytest.approx(1 / 42, rel=1e-13)
    tn, tw = _header_table("RISE_GL21_NODES"), _header_table("RISE_GL21_WEIGHTS")
    assert tn.size == tw.size == 21
    assert np.abs(tn - nodes).max() < 1e-15
    assert np.abs(tw - weights).max() < 3.8e-8
    assert abs(tw.sum() - 1.0 - 7.35e-8) < 1e-10
    # F_avg of a perfect mirror (F = 1 everywhere) is 2 sum(w mu) = 1 + 7.4e-8 with the table.
    assert L.rise_oracle_fresnel_avg(1.0, 1e-3, 1e6) == pytest.approx(float(2 * np.sum(tw * tn)), abs=1e-9)
    assert L.rise_oracle_fresnel_avg(1.0, 1e-3, 1e6) == pytest.approx(1.0, abs=1e-7)


@pytest.mark.parametrize("alpha", [0.1, 0.4, 0.8])
def test_white_furnace_with_multiscatter(alpha):
    # rd = 0, rs = 1, F ~ 1: the mean of kray (f cos / pdf / select) is the albedo, ~1 with the
    # multiscatter compensation (Kulla-Conty), below 1 for the single-scatter lobe alone.
    rng = np.random.default_rng(11)
    n = np.array([0.0, 0.0, 1.0])
    for th in (0.2, 0.9):
        d = np.array([math.sin(th), 0.0, -math.cos(th)])
        ctx = _ctx(d, n, rd=0.0, rs=1.0, alpha=alpha, ior=1e-3, ext=1e6)
        N = 40000
        tot = 0.0
        out = V(*([0.0] * 7))
        for _ in range(N):
            L.rise_oracle_ggx_scatter(ctx, float(rng.random()), float(rng.random()), float(rng.random()), out)
            if out[0]:
                tot += out[4]
        assert tot / N == pytest.approx(1.0, abs=0.03), (alpha, th, tot / N)
