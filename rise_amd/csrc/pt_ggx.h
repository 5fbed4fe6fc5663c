// pt_ggx.h -- GGX conductor and smooth dielectric scattering on the device
// (FP32), for the spectral path tracer.
//
// Each function cites the reference routine whose behaviour it follows
// (paths relative to the reference's src/Library/):
//   Utilities/MicrofacetUtils.h:181-371       Smith Lambda / G1 / G2, GGX D, VNDF sample and pdf
//   Utilities/MicrofacetEnergyLUT.h:40-426    E_ss / E_avg lookups, multiscatter lobe, F_avg, F_ms
//   Utilities/Optics.hpp:35-84, Optics.cpp:20-160   conductor and dielectric Fresnel, reflection, refraction
//   Materials/GGXSPF.cpp:424-790, GGXBRDF.cpp:339-498   GGX sampling, pdf and value (conductor mode)
//   Materials/DielectricSPF.cpp:120-287, 395-432        smooth dielectric scattering
#pragma once

#include "ggx_lut_data.h"
#include "pt_device.h"

#include <array>

namespace rise_hip {

constexpr int kLut = RISE_GGX_LUT_SIZE;

// The multiscatter lobe's normalization Z = 2 * integral_0^1 (1 - Ess(mu)) mu dmu
// per alpha row of the table, over the piecewise-linear Ess model (flat caps
// below the first and above the last cell centre, linear between centres):
// MicrofacetEnergyLUT::MSLobeZ, evaluated at compile time.
constexpr double kEssTable[kLut * kLut] = {RISE_GGX_ESS_VALUES};

constexpr double seg_total(double lo, double hi, double ess_lo, double slope)
{
    const double y = hi - lo;
    const double p = (1.0 - ess_lo) * lo, q = (1.0 - ess_lo) - slope * lo, r = -slope;
    return p * y + q * y * y * 0.5 + r * y * y * y * (1.0 / 3.0);
}

constexpr std::array<float, kLut> make_row_z()
{
    std::array<float, kLut> z{};
    for (int row = 0; row < kLut; ++row) {
        const double* e = kEssTable + row * kLut;
        double I = seg_total(0.0, 0.5 / kLut, e[0], 0.0);
        for (int k = 0; k < kLut - 1; ++k) {
            const double ck = (k + 0.5) / kLut, ck1 = (k + 1.5) / kLut;
            I += seg_total(ck, ck1, e[k], (e[k + 1] - e[k]) / (ck1 - ck));
        }
        I += seg_total((kLut - 0.5) / kLut, 1.0, e[kLut - 1], 0.0);
        z[row] = (float)(2.0 * I);
    }
    return z;
}

struct GgxTables {
    float ess[kLut * kLut];
    float eavg[kLut];
    float row_z[kLut];
    float gl_x[21], gl_w[21];
};

// ---- Smith / GGX (MicrofacetUtils.h:197-371) --------------------------------

RISE_HD inline float ggx_lambda(float a, f3 v)
{
    const float vz2 = v.z * v.z;
    if (vz2 < 1e-20f) return 1e10f;
    a = a < 1e-4f ? 1e-4f : a;
    const float a2 = a * a * v.x * v.x + a * a * v.y * v.y;
    return (-1.0f + sqrtf(1.0f + a2 / vz2)) * 0.5f;
}

RISE_HD inline float ggx_g1(float a, f3 v) { return 1.0f / (1.0f + ggx_lambda(a, v)); }
RISE_HD inline float ggx_g2(float a, f3 wi, f3 wo) { return 1.0f / (1.0f + ggx_lambda(a, wi) + ggx_lambda(a, wo)); }

RISE_HD inline float ggx_d(float a, f3 h)
{
    if (h.z <= 0.0f) return 0.0f;
    a = a < 1e-4f ? 1e-4f : a;
    const float hx = h.x / a, hy = h.y / a;
    const float den = hx * hx + hy * hy + h.z * h.z;
    return 1.0f / (3.14159265358979f * a * a * den * den);
}

RISE_HD inline f3 to_local(const Onb& b, f3 v) { return mk3(dot3(v, b.u), dot3(v, b.v), dot3(v, b.w)); }

// Spherical-cap VNDF sampling (Dupuy and Benyoub 2023), MicrofacetUtils.h:275-318.
RISE_HD inline f3 vndf_sample(f3 wi, const Onb& b, float a, float u1, float u2)
{
    if (a < 1e-6f) return b.w;
    const f3 wl = to_local(b, wi);
    const f3 wh = norm3(mk3(a * wl.x, a * wl.y, wl.z));
    const float phi = kTwoPi * u1;
    const float z = (1.0f - u2) * (1.0f + wh.z) - wh.z;
    const float st = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    const f3 c = mk3(st * cosf(phi) + wh.x, st * sinf(phi) + wh.y, z + wh.z);
    const f3 ml = norm3(mk3(a * c.x, a * c.y, c.z));
    return norm3(b.u * ml.x + b.v * ml.y + b.w * ml.z);
}

RISE_HD inline float vndf_pdf(f3 wi, f3 wo, const Onb& b, float a)
{
    if (a < 1e-6f) return 0.0f;
    const float cos_wi = dot3(wi, b.w);
    if (cos_wi < 1e-10f) return 0.0f;
    const f3 h = norm3(wi + wo);
    const float wih = dot3(wi, h), woh = dot3(wo, h);
    if (wih <= 0.0f || woh < 1e-10f) return 0.0f;
    const f3 hl = to_local(b, h);
    if (hl.z <= 0.0f) return 0.0f;
    return ggx_g1(a, to_local(b, wi)) * wih * ggx_d(a, hl) / cos_wi / (4.0f * woh);
}

// ---- energy tables (MicrofacetEnergyLUT.h) ----------------------------------

RISE_HD inline void alpha_index(float alpha, int& a0, int& a1, float& f)
{
    const float a = fmaxf(0.0f, fminf(1.0f, (alpha - 0.01f) / (1.0f - 0.01f))) * (float)(kLut - 1);
    a0 = (int)a;
    a1 = a0 + 1 < kLut - 1 ? a0 + 1 : kLut - 1;
    f = a - (float)a0;
}

RISE_HD inline float lookup_ess(const GgxTables& T, float c, float alpha)
{
    int a0, a1;
    float af;
    alpha_index(alpha, a0, a1, af);
    float x = fmaxf(0.0f, fminf(1.0f, c)) * (float)kLut - 0.5f;
    if (x < 0.0f) x = 0.0f;
    const int c0 = (int)x;
    const int c1 = c0 + 1 < kLut - 1 ? c0 + 1 : kLut - 1;
    const float cf = x - (float)c0;
    const float* r0 = T.ess + a0 * kLut;
    const float* r1 = T.ess + a1 * kLut;
    return (1 - af) * ((1 - cf) * r0[c0] + cf * r0[c1]) + af * ((1 - cf) * r1[c0] + cf * r1[c1]);
}

RISE_HD inline float lookup_eavg(const GgxTables& T, float alpha)
{
    int a0, a1;
    float af;
    alpha_index(alpha, a0, a1, af);
    return (1 - af) * T.eavg[a0] + af * T.eavg[a1];
}

RISE_HD inline float ms_z(const GgxTables& T, float alpha)
{
    int a0, a1;
    float af;
    alpha_index(alpha, a0, a1, af);
    return (1.0f - af) * T.row_z[a0] + af * T.row_z[a1];
}

RISE_HD inline float ms_pdf(const GgxTables& T, float c, float alpha, float z)
{
    if (z <= 1e-15f) return 0.0f;
    c = fmaxf(0.0f, fminf(1.0f, c));
    return fmaxf(0.0f, (1.0f - lookup_ess(T, c, alpha)) * c / (3.14159265358979f * z));
}

// Segment k of the piecewise Ess model of the alpha-blended row.
RISE_HD inline void ms_segment(const GgxTables& T, int a0, int a1, double af, int k, double& lo, double& hi,
                               double& ess_lo, double& slope)
{
    auto row = [&](int c) { return (1.0 - af) * (double)T.ess[a0 * kLut + c] + af * (double)T.ess[a1 * kLut + c]; };
    if (k == 0) {
        lo = 0.0; hi = 0.5 / kLut; ess_lo = row(0); slope = 0.0;
    } else if (k == kLut) {
        lo = (kLut - 0.5) / kLut; hi = 1.0; ess_lo = row(kLut - 1); slope = 0.0;
    } else {
        lo = (k - 0.5) / kLut; hi = (k + 0.5) / kLut; ess_lo = row(k - 1); slope = (row(k) - row(k - 1)) / (hi - lo);
    }
}

// SampleMSCosTheta: exact per-segment CDF inversion (FP64: the lobe is rare
// and the Newton-bisection tolerance is below FP32 resolution).
RISE_HD inline float sample_ms_cos(const GgxTables& T, float alpha, float u)
{
    int a0, a1;
    float aff;
    alpha_index(alpha, a0, a1, aff);
    const double af = aff;
    double I = 0.0;
    for (int k = 0; k <= kLut; ++k) {
        double lo, hi, e, s;
        ms_segment(T, a0, a1, af, k, lo, hi, e, s);
        I += seg_total(lo, hi, e, s);
    }
    if (I <= 1e-15) return fmaxf(0.0f, fminf(1.0f, u));
    const double target = fmax(0.0, fmin(I, (double)u * I));
    double run = 0.0, lo = 0, hi = 0, e = 0, s = 0, tot = 0;
    int k = 0;
    for (;; ++k) {
        ms_segment(T, a0, a1, af, k, lo, hi, e, s);
        tot = seg_total(lo, hi, e, s);
        if (k == kLut || !(run + tot < target)) break;
        run += tot;
    }
    const double lt = fmax(0.0, fmin(tot, target - run));
    const double w = hi - lo;
    double ylo = 0.0, yhi = w, y = w * 0.5;
    const double P = (1.0 - e) * lo, Q = (1.0 - e) - s * lo, R = -s;
    for (int it = 0; it < 30; ++it) {
        const double fy = P * y + Q * y * y * 0.5 + R * y * y * y * (1.0 / 3.0);
        const double diff = fy - lt;
        if (fabs(diff) < 1e-13) break;
        if (diff > 0) yhi = y;
        else ylo = y;
        const double d = fmax(0.0, (1.0 - (e + s * y)) * (lo + y));
        double yn = (d > 1e-12) ? (y - diff / d) : 0.5 * (ylo + yhi);
        if (!(yn > ylo && yn < yhi)) yn = 0.5 * (ylo + yhi);
        y = yn;
    }
    return (float)fmax(0.0, fmin(1.0, lo + y));
}

// ---- Fresnel (Optics.hpp:35-84, MicrofacetEnergyLUT.h:384-426) -------------

RISE_HD inline float fresnel_conductor(float cos_i, float ni, float nt, float kt)
{
    float c = fabsf(cos_i);
    if (c < kNearZeroF) c = kNearZeroF;
    const float c2 = c * c;
    const float s2 = fmaxf(0.0f, 1.0f - c2);
    const float s = sqrtf(s2), t = s / c, t2 = t * t;
    const float ni2 = ni * ni, nt2 = nt * nt, kt2 = kt * kt;
    const float ns = ni2 * s2, nk = 4.0f * nt2 * kt2;
    float a2 = nt2 - kt2 - ns;
    const float sq = sqrtf(a2 * a2 + nk);
    float b2 = sq - nt2 + kt2 + ns;
    a2 = sq + nt2 - kt2 - ns;
    const float dni = 0.5f / ni2;
    a2 = fmaxf(0.0f, a2 * dni);
    b2 = fmaxf(0.0f, b2 * dni);
    const float a = sqrtf(a2), ab2 = a2 + b2, dcos = 2.0f * a * c, astan = 2.0f * a * s * t, stan = s2 * t2;
    const float rs = (ab2 - dcos + c2) / (ab2 + dcos + c2);
    const float rp = rs * (ab2 - astan + stan) / (ab2 + astan + stan);
    return 0.5f * (rp + rs);
}

RISE_HD inline float fresnel_avg(const GgxTables& T, float ni, float nt, float kt)
{
    float sum = 0.0f;
    for (int i = 0; i < 21; ++i) sum += fresnel_conductor(T.gl_x[i], ni, nt, kt) * (2.0f * T.gl_x[i] * T.gl_w[i]);
    return sum;
}

RISE_HD inline float ggx_fms(float f_avg, float e_avg)
{
    const float den = 1.0f - f_avg * (1.0f - e_avg);
    return den < 1e-10f ? 0.0f : f_avg * f_avg * e_avg / den;
}

// ---- GGX material at one wavelength -------------------------------------------

struct GgxParams {
    float rd, rs, alpha, ior, ext, amb;
};

// Ray-facing ONB (FlipW) and geometric normal oriented against the ray.
RISE_HD inline void ggx_facing(const Surf& h, Onb& b, f3& gn)
{
    b = h.frame;
    if (dot3(h.dir, h.frame.w) > kNearZeroF) {
        b.w = -b.w;
        b.u = -b.u;
    }
    const f3 raw = dot3(h.ng, h.ng) > 1e-12f ? h.ng : b.w;
    gn = dot3(raw, h.dir) < 0.0f ? raw : -raw;
}

// GGXBRDF::valueNM (GGXBRDF.cpp:339-498), conductor Fresnel.
RISE_HD inline float ggx_value(const GgxTables& T, const Surf& h, const GgxParams& g, f3 vin)
{
    Onb b;
    f3 gn;
    ggx_facing(h, b, gn);
    const f3 n = b.w, v = norm3(vin), r = norm3(-h.dir);
    const float nr = dot3(n, r), nv = dot3(n, v);
    if (nr < kNearZeroF || nv < kNearZeroF) return 0.0f;
    if (dot3(v, gn) <= 0.0f || dot3(r, gn) <= 0.0f) return 0.0f;
    const float a = fmaxf(g.alpha, 1e-4f);
    const f3 hv = norm3(v + r);
    const float sf = ggx_d(a, to_local(b, hv)) * ggx_g2(a, to_local(b, v), to_local(b, r)) / (4.0f * nv * nr);
    float spec = 0.0f;
    if (sf > 0.0f) {
        const float f = fresnel_conductor(dot3(h.dir, hv), g.amb, g.ior, g.ext);
        if (f > 0.0f) spec = g.rs * f * sf;
    }
    const float eavg = lookup_eavg(T, a);
    if (1.0f - eavg > 1e-10f) {
        const float shape = (1.0f - lookup_ess(T, nr, a)) * (1.0f - lookup_ess(T, nv, a)) /
                            (3.14159265358979f * (1.0f - eavg));
        spec += ggx_fms(g.rs * fresnel_avg(T, g.amb, g.ior, g.ext), eavg) * shape;
    }
    return g.rd * kInvPi + spec;
}

// GGXSPF::PdfNM (GGXSPF.cpp:740-790).
RISE_HD inline float ggx_pdf(const GgxTables& T, const Surf& h, const GgxParams& g, f3 wo)
{
    Onb b;
    f3 gn;
    ggx_facing(h, b, gn);
    const f3 n = b.w, won = norm3(wo);
    const float ct = dot3(won, n);
    if (ct <= 0.0f || dot3(won, gn) <= 0.0f) return 0.0f;
    const f3 wi = norm3(-h.dir);
    const float a = fmaxf(g.alpha, 1e-4f);
    const float wms = g.rs * (1.0f - lookup_ess(T, dot3(wi, n), a));
    const float total = g.rd + g.rs + wms;
    if (total < 1e-10f) return ct * kInvPi;
    const float spec_pdf = a >= 1e-6f ? vndf_pdf(wi, won, b, a) : 0.0f;
    return (g.rd * ct * kInvPi + wms * ms_pdf(T, ct, a, ms_z(T, a)) + g.rs * spec_pdf) / total;
}

struct ScatterOut {
    bool ok;
    f3 dir;
    float kray;
    float pdf;
    int type;  // 0 diffuse, 1 reflection, 2 refraction
};

// GGXSPF::ScatterNM (GGXSPF.cpp:424-687), conductor Fresnel: lobe by u_lobe,
// direction by (u1, u2).
RISE_HD inline ScatterOut ggx_scatter(const GgxTables& T, const Surf& h, const GgxParams& g, float u_lobe, float u1,
                                      float u2)
{
    ScatterOut out{false, mk3(0, 0, 0), 0.0f, 0.0f, 0};
    Onb b;
    f3 gn;
    ggx_facing(h, b, gn);
    const f3 n = b.w, wi = norm3(-h.dir);
    const float a = fmaxf(g.alpha, 1e-4f);
    const float eavg = lookup_eavg(T, a);
    const float ess_i = lookup_ess(T, dot3(wi, n), a);
    const float wd = g.rd, ws = g.rs, wms = ws * (1.0f - ess_i);
    const float total = wd + ws + wms;
    const float z = ms_z(T, a);
    const float p_diff = total > 1e-10f ? wd / total : 1.0f;
    const float p_spec = total > 1e-10f ? ws / total : 0.0f;
    auto mix = [&](float ct, float spec_pdf, float fallback) {
        return total > 1e-10f ? (wd * ct * kInvPi + wms * ms_pdf(T, ct, a, z) + ws * spec_pdf) / total : fallback;
    };
    if (u_lobe < p_diff) {
        const float phi = kTwoPi * u1, st = sqrtf(u2), ctt = sqrtf(1.0f - u2);
        const f3 wo = onb_xform(b, mk3(cosf(phi) * st, sinf(phi) * st, ctt));
        const float ct = dot3(wo, n);
        if (ct > 0.0f && dot3(wo, gn) > 0.0f) {
            const float sp = a >= 1e-6f ? vndf_pdf(wi, wo, b, a) : 0.0f;
            out = ScatterOut{true, wo, wd / p_diff, mix(ct, sp, ct * kInvPi), 0};
        }
    } else if (u_lobe < p_diff + p_spec) {
        if (a >= 1e-6f) {
            const f3 m = vndf_sample(wi, b, a, u1, u2);
            const float wim = dot3(wi, m);
            if (wim > 0.0f) {
                const f3 wo = norm3(m * (2.0f * wim) - wi);
                const float ct = dot3(wo, n);
                if (ct > 0.0f && dot3(wo, gn) > 0.0f) {
                    const float vp = vndf_pdf(wi, wo, b, a);
                    if (vp > 1e-10f) {
                        const float F = g.rs * fresnel_conductor(dot3(h.dir, m), g.amb, g.ior, g.ext);
                        const float G1 = ggx_g1(a, to_local(b, wi));
                        const float kray = G1 > 1e-10f ? F * ggx_g2(a, to_local(b, wi), to_local(b, wo)) / (G1 * p_spec)
                                                       : 0.0f;
                        if (kray > 0.0f) out = ScatterOut{true, wo, kray, mix(ct, vp, vp), 1};
                    }
                }
            }
        }
    } else {
        const float p_ms = 1.0f - p_diff - p_spec;
        if (p_ms > 1e-10f && 1.0f - eavg > 1e-10f) {
            const float ct = sample_ms_cos(T, a, u1);
            const float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
            const float phi = kTwoPi * u2;
            const f3 wo = onb_xform(b, mk3(st * cosf(phi), st * sinf(phi), ct));
            if (ct > 0.0f && dot3(wo, gn) > 0.0f) {
                const float sp = a >= 1e-6f ? vndf_pdf(wi, wo, b, a) : 0.0f;
                const float mp = ms_pdf(T, ct, a, z);
                const float fms = ggx_fms(g.rs * fresnel_avg(T, g.amb, g.ior, g.ext), eavg);
                const float f_ms = fms * (1.0f - lookup_ess(T, ct, a)) * (1.0f - ess_i) * kInvPi / (1.0f - eavg);
                const float kray = mp > 1e-14f ? f_ms * ct / (mp * p_ms) : 0.0f;
                if (kray > 0.0f) out = ScatterOut{true, wo, kray, mix(ct, sp, ct * kInvPi), 0};
            }
        }
    }
    return out;
}

// ---- dielectric (Optics.cpp:20-160, DielectricSPF.cpp:120-287) -----------------

RISE_HD inline f3 reflect3(f3 vin, f3 n)
{
    const float d = 2.0f * dot3(n, vin);
    return vin - n * d;
}

// CalculateRefractedRay: false on total internal reflection.
RISE_HD inline bool refract3(f3 n, float ni, float nt, f3& vin)
{
    if (ni <= kNearZeroF || nt <= kNearZeroF) return false;
    if (dot3(n, vin) > 0.0f) n = -n;
    const float k0 = dot3(n, vin);
    const f3 s = (vin - n * k0) * (ni / nt);
    const float k = 1.0f - dot3(s, s);
    if (k < kNearZeroF) return false;
    vin = s - n * sqrtf(k);
    return true;
}

RISE_HD inline float dielectric_reflectance(f3 v, f3 tv, f3 n, float ni, float nt)
{
    if (ni <= kNearZeroF || nt <= kNearZeroF) return 1.0f;
    const float ci = fabsf(dot3(v, n)), ct = fabsf(dot3(tv, n));
    const float nn = ni * nt, cc = ct * ci, ni2 = ni * ni, nt2 = nt * nt, ci2 = ci * ci, ct2 = ct * ct;
    const float nit = ni2 - nt2, cit = ci2 - ct2;
    const float num = nit * nit * cc * cc + cit * cit * nn * nn;
    float den = cc * (ni2 + nt2) + nn * (ci2 + ct2);
    den = den * den;
    const float r = (den < 0.000001f && num < 0.000001f) ? 1.0f : num / den;
    return r < 1.0f ? r : 1.0f;
}

}  // namespace rise_hip
