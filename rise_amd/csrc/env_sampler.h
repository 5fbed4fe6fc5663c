// env_sampler.h -- importance sampling of the sky radiance map for the
// environment light sample of next-event estimation: the reference's
// EnvironmentSampler (Rendering/EnvironmentSampler.cpp) as RayCaster builds it
// for a radiance map (RayCaster.cpp:330-360: 64 x 64 cells) over the
// hosek_wilkie_skylight map's painter (HosekWilkieSpectralRadianceMap.cpp:
// HWAdapterPainter), and used by LightSampler::EvaluateDirectLightingNM
// (Lights/LightSampler.cpp:2278-2358) and, for the weight of a BSDF-sampled
// ray that escapes, PathTracingIntegrator.cpp:4259-4282.
//
// The table is the reference's, quirks included: its cells are laid out in
// the light-probe (angular) mapping s = 0.5 + x r, t = 0.5 - y r with
// r = acos(-z) / (2 pi sqrt(x^2 + y^2)), which sampling and the pdf invert,
// but each cell's weight is the painter's colour at the cell centre read as
// equirectangular (u -> azimuth, v -> polar angle) coordinates: the maximum
// Rec.709 component of the sky at that direction, times the probe mapping's
// solid-angle factor sin(theta) / theta.  The mismatch only shapes the
// importance function; sampling and pdf agree with each other, so the
// estimator is unbiased.
//
// The table is built on the host in double precision; sampling and the pdf
// run on the device in double precision over it (a few dozen operations per
// light sample), so the GPU takes the oracle's decisions.
#pragma once

#include "output_color.h"

#include <cmath>

#ifndef RISE_HD
#define RISE_HD __host__ __device__
#endif

namespace rise_hip {

constexpr int kEnvRes = 64;

struct EnvTable {
    double cond[kEnvRes][kEnvRes + 1];  // conditional CDF per row (last entry 1)
    double marg[kEnvRes + 1];           // marginal CDF over rows (last entry 1)
    double total;                       // integrated weight (> 0: valid)
};

// The Preetham sky of HosekWilkieSkyModel in double precision (host):
// Perez coefficients for Y, x, y, zenith values, the solar zenith angle, the
// sun direction.
struct SkyModelD {
    double p[3][5];
    double zen[3];  // Yz, xz, yz
    double theta_s;
    double sun[3];
};

// HosekWilkieSkyModel::PerezF (HosekWilkieSkyModel.cpp:165-175).
inline double perez_d(const double c[5], double cos_theta, double gamma)
{
    const double sc = std::max(1e-3, cos_theta);
    const double t1 = 1.0 + c[0] * std::exp(c[1] / sc);
    const double cg = std::cos(gamma);
    return t1 * (1.0 + c[2] * std::exp(c[3] * gamma) + c[4] * cg * cg);
}

// HosekWilkieSkyModel::IntegrateRGB (HosekWilkieSkyModel.cpp:282-301): the
// sky's xyY in direction d (unit), to XYZ, to Rec.709 through the
// reference's gamut clip (ColorUtils::XYZtoRec709RGB).  Host only.
inline void sky_rgb(const SkyModelD& m, const double d[3], double rgb[3])
{
    rgb[0] = rgb[1] = rgb[2] = 0.0;
    if (d[1] <= 0.0) return;
    const double ct = std::max(0.0, d[1]);
    const double ds = std::max(-1.0, std::min(1.0, d[0] * m.sun[0] + d[1] * m.sun[1] + d[2] * m.sun[2]));
    const double g = std::acos(ds);
    const double Y = m.zen[0] * perez_d(m.p[0], ct, g) / perez_d(m.p[0], 1.0, m.theta_s);
    const double x = m.zen[1] * perez_d(m.p[1], ct, g) / perez_d(m.p[1], 1.0, m.theta_s);
    const double y = m.zen[2] * perez_d(m.p[2], ct, g) / perez_d(m.p[2], 1.0, m.theta_s);
    if (Y < 0 || y < 1e-6) return;
    out::xyz_to_rec709((x / y) * Y, Y, ((1.0 - x - y) / y) * Y, rgb);
}

// EnvironmentSampler::Build (EnvironmentSampler.cpp:150-246).  Host only.
inline void build_env_table(const SkyModelD& m, EnvTable& t)
{
    const double kTwoPiD = 6.283185307179586476925286766559;
    const double kPiD = 3.1415926535897932384626433832795;
    for (int v = 0; v < kEnvRes; ++v) {
        t.cond[v][0] = 0.0;
        for (int u = 0; u < kEnvRes; ++u) {
            const double s = ((double)u + 0.5) / (double)kEnvRes;
            const double tc = ((double)v + 0.5) / (double)kEnvRes;
            // HWAdapterPainter::UVtoDir: equirectangular.
            const double phi = s * (2.0 * kPiD), theta = tc * kPiD;
            const double d[3] = {std::sin(theta) * std::sin(phi), std::cos(theta), std::sin(theta) * std::cos(phi)};
            double rgb[3];
            sky_rgb(m, d, rgb);
            const double lum = std::max(rgb[0], std::max(rgb[1], rgb[2]));  // ColorMath::MaxValue
            const double dx = s - 0.5, dy = -(tc - 0.5);
            const double r = std::sqrt(dx * dx + dy * dy);
            const double th = r * kTwoPiD;
            double w = th < 1e-6 ? 1.0 : std::sin(th) / th;
            if (r > 0.5) w = 0.0;
            t.cond[v][u + 1] = t.cond[v][u] + lum * w;
        }
    }
    t.marg[0] = 0.0;
    for (int v = 0; v < kEnvRes; ++v) t.marg[v + 1] = t.marg[v] + t.cond[v][kEnvRes];
    t.total = t.marg[kEnvRes];
    if (!(t.total > 0.0)) return;
    for (int v = 0; v < kEnvRes; ++v) {
        const double row = t.cond[v][kEnvRes];
        if (row > 0.0)
            for (int u = 0; u <= kEnvRes; ++u) t.cond[v][u] /= row;
        t.cond[v][kEnvRes] = 1.0;
    }
    for (int v = 0; v <= kEnvRes; ++v) t.marg[v] /= t.total;
    t.marg[kEnvRes] = 1.0;
}

// EnvironmentSampler::SampleCDF (:118-148): largest i with cdf[i] <= xi, and
// the offset inside that bin.
RISE_HD inline unsigned env_search(const double* cdf, unsigned n, double xi, double& frac)
{
    unsigned lo = 0, hi = n;
    while (lo + 1 < hi) {
        const unsigned mid = (lo + hi) / 2;
        if (cdf[mid] <= xi) lo = mid;
        else hi = mid;
    }
    const double w = cdf[lo + 1] - cdf[lo];
    frac = w > 0 ? (xi - cdf[lo]) / w : 0.0;
    return lo;
}

// The probe mapping's solid-angle density from the (s, t) density
// (:290-310, :338-352).
RISE_HD inline double env_solid_angle_pdf(double pdf_st, double s, double tc)
{
    const double kPiD = 3.1415926535897932384626433832795;
    const double dx = s - 0.5, dy = -(tc - 0.5);
    const double r = sqrt(dx * dx + dy * dy);
    const double th = r * (2.0 * kPiD);
    if (th < 1e-6 || r > 0.5) return pdf_st / (4.0 * kPiD);
    const double st = sin(th);
    if (st < 1e-10) return 0.0;
    return pdf_st * th / ((4.0 * kPiD) * kPiD * st);
}

// EnvironmentSampler::Sample (:248-312) with an identity map transform:
// direction (unit) and its solid-angle pdf.
RISE_HD inline void env_sample(const EnvTable& t, double u1, double u2, double dir[3], double& pdf)
{
    const double kTwoPiD = 6.283185307179586476925286766559;
    double fv, fu;
    const unsigned v = env_search(t.marg, kEnvRes, u2, fv);
    const unsigned u = env_search(t.cond[v], kEnvRes, u1, fu);
    const double s = ((double)u + fu) / (double)kEnvRes;
    const double tc = ((double)v + fv) / (double)kEnvRes;
    // TexCoordToDirection (:77-108).
    const double dx = s - 0.5, dy = -(tc - 0.5);
    const double rt = sqrt(dx * dx + dy * dy);
    if (rt < 1e-12) {
        dir[0] = 0.0;
        dir[1] = 0.0;
        dir[2] = -1.0;
    } else {
        const double th = rt * kTwoPiD;
        const double sth = sin(th), cth = cos(th);
        dir[0] = dx * sth / rt;
        dir[1] = dy * sth / rt;
        dir[2] = -cth;
    }
    const double pm = (t.marg[v + 1] - t.marg[v]) * (double)kEnvRes;
    const double pc = (t.cond[v][u + 1] - t.cond[v][u]) * (double)kEnvRes;
    pdf = env_solid_angle_pdf(pm * pc, s, tc);
}

// EnvironmentSampler::Pdf (:314-355) with an identity map transform.
RISE_HD inline double env_pdf(const EnvTable& t, const double d[3])
{
    double s = 0.5, tc = 0.5;
    const double den = sqrt(d[0] * d[0] + d[1] * d[1]);
    if (!(den < 1e-12)) {
        const double r = 0.159154943 * acos(fmax(-1.0, fmin(1.0, -d[2]))) / den;
        s = 0.5 + d[0] * r;
        tc = 0.5 - d[1] * r;
    }
    s = fmax(0.0, fmin(0.999999, s));
    tc = fmax(0.0, fmin(0.999999, tc));
    const unsigned u = (unsigned)fmin(s * kEnvRes, (double)(kEnvRes - 1));
    const unsigned v = (unsigned)fmin(tc * kEnvRes, (double)(kEnvRes - 1));
    const double pm = (t.marg[v + 1] - t.marg[v]) * (double)kEnvRes;
    const double pc = (t.cond[v][u + 1] - t.cond[v][u]) * (double)kEnvRes;
    return env_solid_angle_pdf(pm * pc, s, tc);
}

}  // namespace rise_hip
