// ref_pt_pins.cpp -- C entry points over the reference's own path-tracing
// building blocks, compiled from its sources where they lie (see
// oracle/ref/Makefile), so the FP64 oracle's restatements of them can be
// compared value for value.  TEST INFRASTRUCTURE ONLY; never shipped.
//
// Functions exercised (paths relative to src/Library/):
//   HosekWilkieSkyModel::SampleRadiance / IntegrateRGB  Utilities/HosekWilkieSkyModel.cpp:205-301
//   Optics::CalculateConductorReflectance         Utilities/Optics.hpp:35-84
//   Optics::CalculateDielectricReflectance,
//     CalculateRefractedRay, CalculateReflectedRay Utilities/Optics.cpp:20-160
//   GGX_D_Aniso, GGX_G1_Aniso, GGX_G2_Aniso,
//     VNDF_Sample_Aniso, VNDF_Pdf_Aniso           Utilities/MicrofacetUtils.h:188-371
//   MicrofacetEnergyLUT::LookupEss / LookupEavg / MSLobeZ / SampleMSCosTheta /
//     MSPdf / ComputeFresnelAvg / ComputeFms      Utilities/MicrofacetEnergyLUT.h:90-426
//   PathTransportUtilities::EvaluateRussianRoulette / PowerHeuristic
//                                                 Utilities/PathTransportUtilities.h:86-258
//   SampledWavelengths::SampleEquidistant         Utilities/Color/SampledWavelengths.h:56-90
//   OrthonormalBasis3D::CreateFromW               Utilities/OrthonormalBasis3D.cpp
//   SobolSequence::Sample / HashCombine           Sampling/SobolSequence.h:121-200
//   MortonCode::Morton2D                          Utilities/MortonCode.h

#include "Utilities/HosekWilkieSkyModel.h"
#include "Utilities/MicrofacetEnergyLUT.h"
#include "Utilities/MicrofacetUtils.h"
#include "Utilities/Optics.h"
#include "Utilities/OrthonormalBasis3D.h"
#include "Utilities/PathTransportUtilities.h"
#include "Utilities/Color/SampledWavelengths.h"
#include "Utilities/MortonCode.h"
#include "Sampling/SobolSequence.h"

#include <cmath>

using namespace RISE;

namespace {

Vector3 v3(const double* p) { return Vector3(p[0], p[1], p[2]); }

void put(const Vector3& v, double* out)
{
    out[0] = v.x;
    out[1] = v.y;
    out[2] = v.z;
}

OrthonormalBasis3D onb9(const double* p) { return OrthonormalBasis3D(v3(p), v3(p + 3), v3(p + 6)); }

}  // namespace

extern "C" {

// Sky radiance of hosek_wilkie_skylight's radiance map for a unit world
// direction: HosekWilkieSpectralRadianceMap::GetRadianceNM with an identity
// transform (Rendering/HosekWilkieSpectralRadianceMap.cpp:127-133).
double rise_ref_sky_radiance(double elev_deg, double az_deg, double turbidity, double scale, const double dir[3],
                             double nm)
{
    const HosekWilkieSkyModel m(elev_deg * (M_PI / 180.0), az_deg * (M_PI / 180.0), turbidity,
                                RISEPel(0.3, 0.3, 0.3));
    return m.SampleRadiance(Vector3Ops::Normalize(v3(dir)), nm) * scale;
}

// The model's Rec.709 colour for a direction (IntegrateRGB); rgb[3] out.
void rise_ref_sky_rgb(double elev_deg, double az_deg, double turbidity, const double dir[3], double rgb[3])
{
    const HosekWilkieSkyModel m(elev_deg * (M_PI / 180.0), az_deg * (M_PI / 180.0), turbidity,
                                RISEPel(0.3, 0.3, 0.3));
    const RISEPel c = m.IntegrateRGB(v3(dir));
    rgb[0] = c.r;
    rgb[1] = c.g;
    rgb[2] = c.b;
}

void rise_ref_sun_direction(double elev_deg, double az_deg, double out[3])
{
    const HosekWilkieSkyModel m(elev_deg * (M_PI / 180.0), az_deg * (M_PI / 180.0), 3.0, RISEPel(0.3, 0.3, 0.3));
    put(m.SunDirection(), out);
}

double rise_ref_fresnel_conductor(const double v[3], const double n[3], double ni, double nt, double kt)
{
    return Optics::CalculateConductorReflectance<Scalar>(v3(v), v3(n), ni, nt, kt);
}

double rise_ref_dielectric_reflectance(const double v[3], const double tv[3], const double n[3], double ni,
                                       double nt)
{
    return Optics::CalculateDielectricReflectance(v3(v), v3(tv), v3(n), ni, nt);
}

// Returns 1 and writes the refracted direction, 0 on total internal reflection.
int rise_ref_refract(const double n[3], double ni, double nt, double vin[3])
{
    Vector3 v = v3(vin);
    const bool ok = Optics::CalculateRefractedRay(v3(n), ni, nt, v);
    if (ok) put(v, vin);
    return ok ? 1 : 0;
}

void rise_ref_reflect(const double vin[3], const double n[3], double out[3])
{
    put(Optics::CalculateReflectedRay(v3(vin), v3(n)), out);
}

double rise_ref_ggx_d(double ax, double ay, const double h_local[3])
{
    return MicrofacetUtils::GGX_D_Aniso<Scalar>(ax, ay, v3(h_local));
}

double rise_ref_ggx_g1(double ax, double ay, const double v_local[3])
{
    return MicrofacetUtils::GGX_G1_Aniso(ax, ay, v3(v_local));
}

double rise_ref_ggx_g2(double ax, double ay, const double wi_local[3], const double wo_local[3])
{
    return MicrofacetUtils::GGX_G2_Aniso(ax, ay, v3(wi_local), v3(wo_local));
}

void rise_ref_vndf_sample(const double wi[3], const double onb[9], double ax, double ay, double u1, double u2,
                          double m_out[3])
{
    put(MicrofacetUtils::VNDF_Sample_Aniso(v3(wi), onb9(onb), ax, ay, u1, u2), m_out);
}

double rise_ref_vndf_pdf(const double wi[3], const double wo[3], const double onb[9], double ax, double ay)
{
    return MicrofacetUtils::VNDF_Pdf_Aniso(v3(wi), v3(wo), onb9(onb), ax, ay);
}

double rise_ref_ggx_lookup_ess(double cos_theta, double alpha) { return MicrofacetEnergyLUT::LookupEss(cos_theta, alpha); }
double rise_ref_ggx_lookup_eavg(double alpha) { return MicrofacetEnergyLUT::LookupEavg(alpha); }
double rise_ref_ggx_ms_z(double alpha) { return MicrofacetEnergyLUT::MSLobeZ(alpha); }
double rise_ref_ggx_sample_ms_cos(double alpha, double u) { return MicrofacetEnergyLUT::SampleMSCosTheta(alpha, u); }
double rise_ref_ggx_ms_pdf(double cos_theta, double alpha, double z)
{
    return MicrofacetEnergyLUT::MSPdf(cos_theta, alpha, z);
}

double rise_ref_fresnel_avg(const double n[3], double ni, double nt, double kt)
{
    return MicrofacetEnergyLUT::ComputeFresnelAvg<Scalar>(v3(n), ni, nt, kt);
}

double rise_ref_ggx_fms(double f_avg, double e_avg) { return MicrofacetEnergyLUT::ComputeFms<Scalar>(f_avg, e_avg); }

// out: terminate flag, survival probability.
void rise_ref_russian_roulette(unsigned depth, unsigned rr_min_depth, double rr_threshold, double cur_max,
                               double prev_max, double u, double out[2])
{
    const PathTransportUtilities::RussianRouletteResult r =
        PathTransportUtilities::EvaluateRussianRoulette(depth, rr_min_depth, rr_threshold, cur_max, prev_max, u);
    out[0] = r.terminate ? 1.0 : 0.0;
    out[1] = r.survivalProb;
}

double rise_ref_power_heuristic(double pa, double pb) { return PathTransportUtilities::PowerHeuristic(pa, pb); }

void rise_ref_sample_equidistant(double u, double lambda_min, double lambda_max, double out[4])
{
    const SampledWavelengths s = SampledWavelengths::SampleEquidistant(u, lambda_min, lambda_max);
    for (int i = 0; i < 4; ++i) out[i] = s.lambda[i];
}

// OrthonormalBasis3D::CreateFromW: onb9 out = u, v, w.
void rise_ref_onb_from_w(const double w[3], double onb[9])
{
    OrthonormalBasis3D b;
    b.CreateFromW(v3(w));
    put(b.u(), onb);
    put(b.v(), onb + 3);
    put(b.w(), onb + 6);
}

double rise_ref_sobol_sample(uint32_t index, uint32_t dim, uint32_t seed)
{
    return SobolSequence::Sample(index, dim, seed);
}

uint32_t rise_ref_hash_combine(uint32_t a, uint32_t b) { return SobolSequence::HashCombine(a, b); }

uint32_t rise_ref_morton2d(uint32_t x, uint32_t y) { return MortonCode::Morton2D(x, y); }

}  // extern "C"
