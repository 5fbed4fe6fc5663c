/*
 * pt_bsdf.h -- FP64 restatement of the reference's GGX conductor and smooth
 * dielectric scattering math (TEST INFRASTRUCTURE ONLY; see pt_oracle.h).
 *
 * Follows (paths relative to the reference's src/Library/):
 *   Utilities/MicrofacetUtils.h:181-371       Smith Lambda / G1 / G2, GGX D, VNDF sample and pdf
 *   Utilities/MicrofacetEnergyLUT.h:40-426    E_ss / E_avg lookups, multiscatter lobe (Z, sampler,
 *                                             pdf), F_avg by 21-point Gauss-Legendre, F_ms
 *   Utilities/Optics.hpp:35-84                conductor Fresnel
 *   Utilities/Optics.cpp:20-160               reflection, refraction, dielectric Fresnel
 *   Materials/GGXSPF.cpp:424-790              ScatterNM / PdfNM (conductor Fresnel mode)
 *   Materials/GGXBRDF.cpp:339-498             valueNM (conductor Fresnel mode)
 *   Materials/DielectricSPF.cpp:120-287, 395-432   GenerateScatteredRay / ScatterNM (smooth)
 * The GGX tables come from rise_amd/csrc/ggx_lut_data.h (data).
 */
#ifndef RISE_PT_BSDF_H
#define RISE_PT_BSDF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double x, y, z; } rb_v3;
typedef struct { rb_v3 u, v, w; } rb_onb;

/* Surface context for the GGX functions: incoming ray direction, shading ONB
 * (w = shading normal), geometric normal, wavelength-resolved parameters. */
typedef struct rb_ggx {
    rb_v3 dir;        /* ri.ray.Dir() (unit) */
    rb_onb onb;       /* ri.onb */
    rb_v3 ng;         /* ri.vGeomNormal */
    double rd, rs;    /* diffuse and specular painter values at nm */
    double alpha;     /* isotropic roughness (alphaX = alphaY) */
    double ior, ext;  /* conductor n, k at nm */
    double amb;       /* ambient (incident-medium) IOR */
} rb_ggx;

typedef struct rb_sample {
    int ok;           /* a ray was produced */
    rb_v3 dir;
    double kray;      /* krayNM */
    double pdf;       /* mixture pdf (0 for delta) */
    int type;         /* 0 diffuse, 1 reflection, 2 refraction */
} rb_sample;

double rise_oracle_ggx_lookup_ess(double cos_theta, double alpha);
double rise_oracle_ggx_lookup_eavg(double alpha);
double rise_oracle_ggx_ms_z(double alpha);
double rise_oracle_ggx_sample_ms_cos(double alpha, double u);
double rise_oracle_ggx_ms_pdf(double cos_theta, double alpha, double z);
double rise_oracle_fresnel_conductor(double cos_i, double ni, double nt, double kt);
double rise_oracle_fresnel_avg(double ni, double nt, double kt);
double rise_oracle_ggx_fms(double f_avg, double e_avg);
double rise_oracle_ggx_d(double ax, double ay, const double h_local[3]);
double rise_oracle_ggx_g1(double ax, double ay, const double v_local[3]);
double rise_oracle_ggx_g2(double ax, double ay, const double wi_local[3], const double wo_local[3]);
/* VNDF sample (world-space micronormal) and pdf; onb9 = u, v, w as 9 doubles. */
void rise_oracle_vndf_sample(const double wi[3], const double onb9[9], double ax, double ay, double u1, double u2,
                             double m_out[3]);
double rise_oracle_vndf_pdf(const double wi[3], const double wo[3], const double onb9[9], double ax, double ay);

/* GGX material (conductor Fresnel): value, pdf and sampling at one wavelength. */
double rb_ggx_value(const rb_ggx* g, rb_v3 vin);
double rb_ggx_pdf(const rb_ggx* g, rb_v3 wo);
rb_sample rb_ggx_scatter(const rb_ggx* g, double u_lobe, double u1, double u2);

/* Flat-argument wrappers for tests: ctx = dir[3], onb u[3] v[3] w[3], ng[3],
 * rd, rs, alpha, ior, ext, amb (21 doubles). */
double rise_oracle_ggx_value(const double ctx[21], const double vin[3]);
double rise_oracle_ggx_pdf(const double ctx[21], const double wo[3]);
/* out: ok, dir[3], kray, pdf, type (7 doubles). */
void rise_oracle_ggx_scatter(const double ctx[21], double u_lobe, double u1, double u2, double out[7]);

/* Optics */
rb_v3 rb_reflect(rb_v3 vin, rb_v3 n);
int rb_refract(rb_v3 n, double ni, double nt, rb_v3* vin);
double rb_dielectric_reflectance(rb_v3 v, rb_v3 tv, rb_v3 n, double ni, double nt);
double rise_oracle_dielectric_reflectance(const double v[3], const double tv[3], const double n[3], double ni,
                                          double nt);
/* Returns 1 and rewrites vin on refraction, 0 on total internal reflection. */
int rise_oracle_refract(const double n[3], double ni, double nt, double vin[3]);

#ifdef __cplusplus
}
#endif

#endif
