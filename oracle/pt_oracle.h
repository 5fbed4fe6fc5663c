/*
 * pt_oracle.h -- FP64 CPU oracle for RISE's hero-wavelength spectral path
 * tracer (TEST INFRASTRUCTURE ONLY: tests/, smoke() and bench.py's CPU
 * baseline leg use it as the checker or as the timed CPU port; the product
 * path never links it).
 *
 * Restates, for Lambertian, GGX conductor (pt_bsdf.c) and smooth dielectric
 * surfaces and Lambertian area luminaires (paths relative to the reference's
 * src/Library/):
 *   Rendering/PathTracingSpectralRasterizer.cpp:121-198, 206-479  pixel/sample loop, XYZ film
 *   Shaders/PathTracingIntegrator.cpp:3939-4822, 4855-5126          IntegrateRayHWSS / IntegrateFromHitHWSS
 *   Shaders/PathTracingIntegrator.cpp:1561-3377                     IntegrateFromHitNM (single-wavelength
 *                                                                   loop after a dielectric hit)
 *   Materials/ScatteredRayContainer.cpp:43-92                       lobe selection by krayNM
 *   Utilities/IORStack.h                                            index-of-refraction stack
 *   Lights/LightSampler.cpp:1942-2362                               EvaluateDirectLightingNM (area luminaries)
 *   Materials/LambertianBRDF.cpp:44-93, LambertianSPF.cpp:90-172    BRDF, sampling, pdf
 *   Materials/LambertianEmitter.cpp:82-89                           one-sided emission
 *   Utilities/PathTransportUtilities.h:86-210, 251-258              RR, bounce limits, power heuristic
 *   Utilities/Color/SampledWavelengths.h:79-106                     equidistant hero wavelengths
 *   Utilities/Color/ColorUtils.cpp:270-360                          XYZFromNM, CIE_Y_Integral
 *   Utilities/Color/SpectralPacket.h:285-297                        step-SPD lookup
 *   Utilities/GeometricUtilities.cpp:85-98, 543-568                 cosine sampling, point on triangle
 *   Utilities/OrthonormalBasis3D.cpp:37-56, 138-144                 ONB from W
 *   Objects/Object.cpp:641-643                                      hit point back-off
 *   Utilities/HosekWilkieSkyModel.cpp:20-301, Rendering/HosekWilkieSpectralRadianceMap.cpp
 *                                                                   sky radiance on escape
 *   Materials/HomogeneousMedium.cpp:95-240                          coefficients, free flight, Tr
 *   Materials/HenyeyGreensteinPhaseFunction.cpp:20-110              phase function (GeometricUtilities.cpp:100-118 Perturb)
 *   Shaders/PathTracingIntegrator.cpp:364-530, 1686-1925            medium transport in the NM loop
 *   Utilities/MediumTracking.h, Utilities/MediumTransport.cpp:139-174  current medium, in-scattering NEE
 *   Lights/LightSampler.cpp:140-470                                 shadow-ray transmittance walk
 * Random numbers: the reference draws every sampler.Get1D() of its bounce
 * loops from an Owen-scrambled Sobol sequence (Sampling/SobolSequence.h,
 * Utilities/SobolSampler.h, ZSobolSampler.h; stream 16 + depth, 32
 * dimensions per stream), the pixel positions from Owen-scrambled Sobol
 * points under a per-pixel seed (Sampling/SobolSampling2D.cpp), and the hero
 * wavelength, that per-pixel seed and free-flight distances from a Mersenne
 * Twister seeded per thread, so its renders are not reproducible.  The oracle
 * draws the same Sobol numbers at the same dimensions, in the order the
 * reference consumes them (pt_oracle.c: make_stream, NEE_DIMS, bsdf_draws),
 * with the index shuffled per dimension pair unless the unshuffled sampler
 * is selected, and takes the Mersenne-Twister draws from a counter-based hash
 * of (seed, path, dimension).  The HIP renderer uses the same functions, so
 * both follow the same random decisions and their per-sample results can be
 * compared directly.  Per bounce (base 32 * (16 + depth)): a shared
 * four-wavelength vertex takes +4w..+4w+3 for the light sample of wavelength
 * w and its BSDF draws from +16; a single-wavelength vertex (after a
 * dielectric hit every wavelength goes on as its own path) takes +0..+3 for
 * its light sample and its BSDF draws from +4, or +0..+2 at a dielectric;
 * the lanes of one path share these draws, as the reference's branches do.
 */
#ifndef RISE_PT_ORACLE_H
#define RISE_PT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ro_spd {
    double lambda_begin, lambda_end; /* SpectralPacket range */
    int32_t n;                       /* number of step bins; -1: sigmoid polynomial */
    const double* amp;               /* n amplitudes (sigmoid: c0, c1, c2, scale) */
} ro_spd;

typedef struct ro_material {
    int32_t type;          /* 0 Lambertian, 1 GGX conductor, 2 smooth dielectric */
    int32_t reflect_spd;   /* Lambertian reflectance / GGX diffuse (-1: none) */
    int32_t emit_spd;      /* -1: not an emitter */
    double emit_scale;     /* LambertianEmitter scale */
    int32_t spec_spd;      /* GGX specular tint */
    int32_t ior_spd;       /* GGX conductor n / dielectric index of refraction */
    int32_t ext_spd;       /* GGX conductor k / dielectric transmittance per unit length (-1: 1) */
    double alpha;          /* GGX roughness (isotropic) */
} ro_material;

/* HomogeneousMedium with RGB coefficients and a Henyey-Greenstein phase
 * function.  Its single-wavelength coefficients are the luminances of the
 * RGB triples at every wavelength (GetCoefficientsNM, HomogeneousMedium.cpp:
 * 95-115): sigma_t = Y(sigma_a + sigma_s), sigma_s = Y(sigma_s). */
typedef struct ro_medium {
    double sigma_a[3], sigma_s[3];
    double g;  /* HG asymmetry */
} ro_medium;

typedef struct ro_scene {
    const float* tris;          /* n_tris * 9 */
    int64_t n_tris;
    const int32_t* tri_object;  /* object index per triangle */
    int32_t n_objects;
    const int32_t* object_material;
    const uint8_t* object_double_sided;
    const ro_material* materials;
    int32_t n_materials;
    const ro_spd* spds;
    int32_t n_spds;
    double cam_pos[3], cam_look[3], cam_up[3], cam_fov_deg;
    int32_t width, height;
    int32_t spp_total;          /* samples per pixel of the whole render (path id layout) */
    double lambda_begin, lambda_end;
    int32_t max_diffuse_bounce, max_glossy_bounce, max_transmission_bounce;
    int32_t rr_min_depth;
    double rr_threshold;
    uint64_t seed;
    /* Sky radiance map (HosekWilkieSpectralRadianceMap: the Preetham model,
     * Utilities/HosekWilkieSkyModel.cpp), seen by rays that leave the scene. */
    int32_t sky_enabled;
    int32_t sky_background;     /* camera rays that miss see the sky */
    double sky_elevation_deg, sky_azimuth_deg, sky_turbidity, sky_scale;
    /* Participating media inside objects (standard_object interior_medium),
     * entered through a dielectric boundary; objects that do not cast
     * shadows (casts_shadows FALSE) are skipped by shadow rays.  NULL: none. */
    const int32_t* object_medium;        /* medium index per object, -1: none */
    const uint8_t* object_casts_shadows; /* NULL: every object casts shadows */
    const ro_medium* media;
    int32_t n_media;
    /* Sampler of the bounce loops: 0 Z-order Owen-scrambled Sobol (the
     * reference's default, blue_noise_sampler), 1 per-pixel Owen-scrambled
     * Sobol, both with the index shuffled per dimension pair; 2 the Z-order
     * sampler exactly as the reference draws it (no shuffle). */
    int32_t sampler;
    /* The sky's matched sun (hosek_wilkie_skylight create_sun, power
     * sun_intensity_scale) and the scene's directional lights (7 doubles
     * each: direction toward the light, RGB colour, power). */
    int32_t sun_enabled;
    double sun_power;
    int32_t n_dir_lights;
    const double* dir_lights;
} ro_scene;

/* Counter-based uniform in [0,1) with 24 random bits (exactly representable
 * in float): splitmix64 finalizer of key * golden + seed, key = path << 32 | dim. */
double rise_oracle_rng(uint64_t seed, uint32_t path, uint32_t dim);

/* SobolSequence::HashCombine and SobolSequence::Sample (as 32 bits, and in
 * [0, 1) with 24 bits) (Sampling/SobolSequence.h:121-200). */
uint32_t rise_oracle_hash_combine(uint32_t a, uint32_t b);
uint32_t rise_oracle_sobol_u32(uint32_t index, uint32_t dim, uint32_t seed);
double rise_oracle_sobol(uint32_t index, uint32_t dim, uint32_t seed);

/* Renders samples [s0, s1) of every pixel.  xyz_sum (W*H*3) receives the
 * sum of per-sample XYZ, xyz_sq (optional) the sum of squares; counters
 * (optional, 3 uint64) extension rays, shadow rays, paths.  n_threads <= 0:
 * all online CPUs.  Returns 0 on success, -1 on invalid input. */
int rise_oracle_render(const ro_scene* scene, int32_t s0, int32_t s1, double* xyz_sum, double* xyz_sq,
                       uint64_t* counters, int n_threads);

/* XYZFromNM (ColorUtils.cpp:313-325); returns 0 outside 380-780 nm. */
int rise_oracle_xyz_from_nm(double nm, double xyz[3]);

/* Sky radiance at wavelength nm in direction dir (unit), scale included. */
double rise_oracle_sky_radiance(const ro_scene* scene, const double dir[3], double nm);

/* HenyeyGreensteinPhaseFunction: Evaluate at cos(theta) and Sample around
 * wo from two uniforms (out: unit direction). */
double rise_oracle_hg_eval(double cos_theta, double g);
void rise_oracle_hg_sample(const double wo[3], double g, double u1, double u2, double out[3]);

/* Transmittance along a shadow ray at wavelength nm from o (unit dir d) over
 * maxDist, starting inside the medium of object origin_object (-1: none):
 * the boundary walk of EvalShadowTransmittanceNM. */
double rise_oracle_shadow_transmittance(const ro_scene* scene, const double o[3], const double d[3], double max_dist,
                                        int32_t origin_object, double nm);

/* CIE_Y_Integral (ColorUtils.cpp:327-350). */
double rise_oracle_cie_y_integral(double a, double b);

/* Refractions of the last rise_oracle_render into a 4th nested dielectric:
 * the pushes a 3-object index-of-refraction stack would have to drop. */
uint64_t rise_oracle_last_ior_deep_pushes(void);

#ifdef __cplusplus
}
#endif

#endif
