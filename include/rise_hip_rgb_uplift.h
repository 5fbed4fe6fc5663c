/*
 * rise_hip_rgb_uplift.h -- RGB to spectrum uplift (Jakob & Hanika 2019
 * sigmoid-polynomial spectra) for RGB-authored materials.
 *
 * Replaces the reference's RGBToSpectrumTable (Utilities/Color/
 * RGBToSpectrumTable.cpp: table lookup with trilinear interpolation),
 * RGBSigmoidPolynomial::Eval (Utilities/Color/RGBSigmoidPolynomial.h:55-69)
 * and the offline LUT generator tools/JakobHanikaLUTGen.cpp, whose baked
 * output (RGBToSpectrumTable_LUTData.cpp) the reference tree does not carry:
 * the table is regenerated here by the same solver (Rec.709 / D65 target,
 * Gauss-Newton with Levenberg-Marquardt damping and warm starts along each
 * plane), cell for cell equal to the generator's output.
 *
 * A spectrum S(lambda) = scale * sigmoid(c0 t^2 + c1 t + c2), t = (lambda -
 * 380) / 400, sigmoid(x) = 1/2 + x / (2 sqrt(1 + x^2)), zero outside
 * [380, 780] nm, is passed to the path tracer as a rise_hip_spd_desc with
 * n = RISE_HIP_SPD_SIGMOID and amplitudes = {c0, c1, c2, scale}.
 * Host code; no GPU.
 */
#ifndef RISE_HIP_RGB_UPLIFT_H
#define RISE_HIP_RGB_UPLIFT_H

#include "rise_hip_render.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rise_hip_rgb_lut rise_hip_rgb_lut;

typedef struct rise_hip_rgb_lut_stats {
    int64_t cells;
    int64_t converged;      /* residual < 1e-4 in target RGB (black cells count as converged) */
    int64_t failed;
    double max_residual;
    double mean_residual;
    double build_ms;
} rise_hip_rgb_lut_stats;

/* Generates the Rec.709 table at `resolution` (8..256) cells per axis
 * (3 * resolution^3 cells), on `threads` host threads (0: all). */
rise_hip_status rise_hip_rgb_lut_generate(int32_t resolution, int32_t threads, rise_hip_rgb_lut** out_lut,
                                          rise_hip_rgb_lut_stats* stats);

/* Reads / writes the generator's binary ".coeff" format ("RJHL", version
 * 0x00010000, resolution, 3, 3, then 3 * resolution^3 float triples). */
rise_hip_status rise_hip_rgb_lut_read(const char* path, rise_hip_rgb_lut** out_lut);
rise_hip_status rise_hip_rgb_lut_write(const rise_hip_rgb_lut* lut, const char* path);

/* The table body (owned by the lut): index ((max_channel * res + iz) * res +
 * ix) * res + iy, three floats per cell. */
rise_hip_status rise_hip_rgb_lut_get(const rise_hip_rgb_lut* lut, int32_t* resolution, const float** coeffs);

/* RGBToSpectrumTable::operator(): linear Rec.709 RGB (clamped to [0, 1])
 * to sigmoid coefficients. */
rise_hip_status rise_hip_rgb_lut_lookup(const rise_hip_rgb_lut* lut, const double rgb[3], float coeffs[3]);

rise_hip_status rise_hip_rgb_lut_release(rise_hip_rgb_lut* lut);

/* The process-wide 64^3 Rec.709 table (the reference's resolution),
 * generated on first use (a few seconds on all host threads) and kept. */
rise_hip_status rise_hip_rgb_lut_default(const rise_hip_rgb_lut** out_lut);

/* RGBSigmoidPolynomial::Eval times scale, in double (for checks). */
double rise_hip_sigmoid_spectrum_eval(const float coeffs[3], double scale, double lambda_nm);

#ifdef __cplusplus
}
#endif

#endif /* RISE_HIP_RGB_UPLIFT_H */
