// output_color.h -- the output stage's per-pixel colour pipeline, compiled for
// host and device: film XYZ -> linear Rec.709 (with the reference's gamut
// clip) -> exposure and display transform (LDR formats) -> output colour
// space -> clamp and integerize.  Double precision throughout, in the
// reference's operation order, so that quantized bytes match its writers.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#ifndef RISE_HD
#define RISE_HD __host__ __device__
#endif

namespace rise_hip {
namespace out {

struct P2 {
    double x, y;
};

constexpr double kNearZero = 1e-12;

// LineIntersection2D (Utilities/Color/ColorUtils.cpp:150-175).
RISE_HD inline bool line_intersection(P2 a0, P2 a1, P2 b0, P2 b1, P2& out)
{
    const double denom = (b1.y - b0.y) * (a1.x - a0.x) - (b1.x - b0.x) * (a1.y - a0.y);
    if (denom > -kNearZero && denom < kNearZero) return false;
    const double ua = ((b1.x - b0.x) * (a0.y - b0.y) - (b1.y - b0.y) * (a0.x - b0.x)) / denom;
    out.x = a0.x + ua * (a1.x - a0.x);
    out.y = a0.y + ua * (a1.y - a0.y);
    return true;
}

// IsPointOnLineSegment (:198-205): the x-parameter strictly inside (0, 1).
RISE_HD inline bool on_segment(P2 s, P2 e, P2 p)
{
    const double v = (p.x - s.x) / (e.x - s.x);
    return v >= kNearZero && v <= 1.0 - kNearZero;
}

// ClipAgainstGamutEdge (:207-222).
RISE_HD inline bool clip_edge(P2 c, P2 a, P2 b, P2 w, P2& out)
{
    if (!line_intersection(w, c, a, b, out)) return false;
    if (!on_segment(a, b, out)) return false;
    if (on_segment(w, out, c)) {
        out = c;
        return true;
    }
    return on_segment(w, c, out);
}

// MoveXYZIntoRec709RGBGamut + XYZtoRec709RGB (ColorUtils.cpp:351-363,
// Color.cpp:214-220): chromaticities outside the Rec.709 triangle move
// toward the white point onto its edge (red-green, red-blue, blue-green
// tried in that order), luminance kept; then the XYZ -> Rec.709 matrix.
RISE_HD inline void xyz_to_rec709(double X, double Y, double Z, double rgb[3])
{
    const P2 red{0.640, 0.330}, green{0.300, 0.600}, blue{0.150, 0.060}, white{0.3127, 0.3290};
    const double ovs = 1.0 / (X + Y + Z);
    const P2 c{X * ovs, Y * ovs};
    P2 hit{0.0, 0.0};
    if (clip_edge(c, red, green, white, hit) || clip_edge(c, red, blue, white, hit) ||
        clip_edge(c, blue, green, white, hit)) {
        const double prod = (1.0 / hit.y) * Y;
        X = prod * hit.x;
        Z = prod * (1 - hit.x - hit.y);
    }
    rgb[0] = 3.240479 * X + -1.537150 * Y + -0.498535 * Z;
    rgb[1] = -0.969256 * X + 1.875992 * Y + 0.041556 * Z;
    rgb[2] = 0.055648 * X + -0.204043 * Y + 1.057311 * Z;
}

// Display transforms (Rendering/DisplayTransform.h): non-finite and
// non-positive inputs are 0.
RISE_HD inline double sanitise(double x) { return ((x - x) == 0.0 && x > 0.0) ? x : 0.0; }  // x - x: NaN unless finite

RISE_HD inline double clamp01(double y) { return fmin(fmax(y, 0.0), 1.0); }

RISE_HD inline double display_curve(int dt, double x)
{
    const double s = sanitise(x);
    switch (dt) {
    case 0:  // none
        return s;
    case 1:  // Reinhard
        return s / (1.0 + s);
    case 3: {  // AgX (scalar sigmoid in log10)
        const double lx = log10(s + 1e-10);
        return clamp01(0.5 + 0.5 * tanh(1.0 * (lx - 0.0)));
    }
    case 4: {  // Hable, white point 11.2
        auto curve = [](double v) {
            const double A = 0.15, B = 0.50, C = 0.10, D = 0.20, E = 0.02, F = 0.30;
            return ((v * (A * v + C * B) + D * E) / (v * (A * v + B) + D * F)) - E / F;
        };
        return clamp01(curve(s * 2.0) / curve(11.2));
    }
    case 2:  // ACES (Narkowicz fit), also the fallback of unknown values
    default: {
        const double num = s * (2.51 * s + 0.03);
        const double den = s * (2.43 * s + 0.59) + 0.14;
        return clamp01(num / den);
    }
    }
}

// SRGBTransferFunction (ColorUtils.cpp:20-31).
RISE_HD inline double srgb_oetf(double x)
{
    if (x == 0) return 0;
    if (x <= 0.003040247678) return 12.92 * x;
    return 1.055 * pow(x, 1.0 / 2.4) - 0.055;
}

struct Params {
    int color_space;    // 0 linear Rec.709, 1 sRGB
    int display;        // display transform (applied when `tone` is set)
    int tone;           // LDR with a display transform or exposure: DisplayTransformWriter
    double exposure;    // 2^EV
};

// The pixel as the writer sees it before integerizing: RISEPel after the
// optional DisplayTransformWriter, converted to the output colour space.
RISE_HD inline void pixel_colour(const Params& p, float fx, float fy, float fz, double out[3])
{
    double c[3];
    xyz_to_rec709((double)fx, (double)fy, (double)fz, c);
    if (p.tone)
        for (int k = 0; k < 3; ++k) c[k] = display_curve(p.display, c[k] * p.exposure);
    if (p.color_space == 1)
        for (int k = 0; k < 3; ++k) c[k] = srgb_oetf(c[k]);
    for (int k = 0; k < 3; ++k) out[k] = c[k];
}

// Color_Template::Integerize (Color_Template.h:129-134): clamp to [0, 1],
// scale, truncate.
template <class T>
RISE_HD inline T integerize(double v, double maxv)
{
    if (v > 1.0) v = 1.0;
    if (v < 0.0) v = 0.0;
    return (T)(v * maxv);
}

// HDRWriter::WriteHDRPixelToBuffer (RasterImages/HDRWriter.h:37-58): shared
// exponent of the largest channel, mantissas scaled by 255.9999.
RISE_HD inline void rgbe(const double c[3], uint8_t out[4])
{
    double v = c[0] > c[1] ? (c[0] > c[2] ? c[0] : c[2]) : (c[1] > c[2] ? c[1] : c[2]);
    if (v < 1e-32) {
        out[0] = out[1] = out[2] = out[3] = 0;
        return;
    }
    int e;
    v = frexp(v, &e) * 255.9999 / v;
    out[0] = (uint8_t)(c[0] < 0 ? 0 : c[0] * v);
    out[1] = (uint8_t)(c[1] < 0 ? 0 : c[1] * v);
    out[2] = (uint8_t)(c[2] < 0 ? 0 : c[2] * v);
    out[3] = (uint8_t)(e + 128);
}

}  // namespace out
}  // namespace rise_hip
