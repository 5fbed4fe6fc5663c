// pt_device.h -- device-side scene tables and shading functions of the
// spectral path tracer (Lambertian surfaces, Lambertian area luminaries,
// hero-wavelength spectral sampling with 4 wavelengths).
//
// Each function cites the reference routine whose behaviour it follows
// (paths relative to the reference's src/Library/).  Arithmetic is FP32;
// intersection certification stays FP64 (trace_core.h).
#pragma once

#include "trace_core.h"

namespace rise_hip {

constexpr int kHwss = 4;             // SampledWavelengths::N (Utilities/Color/SampledWavelengths.h:57)
constexpr int kMaxPathDepth = 128;   // PathTracingIntegrator mMaxPathDepth
constexpr float kInvPi = 0.318309886183790671538f;
constexpr float kTwoPi = 6.28318530717958647692f;
constexpr float kNearZeroF = 1e-12f;

struct DevSpd {
    float begin, end, inv_delta;
    int n;
    int offset;  // into the amplitude table
    int pad[3];
};

enum : int { kMatLambert = 0, kMatGgx = 1, kMatDielectric = 2 };

struct DevMaterial {
    int type;
    int refl_spd;   // Lambertian reflectance / GGX diffuse (-1: none)
    int emit_spd;   // -1: none
    float emit_scale;
    int spec_spd;   // GGX specular tint
    int ior_spd;    // GGX conductor n / dielectric index (-1: the reference's defaults)
    int ext_spd;    // GGX conductor k / dielectric transmittance per unit length (-1: none)
    float alpha;    // GGX roughness
};

struct DevObject {
    int material;
    int double_sided;
    int luminary;   // -1: not a luminary
    int medium;     // interior medium, -1: none
};

// HomogeneousMedium with RGB coefficients and a Henyey-Greenstein phase
// function, at one wavelength: the luminances of sigma_a + sigma_s and of
// sigma_s (GetCoefficientsNM, Materials/HomogeneousMedium.cpp:95-115), the same
// at every wavelength.
struct DevMedium {
    float sigma_t, sigma_s, g;
    int pad;
};

// Tri48 v2.w bit 0: the triangle's object casts no shadow (skipped by the
// any-hit query of shadow rays, ObjectManager.cpp:119-124).
constexpr unsigned kTriNoShadow = 1u;

struct DevLuminary {
    int object;
    int tri_first;   // into lum_tris / lum_cdf
    int tri_count;
    float area;
    float pdf_select;   // weight / total weight
    float cdf_select;   // cumulative selection probability up to and including this luminary
    int material;
    int pad;
};

// Sky radiance map (HosekWilkieSpectralRadianceMap over the Preetham model,
// Utilities/HosekWilkieSkyModel.cpp): Perez coefficients for Y, x, y, the
// zenith values divided by the zenith Perez term, the sun direction, scale.
struct DevSky {
    int on;
    float theta_s;
    float sun[3];
    float pz[3][5];   // A..E for Y, x, y
    float zen[3];     // Yz / F_Y(zenith), xz / F_x(zenith), yz / F_y(zenith)
    float scale;
};

// A directional light (Lights/DirectionalLight.cpp): the direction toward the
// light (unit) and Rec.709 luminance of its colour times its energy, the value
// ComputeDirectLightingNM multiplies by f * cos at every wavelength.
constexpr int kMaxDirLights = 4;
struct DevDirLight {
    float d[3];
    float le;
};

struct EnvTable;

struct PtSceneView {
    const NodeSlot* nodes;
    const Tri48* tris;      // leaf order; v0.w = global id, v1.w = object id
    int depth;
    int n_lums;
    const DevObject* objects;
    const DevMaterial* materials;
    const DevSpd* spds;
    const float* spd_amp;
    const DevLuminary* lums;
    const unsigned* lum_tris;   // leaf-order triangle indices
    const float* lum_cdf;       // area CDF per luminary
    DevSky sky;
    // Participating media: table, and a BVH over the shells of the objects
    // holding one (v0.w = index, v1.w = object id) for the shadow-ray
    // transmittance walk.  has_media = 0: none.
    const DevMedium* media;
    const NodeSlot* mnodes;
    const Tri48* mtris;
    int has_media;
    // Environment light sample over the sky (env_sampler.h; null: none) and
    // the directional lights (the sky's matched sun, directional_light).
    const EnvTable* env;
    int n_dir;
    DevDirLight dir[kMaxDirLights];
};

// ---- random numbers: counter-based hash ------------------------------------
// u = top 24 bits of splitmix64(key * golden + seed), key = path << 32 | dim:
// the draws the reference takes from its Mersenne Twister (hero wavelength,
// the pixel-position points' per-pixel seed, free-flight distances).  The
// oracle (oracle/pt_oracle.c) uses the same function, so both follow the
// same random decisions.
RISE_HD inline float rng_u(unsigned long long seed, unsigned path, unsigned dim)
{
    unsigned long long x = ((((unsigned long long)path) << 32) | dim) * 0x9E3779B97F4A7C15ull + seed;
    x ^= x >> 30;
    x *= 0xBF58476D1CE4E5B9ull;
    x ^= x >> 27;
    x *= 0x94D049BB133111EBull;
    x ^= x >> 31;
    return (float)(unsigned)(x >> 40) * (1.0f / 16777216.0f);
}

// The same hash's top 32 bits (a per-pixel scramble seed).
RISE_HD inline unsigned rng_u32(unsigned long long seed, unsigned path, unsigned dim)
{
    unsigned long long x = ((((unsigned long long)path) << 32) | dim) * 0x9E3779B97F4A7C15ull + seed;
    x ^= x >> 30;
    x *= 0xBF58476D1CE4E5B9ull;
    x ^= x >> 27;
    x *= 0x94D049BB133111EBull;
    x ^= x >> 31;
    return (unsigned)(x >> 32);
}

// ---- Owen-scrambled Sobol (Sampling/SobolSequence.h:121-200) ----------------
// Every sampler.Get1D() of the reference's bounce loops.  Padded Sobol:
// dimension d uses Sobol dimension d & 1 (van der Corput, or the (0, 2)
// sequence's second dimension) scrambled by the Laine-Karras hash with seed
// HashCombine(seed, d).  The oracle (oracle/pt_oracle.c) computes the same
// numbers; both quantize to 24 bits, as rng_u does.
RISE_HD inline unsigned reverse_bits32(unsigned n) { return __builtin_bitreverse32(n); }

RISE_HD inline unsigned sobol_dim1(unsigned index)
{
    // Direction number i has bit 31 - j set iff C(i, j) is odd, i.e. iff j is
    // a bit-subset of i (Lucas), so bit 31 - j of the result is the XOR of the
    // index bits at every superset of j: a superset-sum over the five bits of
    // the position, then a bit reversal.  Equal to the XOR of the direction
    // numbers of the set bits for every 32-bit index.
    index ^= (index >> 1) & 0x55555555u;
    index ^= (index >> 2) & 0x33333333u;
    index ^= (index >> 4) & 0x0f0f0f0fu;
    index ^= (index >> 8) & 0x00ff00ffu;
    index ^= (index >> 16) & 0x0000ffffu;
    return reverse_bits32(index);
}

RISE_HD inline unsigned hash_combine(unsigned a, unsigned b)
{
    a ^= b + 0x9e3779b9u + (a << 6) + (a >> 2);
    a ^= a >> 16;
    a *= 0x85ebca6bu;
    a ^= a >> 13;
    a *= 0xc2b2ae35u;
    a ^= a >> 16;
    return a;
}

RISE_HD inline unsigned owen_scramble(unsigned v, unsigned seed)
{
    v = reverse_bits32(v);
    v ^= v * 0x3d20adeau;
    v += seed;
    v *= (seed >> 16) | 1u;
    v ^= v * 0x05526c56u;
    v ^= v * 0x53a22864u;
    return reverse_bits32(v);
}

RISE_HD inline float sobol_f(unsigned index, unsigned dim, unsigned seed)
{
    const unsigned v = (dim & 1u) ? sobol_dim1(index) : reverse_bits32(index);
    return (float)(owen_scramble(v, hash_combine(seed, dim)) >> 8) * (1.0f / 16777216.0f);
}

// The sampler of one camera sample: Sobol index, scramble seed, and the
// pixel's block of indices inside which the index is shuffled per dimension
// pair (mask 0: no shuffle, the reference's own draws).  See sample_key in
// pt_render.hip for how a path's key is made.
struct SampleKey {
    unsigned index, scramble, mask;
};

RISE_HD inline float sample_u(const SampleKey& k, unsigned dim)
{
    unsigned index = k.index;
    if (k.mask) {
        const unsigned pair_seed = hash_combine(k.scramble, 0x51ed270bu + (dim >> 1));
        index = (index & ~k.mask) | (owen_scramble(index, pair_seed) & k.mask);
    }
    return sobol_f(index, dim, k.scramble);
}

// MortonCode::Part1By1 (Utilities/MortonCode.h).
RISE_HD inline unsigned morton_part1by1(unsigned x)
{
    x &= 0x0000ffffu;
    x = (x ^ (x << 8)) & 0x00ff00ffu;
    x = (x ^ (x << 4)) & 0x0f0f0f0fu;
    x = (x ^ (x << 2)) & 0x33333333u;
    x = (x ^ (x << 1)) & 0x55555555u;
    return x;
}

// Sampler dimensions per vertex, in the reference's consumption order after
// StartStream(16 + depth): 4 per light sample (alias pick, then the three
// numbers of UniformRandomPoint), then the BSDF's draws (Lambertian 2; GGX 3:
// the lobe, then 2; smooth dielectric 2, unused; the phase function 2), the
// lobe choice (not at a volume scatter), Russian roulette (not at a
// dielectric).
constexpr unsigned kNeeDims = 4u;
RISE_HD inline unsigned stream_base(int depth) { return 32u * (unsigned)(16 + depth); }

// ---- small float vector ----------------------------------------------------
struct f3 {
    float x, y, z;
};
RISE_HD inline f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
RISE_HD inline f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
RISE_HD inline f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
RISE_HD inline f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
RISE_HD inline f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
RISE_HD inline float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RISE_HD inline f3 cross3(f3 a, f3 b)
{
    return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
RISE_HD inline float len3(f3 a) { return sqrtf(dot3(a, a)); }
RISE_HD inline f3 norm3(f3 a)
{
    const float m = len3(a);
    return m > 0.0f ? a * (1.0f / m) : a;
}

struct Onb {
    f3 u, v, w;
};

// OrthonormalBasis3D::CreateFromW (Utilities/OrthonormalBasis3D.cpp:37-56).
RISE_HD inline Onb onb_from_w(f3 n)
{
    Onb b;
    b.w = norm3(n);
    const f3 cu = mk3(1, 0, 0), cv = mk3(0, 1, 0), cw = mk3(0, 0, 1);
    if (fabsf(dot3(b.w, cv) - 1.0f) < kNearZeroF) {
        b.u = cw;
        b.v = cu;
    } else if (fabsf(dot3(b.w, cv) + 1.0f) < kNearZeroF) {
        b.u = cu;
        b.v = cw;
    } else {
        b.u = norm3(cross3(b.w, cv));
        b.v = norm3(cross3(b.w, b.u));
    }
    return b;
}

// OrthonormalBasis3D::Transform (OrthonormalBasis3D.cpp:138-144).
RISE_HD inline f3 onb_xform(const Onb& b, f3 a) { return b.u * a.x + b.v * a.y + b.w * a.z; }

// PowerHeuristic (Utilities/PathTransportUtilities.h:251-258).
RISE_HD inline float power_heuristic(float pa, float pb)
{
    const float a2 = pa * pa;
    return a2 / (a2 + pb * pb);
}

// SpectralPacket::ValueAtNM (Utilities/Color/SpectralPacket.h:285-297).
RISE_HD inline float spd_value(const PtSceneView& s, int spd, float nm)
{
    const DevSpd d = s.spds[spd];
    if (nm < d.begin || nm > d.end) return 0.0f;
    if (d.n < 0) {
        // Sigmoid-polynomial spectrum (RGBSigmoidPolynomial::Eval): amplitudes
        // hold c0, c1, c2 and a scale; the range is [380, 780] nm.
        const float* c = s.spd_amp + d.offset;
        const float t = (nm - 380.0f) * (1.0f / 400.0f);
        const float x = (c[0] * t + c[1]) * t + c[2];
        return c[3] * (0.5f + x / (2.0f * sqrtf(1.0f + x * x)));
    }
    int idx = (int)((nm - d.begin) * d.inv_delta);
    if (idx >= d.n) idx = d.n - 1;
    if (idx < 0) idx = 0;
    return s.spd_amp[d.offset + idx];
}

// Offsets a surface point along the geometric normal so a ray leaving it
// does not re-hit the surface it starts on: Waechter and Binder, "A Fast and
// Robust Method for Avoiding Self-Intersection", Ray Tracing Gems (2019),
// ch. 6 (integer ULP offset, a fixed epsilon near the origin).  The reference
// instead backs off 1e-12 and advances 1e-8 in FP64 (Objects/Object.cpp:641,
// PathTracingIntegrator.cpp:4812), which FP32 cannot resolve.
RISE_HD inline f3 offset_ray(f3 p, f3 n)
{
    const float origin = 1.0f / 32.0f, float_scale = 1.0f / 65536.0f, int_scale = 256.0f;
    const int ox = (int)(int_scale * n.x), oy = (int)(int_scale * n.y), oz = (int)(int_scale * n.z);
    const float px = bits_to_float((unsigned)(float_to_bits(p.x) + (p.x < 0 ? -ox : ox)));
    const float py = bits_to_float((unsigned)(float_to_bits(p.y) + (p.y < 0 ? -oy : oy)));
    const float pz = bits_to_float((unsigned)(float_to_bits(p.z) + (p.z < 0 ? -oz : oz)));
    return mk3(fabsf(p.x) < origin ? p.x + float_scale * n.x : px, fabsf(p.y) < origin ? p.y + float_scale * n.y : py,
               fabsf(p.z) < origin ? p.z + float_scale * n.z : pz);
}

// Surface record at a hit (the fields of RayIntersectionGeometric the
// Lambertian routines read).
struct Surf {
    f3 p;       // hit point
    f3 dir;     // incoming direction (ri.ray.Dir())
    f3 n;       // shading normal (face normal; ...Specializations.h:194-200)
    f3 ng;      // geometric normal (:207-213)
    Onb frame;  // ri.onb
    int object;
    int material;
};

// Geometric normal facing against the incoming ray (LambertianBRDF.cpp:51-53).
RISE_HD inline f3 geom_against_ray(const Surf& h, f3 fallback)
{
    const f3 raw = dot3(h.ng, h.ng) > 1e-12f ? h.ng : fallback;
    return dot3(raw, h.dir) < 0.0f ? raw : -raw;
}

// ShouldReflect (Materials/LambertianBRDF.cpp:44-64).
RISE_HD inline bool should_reflect(const Surf& h, f3 vin)
{
    const f3 n = h.frame.w;
    const f3 v = norm3(vin);
    const f3 r = norm3(-h.dir);
    const float nr = dot3(n, r), nv = dot3(n, v);
    const f3 g = geom_against_ray(h, n);
    if ((nr <= -kNearZeroF && nv <= -kNearZeroF) || (nr >= kNearZeroF && nv >= kNearZeroF))
        return dot3(v, g) > 0.0f && dot3(r, g) > 0.0f;
    return false;
}

// LambertianBRDF::valueNM (LambertianBRDF.cpp:76-84).
RISE_HD inline float lambert_value(const PtSceneView& s, const Surf& h, int refl_spd, f3 vin, float nm)
{
    if (!should_reflect(h, vin)) return 0.0f;
    return spd_value(s, refl_spd, nm) * kInvPi;
}

// LambertianSPF::Pdf (Materials/LambertianSPF.cpp:147-163).
RISE_HD inline float lambert_pdf(const Surf& h, f3 wo)
{
    const bool front = dot3(h.dir, h.frame.w) <= kNearZeroF;
    const float c = front ? dot3(wo, h.frame.w) : -dot3(wo, h.frame.w);
    const f3 neff = front ? h.frame.w : -h.frame.w;
    const f3 g = geom_against_ray(h, neff);
    return (c > 0.0f && dot3(wo, g) > 0.0f) ? c * kInvPi : 0.0f;
}

// LambertianEmitter::emittedRadianceNM (Materials/LambertianEmitter.cpp:82-89).
RISE_HD inline float emitted(const PtSceneView& s, const DevMaterial& m, f3 out, f3 N, float nm)
{
    if (dot3(out, N) <= 0.0f) return 0.0f;
    return spd_value(s, m.emit_spd, nm) * kInvPi * m.emit_scale;
}

// SampledWavelengths::SampleEquidistant (SampledWavelengths.h:79-106).
RISE_HD inline void wavelengths(float lambda0, float a, float b, float lam[kHwss])
{
    const float range = b - a, delta = range / (float)kHwss;
    lam[0] = lambda0;
    for (int i = 1; i < kHwss; ++i) lam[i] = a + fmodf(lambda0 - a + (float)i * delta, range);
}

}  // namespace rise_hip
