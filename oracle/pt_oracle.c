/*
 * pt_oracle.c -- FP64 restatement of RISE's HWSS spectral path tracer for
 * Lambertian, GGX conductor and smooth dielectric surfaces and Lambertian
 * area luminaires.  TEST INFRASTRUCTURE
 * ONLY.  See pt_oracle.h for the reference lines followed and the random
 * number contract.  Intersections are brute force over all triangles with
 * the ray-cast oracle's semantics (rt_oracle.c), ties to the lowest id.
 */
#include "pt_oracle.h"
#include "pt_bsdf.h"
#include "rt_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#define NZ 1e-12
#define INV_PI 0.31830988618379067154
#define TWO_PI 6.28318530717958647692
#define HWSS_N 4
#define MAX_PATH_DEPTH 128 /* PathTracingIntegrator mMaxPathDepth */

/* ---- random numbers ------------------------------------------------------ */

double rise_oracle_rng(uint64_t seed, uint32_t path, uint32_t dim)
{
    uint64_t x = ((((uint64_t)path) << 32) | dim) * 0x9E3779B97F4A7C15ull + seed;
    x ^= x >> 30;
    x *= 0xBF58476D1CE4E5B9ull;
    x ^= x >> 27;
    x *= 0x94D049BB133111EBull;
    x ^= x >> 31;
    return (double)(x >> 40) * (1.0 / 16777216.0);
}

/* The same hash's top 32 bits: a per-pixel scramble seed. */
static uint32_t rng_u32(uint64_t seed, uint32_t path, uint32_t dim)
{
    uint64_t x = ((((uint64_t)path) << 32) | dim) * 0x9E3779B97F4A7C15ull + seed;
    x ^= x >> 30;
    x *= 0xBF58476D1CE4E5B9ull;
    x ^= x >> 27;
    x *= 0x94D049BB133111EBull;
    x ^= x >> 31;
    return (uint32_t)(x >> 32);
}

/* ---- Owen-scrambled Sobol (Sampling/SobolSequence.h:121-200) -------------- */

static uint32_t reverse_bits(uint32_t n)
{
    n = (n << 16) | (n >> 16);
    n = ((n & 0x00ff00ffu) << 8) | ((n & 0xff00ff00u) >> 8);
    n = ((n & 0x0f0f0f0fu) << 4) | ((n & 0xf0f0f0f0u) >> 4);
    n = ((n & 0x33333333u) << 2) | ((n & 0xccccccccu) >> 2);
    n = ((n & 0x55555555u) << 1) | ((n & 0xaaaaaaaau) >> 1);
    return n;
}

/* Sobol dimension 1: direction numbers d_0 = 2^31, d_{i+1} = d_i ^ (d_i >> 1). */
static uint32_t sobol_dim1(uint32_t index)
{
    uint32_t v = 0, d = 1u << 31;
    for (; index; index >>= 1, d ^= d >> 1)
        if (index & 1u) v ^= d;
    return v;
}

/* SobolSequence::HashCombine: golden-ratio mix, then the MurmurHash3 fmix32 finalizer. */
uint32_t rise_oracle_hash_combine(uint32_t a, uint32_t b)
{
    a ^= b + 0x9e3779b9u + (a << 6) + (a >> 2);
    a ^= a >> 16;
    a *= 0x85ebca6bu;
    a ^= a >> 13;
    a *= 0xc2b2ae35u;
    a ^= a >> 16;
    return a;
}

/* Laine-Karras hash on the bit-reversed value (Burley 2020, Listing 1). */
static uint32_t owen_scramble(uint32_t v, uint32_t seed)
{
    v = reverse_bits(v);
    v ^= v * 0x3d20adeau;
    v += seed;
    v *= (seed >> 16) | 1u;
    v ^= v * 0x05526c56u;
    v ^= v * 0x53a22864u;
    return reverse_bits(v);
}

/* SobolSequence::Sample as 32 bits: padded Sobol (dimension & 1), scrambled
 * with HashCombine(seed, dimension). */
uint32_t rise_oracle_sobol_u32(uint32_t index, uint32_t dim, uint32_t seed)
{
    const uint32_t v = (dim & 1u) ? sobol_dim1(index) : reverse_bits(index);
    return owen_scramble(v, rise_oracle_hash_combine(seed, dim));
}

/* The sample in [0, 1), quantized to 24 bits as the HIP renderer's float draws
 * are (the reference keeps 32; the difference is below 2^-24). */
double rise_oracle_sobol(uint32_t index, uint32_t dim, uint32_t seed)
{
    return (double)(rise_oracle_sobol_u32(index, dim, seed) >> 8) * (1.0 / 16777216.0);
}

/* MortonCode::Part1By1 / Morton2D / RoundUpPow2 / Log2Int (Utilities/MortonCode.h). */
static uint32_t part1by1(uint32_t x)
{
    x &= 0x0000ffffu;
    x = (x ^ (x << 8)) & 0x00ff00ffu;
    x = (x ^ (x << 4)) & 0x0f0f0f0fu;
    x = (x ^ (x << 2)) & 0x33333333u;
    x = (x ^ (x << 1)) & 0x55555555u;
    return x;
}

static uint32_t log2_pow2_ceil(uint32_t v)
{
    uint32_t p = 1, l = 0;
    while (p < v) {
        p <<= 1;
        ++l;
    }
    return l;
}

/* The random streams of one camera sample.  Sampler draws (every
 * sampler.Get1D() of the bounce loops) come from the Owen-scrambled Sobol
 * sequence at (index, scramble): PathTracingSpectralRasterizer::IntegratePixel
 * (:229-345) takes index = (Morton(x, y) << log2(spp rounded up to a power of
 * two)) | sample and pixel seed HashCombine(Morton, 0) for the Z-order sampler
 * (the default, blue_noise_sampler), index = sample and seed HashCombine(x, y)
 * for the plain one; the render seed, when not 0, is folded into the pixel
 * seed with HashCombine so independent renders scramble differently.
 *
 * Shuffling: the reference pads with Sobol dimensions 0 and 1 at the SAME
 * index for every dimension pair, so dimensions 0, 2, 4, ... are scrambles of
 * one van der Corput point and are strongly dependent; with the Laine-Karras
 * hash standing in for random Owen scrambling this is biased (-1.0% on the
 * spectral furnace's closed form, tests/test_oracle_pt.py).  The default
 * samplers therefore also shuffle the index per dimension pair, as Burley
 * (JCGT 2020, section 4) does: the sample's position inside the pixel's block
 * of 2^log2(spp) indices is Owen-scrambled with a per-pair seed, a permutation
 * of the block, so each pixel still uses exactly its own block of the
 * sequence.  SAMPLER_ZSOBOL_UNSHUFFLED keeps the reference's draws unchanged.
 *
 * The draws the reference takes from its Mersenne Twister (hero wavelength,
 * the per-pixel seed of the pixel-position Sobol points, free-flight
 * distances) come from the counter hash at (seed, path). */
typedef struct {
    uint64_t seed;
    uint32_t path;
    uint32_t index, scramble;
    uint32_t mask; /* the pixel's index block (2^log2(spp) - 1); 0: no shuffle */
} rstream;

static rstream make_stream(const ro_scene* s, uint32_t px, uint32_t py, uint32_t si)
{
    rstream r;
    r.seed = s->seed;
    r.path = (uint32_t)(((uint64_t)py * (uint64_t)s->width + px) * (uint64_t)s->spp_total + si);
    uint32_t pseed = rise_oracle_hash_combine(px, py);
    const uint32_t l2 = log2_pow2_ceil((uint32_t)(s->spp_total > 0 ? s->spp_total : 1));
    r.index = si;
    r.mask = s->sampler == 2 ? 0u : (l2 >= 32 ? 0xffffffffu : (1u << l2) - 1u);
    if (s->sampler != 1 && px <= 0xffffu && py <= 0xffffu) {
        const uint32_t mi = part1by1(px) | (part1by1(py) << 1);
        if (l2 < 32 && ((uint64_t)mi << l2) < (1ull << 32)) {
            r.index = (mi << l2) | si;
            pseed = rise_oracle_hash_combine(mi, 0u);
        }
    }
    const uint32_t s32 = (uint32_t)s->seed ^ (uint32_t)(s->seed >> 32);
    r.scramble = s32 ? rise_oracle_hash_combine(pseed, s32) : pseed;
    return r;
}

static double u_sampler(const rstream* r, uint32_t dim)
{
    uint32_t index = r->index;
    if (r->mask) {
        const uint32_t pair_seed = rise_oracle_hash_combine(r->scramble, 0x51ed270bu + (dim >> 1));
        index = (index & ~r->mask) | (owen_scramble(index, pair_seed) & r->mask);
    }
    return rise_oracle_sobol(index, dim, r->scramble);
}
static double u_indep(const rstream* r, uint32_t dim) { return rise_oracle_rng(r->seed, r->path, dim); }

/* Sampler dimensions of one vertex, in the order the reference consumes them
 * after StartStream(16 + depth) (PathTracingIntegrator.cpp:4112, :1686):
 * per light sample 4 (alias-table pick, then the three numbers of
 * UniformRandomPoint; LightSampler.cpp:2090-2215), then the BSDF's own draws
 * (Lambertian 2: LambertianSPF.cpp:90-99; GGX 3: the lobe, then 2,
 * GGXSPF.cpp:479-519; smooth dielectric 2, unused: DielectricSPF.cpp:421; the
 * phase function 2: HenyeyGreensteinPhaseFunction.cpp:81-93), then the
 * lobe choice (RandomlySelect, :4619 / :2670 / :2945) -- not at a volume
 * scatter -- then Russian roulette (:4757 / :3224 / :1870) -- not at a
 * dielectric.  A shared (four-wavelength) vertex takes four light samples,
 * a single-wavelength one (IntegrateFromHitNM) one; every single-wavelength
 * lane of a path restarts the same stream at each depth, so the lanes of one
 * path share their sampler draws, as the reference's do. */
#define NEE_DIMS 4u
#define STREAM_BASE(depth) (32u * (uint32_t)(16 + (depth)))

/* ---- CIE 1931 2-degree colour matching functions, 380-780 nm, 5 nm ------- */
/* (ColorUtils.cpp:81; the z table is zero past 630 nm there, as here.) */
static const double CIE_X[81] = {
    0.0014, 0.0022, 0.0042, 0.0077, 0.0143, 0.0232, 0.0435, 0.0776, 0.1344, 0.2148, 0.2839, 0.3285, 0.3483, 0.3481,
    0.3362, 0.3187, 0.2908, 0.2511, 0.1954, 0.1421, 0.0956, 0.0580, 0.0320, 0.0147, 0.0049, 0.0024, 0.0093, 0.0291,
    0.0633, 0.1096, 0.1655, 0.2257, 0.2904, 0.3597, 0.4334, 0.5121, 0.5945, 0.6784, 0.7621, 0.8425, 0.9163, 0.9786,
    1.0263, 1.0567, 1.0622, 1.0456, 1.0026, 0.9384, 0.8544, 0.7514, 0.6424, 0.5419, 0.4479, 0.3608, 0.2835, 0.2187,
    0.1649, 0.1212, 0.0874, 0.0636, 0.0468, 0.0329, 0.0227, 0.0158, 0.0114, 0.0081, 0.0058, 0.0041, 0.0029, 0.0020,
    0.0014, 0.0010, 0.0007, 0.0005, 0.0003, 0.0002, 0.0002, 0.0001, 0.0001, 0.0001, 0.0000};
static const double CIE_Y[81] = {
    0, 0.0001, 0.0001, 0.0002, 0.0004, 0.0006, 0.0012, 0.0022, 0.004, 0.0073, 0.0116, 0.0168, 0.023, 0.0298,
    0.038, 0.048, 0.06, 0.0739, 0.091, 0.1126, 0.139, 0.1693, 0.208, 0.2586, 0.323, 0.4073, 0.503, 0.6082,
    0.71, 0.7932, 0.862, 0.9149, 0.954, 0.9803, 0.995, 1, 0.995, 0.9786, 0.952, 0.9154, 0.87, 0.8163,
    0.757, 0.6949, 0.631, 0.5668, 0.503, 0.4412, 0.381, 0.321, 0.265, 0.217, 0.175, 0.1382, 0.107, 0.0816,
    0.061, 0.0446, 0.032, 0.0232, 0.017, 0.0119, 0.0082, 0.0057, 0.0041, 0.0029, 0.0021, 0.0015, 0.001, 0.0007,
    0.0005, 0.0004, 0.0002, 0.0002, 0.0001, 0.0001, 0.0001, 0, 0, 0, 0};
static const double CIE_Z[81] = {
    0.0065, 0.0105, 0.0201, 0.0362, 0.0679, 0.1102, 0.2074, 0.3713, 0.6456, 1.0391, 1.3856, 1.623, 1.7471, 1.7826,
    1.7721, 1.7441, 1.6692, 1.5281, 1.2876, 1.0419, 0.813, 0.6162, 0.4652, 0.3533, 0.272, 0.2123, 0.1582, 0.1117,
    0.0782, 0.0573, 0.0422, 0.0298, 0.0203, 0.0134, 0.0087, 0.0057, 0.0039, 0.0027, 0.0021, 0.0018, 0.0017, 0.0014,
    0.0011, 0.001, 0.0008, 0.0006, 0.0003, 0.0002, 0.0002, 0.0001, 0};

/* InterpSPDIndices + ApplySPDFunction (ColorUtils.cpp:270-298, ColorUtils.h:52-62). */
int rise_oracle_xyz_from_nm(double nm, double xyz[3])
{
    if (nm < 380.0 || nm > 780.0) return 0;
    double x = (nm - 380.0) / 5.0;
    unsigned lo = (unsigned)floor(x);
    unsigned hi = lo + 1;
    if (hi > 80) hi = 80;
    const double w = x - lo;
    xyz[0] = CIE_X[lo] + (CIE_X[hi] - CIE_X[lo]) * w;
    xyz[1] = CIE_Y[lo] + (CIE_Y[hi] - CIE_Y[lo]) * w;
    xyz[2] = CIE_Z[lo] + (CIE_Z[hi] - CIE_Z[lo]) * w;
    return 1;
}

double rise_oracle_cie_y_integral(double a, double b)
{
    const double lo = a < 380.0 ? 380.0 : a;
    const double hi = b > 780.0 ? 780.0 : b;
    if (hi <= lo) return 0.0;
    const int first = (int)((lo - 380.0) / 5.0);
    const int last = (int)((hi - 380.0) / 5.0);
    double s = 0.0;
    for (int i = first; i < last; ++i) s += 0.5 * (CIE_Y[i] + CIE_Y[i + 1]) * 5.0;
    return s;
}

/* SpectralPacket::ValueAtNM (SpectralPacket.h:285-297): step bins. */
static double spd_value(const ro_spd* s, double nm)
{
    if (s->n < 0) {
        /* Sigmoid-polynomial spectrum, RGBSigmoidPolynomial::Eval
         * (Utilities/Color/RGBSigmoidPolynomial.h:55-69): amp = c0, c1, c2,
         * scale; zero outside [380, 780] nm. */
        if (nm < 380.0 || nm > 780.0) return 0.0;
        const double t = (nm - 380.0) / 400.0;
        const double x = s->amp[0] * t * t + s->amp[1] * t + s->amp[2];
        return s->amp[3] * (0.5 + x / (2.0 * sqrt(1.0 + x * x)));
    }
    if (nm < s->lambda_begin || nm > s->lambda_end) return 0.0;
    const double delta = (s->lambda_end - s->lambda_begin) / (double)s->n;
    int idx = (int)((nm - s->lambda_begin) / delta);
    if (idx >= s->n) idx = s->n - 1; /* nm == lambda_end */
    return s->amp[idx];
}

/* ---- small vector helpers ------------------------------------------------ */

typedef struct { double x, y, z; } v3;
static v3 mk(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static v3 mul(v3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
static v3 neg(v3 a) { return mk(-a.x, -a.y, -a.z); }
static double dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static v3 cross(v3 a, v3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static double mag(v3 a) { return sqrt(dot(a, a)); }
static v3 normalize(v3 a)
{
    const double m = mag(a);
    if (m > 0.0) {
        const double o = 1.0 / m;
        return mk(a.x * o, a.y * o, a.z * o);
    }
    return a;
}

typedef struct { v3 u, v, w; } onb;

/* OrthonormalBasis3D::CreateFromW (OrthonormalBasis3D.cpp:37-56). */
static onb onb_from_w(v3 n)
{
    onb b;
    b.w = normalize(n);
    const v3 cu = mk(1, 0, 0), cv = mk(0, 1, 0), cw = mk(0, 0, 1);
    if (fabs(dot(b.w, cv) - 1) < NZ) {
        b.u = cw;
        b.v = cu;
    } else if (fabs(dot(b.w, cv) + 1) < NZ) {
        b.u = cu;
        b.v = cw;
    } else {
        b.u = normalize(cross(b.w, cv));
        b.v = normalize(cross(b.w, b.u));
    }
    return b;
}

static v3 onb_xform(const onb* b, v3 a)
{
    return mk(b->u.x * a.x + b->v.x * a.y + b->w.x * a.z, b->u.y * a.x + b->v.y * a.y + b->w.y * a.z,
              b->u.z * a.x + b->v.z * a.y + b->w.z * a.z);
}

/* PathTransportUtilities::PowerHeuristic (PathTransportUtilities.h:251-258). */
static double power_heuristic(double pa, double pb)
{
    const double a2 = pa * pa;
    return a2 / (a2 + pb * pb);
}

/* PathTransportUtilities::EvaluateRussianRoulette (PathTransportUtilities.h:86-110):
 * from depth min_depth on, survive with p = min(1, cur_max / max(prev_max,
 * threshold)); a survivor's weight is divided by `survival` (p when 0 < p < 1,
 * else 1). */
typedef struct { int terminate; double survival; } rr_result;

static rr_result russian_roulette(int depth, int min_depth, double threshold, double cur_max, double prev_max, double u)
{
    rr_result r = {0, 1.0};
    if (depth >= min_depth) {
        const double den = prev_max > threshold ? prev_max : threshold;
        double p = cur_max / den;
        if (p > 1.0) p = 1.0;
        if (u >= p) r.terminate = 1;
        else if (p > 0 && p < 1.0) r.survival = p;
    }
    return r;
}

/* SampledWavelengths::SampleEquidistant (SampledWavelengths.h:56-90): the hero
 * uniform in [a, b), companions at equal offsets with wrap-around. */
static void sample_wavelengths(double u, double a, double b, double lam[HWSS_N])
{
    const double range = b - a, delta = range / HWSS_N;
    lam[0] = a + u * range;
    for (int i = 1; i < HWSS_N; ++i) lam[i] = a + fmod(lam[0] - a + i * delta, range);
}


/* ---- sky (HosekWilkieSkyModel.cpp: Preetham 1999) -------------------------- */

typedef struct { double A, B, C, D, E; } perez;

typedef struct {
    int on;
    double theta_s, cos_theta_s, sun[3], yz, xz, yzc, scale;
    perez pY, px, py;
} sky_model;

static perez perez_of(double T, const double t[5], const double c[5])
{
    perez p = {t[0] * T + c[0], t[1] * T + c[1], t[2] * T + c[2], t[3] * T + c[3], t[4] * T + c[4]};
    return p;
}

static void sky_init(sky_model* m, const ro_scene* s)
{
    memset(m, 0, sizeof(*m));
    if (!s->sky_enabled) return;
    m->on = 1;
    const double el = s->sky_elevation_deg * (M_PI / 180.0), az = s->sky_azimuth_deg * (M_PI / 180.0);
    const double T = s->sky_turbidity;
    /* Preetham 1999 Table 2 (as HosekWilkieSkyModel.cpp:31-53 states them). */
    static const double Yt[5] = {0.1787, -0.3554, -0.0227, 0.1206, -0.0670};
    static const double Yc[5] = {-1.4630, 0.4275, 5.3251, -2.5771, 0.3703};
    static const double xt[5] = {-0.0193, -0.0665, -0.0004, -0.0641, -0.0033};
    static const double xc[5] = {-0.2592, 0.0008, 0.2125, -0.8989, 0.0452};
    static const double yt[5] = {-0.0167, -0.0950, -0.0079, -0.0441, -0.0109};
    static const double yc[5] = {-0.2608, 0.0092, 0.2102, -1.6537, 0.0529};
    double ts = M_PI / 2.0 - el;
    if (ts < 0) ts = 0;
    if (ts > M_PI / 2.0) ts = M_PI / 2.0;
    m->theta_s = ts;
    m->cos_theta_s = cos(ts);
    m->sun[0] = cos(el) * sin(az);
    m->sun[1] = sin(el);
    m->sun[2] = cos(el) * cos(az);
    m->pY = perez_of(T, Yt, Yc);
    m->px = perez_of(T, xt, xc);
    m->py = perez_of(T, yt, yc);
    const double chi = (4.0 / 9.0 - T / 120.0) * (M_PI - 2.0 * ts);
    m->yz = (4.0453 * T - 4.9710) * tan(chi) - 0.2155 * T + 2.4192;
    if (m->yz < 0) m->yz = 0;
    const double t1 = ts, t2 = ts * ts, t3 = t2 * ts;
    m->xz = (0.00166 * t3 - 0.00375 * t2 + 0.00209 * t1 + 0) * T * T +
            (-0.02903 * t3 + 0.06377 * t2 - 0.03202 * t1 + 0.00394) * T +
            (0.11693 * t3 - 0.21196 * t2 + 0.06052 * t1 + 0.25886);
    m->yzc = (0.00275 * t3 - 0.00610 * t2 + 0.00317 * t1 + 0) * T * T +
             (-0.04214 * t3 + 0.08970 * t2 - 0.04153 * t1 + 0.00516) * T +
             (0.15346 * t3 - 0.26756 * t2 + 0.06670 * t1 + 0.26688);
    m->scale = s->sky_scale;
}

static double perez_f(const perez* c, double cos_theta, double gamma)
{
    const double sc = cos_theta > 1e-3 ? cos_theta : 1e-3;
    const double t1 = 1.0 + c->A * exp(c->B / sc);
    const double cg = cos(gamma);
    const double t2 = 1.0 + c->C * exp(c->D * gamma) + c->E * cg * cg;
    return t1 * t2;
}

/* SampleRadiance (HosekWilkieSkyModel.cpp:220-256) times the map's scale. */
static double sky_radiance(const sky_model* m, v3 d, double nm)
{
    if (!m->on) return 0.0;
    if (nm < 380.0 || nm > 780.0) return 0.0;
    d = normalize(d);
    if (d.y <= 0) return 0.0;
    const double ct = d.y > 0 ? d.y : 0;
    double ds = d.x * m->sun[0] + d.y * m->sun[1] + d.z * m->sun[2];
    if (ds < -1) ds = -1;
    if (ds > 1) ds = 1;
    const double g = acos(ds);
    const double Y = m->yz * perez_f(&m->pY, ct, g) / perez_f(&m->pY, 1.0, m->theta_s);
    const double x = m->xz * perez_f(&m->px, ct, g) / perez_f(&m->px, 1.0, m->theta_s);
    const double y = m->yzc * perez_f(&m->py, ct, g) / perez_f(&m->py, 1.0, m->theta_s);
    if (Y < 0) return 0.0;
    if (y < 1e-6) return 0.0;
    const double X = (x / y) * Y, Z = ((1.0 - x - y) / y) * Y;
    double obs[3];
    if (!rise_oracle_xyz_from_nm(nm, obs)) return 0.0;
    const double den = obs[0] + obs[1] + obs[2];
    if (den < 1e-9) return 0.0;
    return (X * obs[0] + Y * obs[1] + Z * obs[2]) / den * m->scale;
}

double rise_oracle_sky_radiance(const ro_scene* scene, const double dir[3], double nm)
{
    sky_model m;
    sky_init(&m, scene);
    return sky_radiance(&m, mk(dir[0], dir[1], dir[2]), nm);
}

/* ---- environment light sample (Rendering/EnvironmentSampler.cpp) ---------- */
/* The importance table RayCaster builds over the sky map's painter
 * (RayCaster.cpp:330-360, 64 x 64 cells; HosekWilkieSpectralRadianceMap.cpp
 * HWAdapterPainter): cells laid out in the light-probe mapping, each weighted
 * by the maximum Rec.709 component of the sky at the cell centre read as
 * equirectangular coordinates, times the probe's sin(theta) / theta. */

#define ENV_RES 64

typedef struct {
    int on;
    double cond[ENV_RES][ENV_RES + 1];
    double marg[ENV_RES + 1];
} env_table;

/* ColorUtils::LineIntersection2D / IsPointOnLineSegment / ClipAgainstGamutEdge /
 * MoveXYZIntoRec709RGBGamut and the XYZ -> Rec.709 matrix (ColorUtils.cpp:150-222,
 * 351-363; Color.cpp:214-220). */
typedef struct { double x, y; } p2;
static int line_isect2(p2 a0, p2 a1, p2 b0, p2 b1, p2* out)
{
    const double den = (b1.y - b0.y) * (a1.x - a0.x) - (b1.x - b0.x) * (a1.y - a0.y);
    if (den > -NZ && den < NZ) return 0;
    const double ua = ((b1.x - b0.x) * (a0.y - b0.y) - (b1.y - b0.y) * (a0.x - b0.x)) / den;
    out->x = a0.x + ua * (a1.x - a0.x);
    out->y = a0.y + ua * (a1.y - a0.y);
    return 1;
}
static int on_seg2(p2 s0, p2 e, p2 q)
{
    const double v = (q.x - s0.x) / (e.x - s0.x);
    return v >= NZ && v <= 1.0 - NZ;
}
static int clip_edge2(p2 c, p2 a, p2 b, p2 w, p2* out)
{
    if (!line_isect2(w, c, a, b, out)) return 0;
    if (!on_seg2(a, b, *out)) return 0;
    if (on_seg2(w, *out, c)) {
        *out = c;
        return 1;
    }
    return on_seg2(w, c, *out);
}
static void xyz_to_rec709(double X, double Y, double Z, double rgb[3])
{
    const p2 red = {0.640, 0.330}, green = {0.300, 0.600}, blue = {0.150, 0.060}, white = {0.3127, 0.3290};
    const double ovs = 1.0 / (X + Y + Z);
    const p2 c = {X * ovs, Y * ovs};
    p2 hit = {0.0, 0.0};
    if (clip_edge2(c, red, green, white, &hit) || clip_edge2(c, red, blue, white, &hit) ||
        clip_edge2(c, blue, green, white, &hit)) {
        const double prod = (1.0 / hit.y) * Y;
        X = prod * hit.x;
        Z = prod * (1 - hit.x - hit.y);
    }
    rgb[0] = 3.240479 * X + -1.537150 * Y + -0.498535 * Z;
    rgb[1] = -0.969256 * X + 1.875992 * Y + 0.041556 * Z;
    rgb[2] = 0.055648 * X + -0.204043 * Y + 1.057311 * Z;
}

/* HosekWilkieSkyModel::IntegrateRGB (HosekWilkieSkyModel.cpp:282-301). */
static void sky_rgb(const sky_model* m, v3 d, double rgb[3])
{
    rgb[0] = rgb[1] = rgb[2] = 0.0;
    if (d.y <= 0) return;
    const double ct = d.y > 0 ? d.y : 0;
    double ds = d.x * m->sun[0] + d.y * m->sun[1] + d.z * m->sun[2];
    if (ds < -1) ds = -1;
    if (ds > 1) ds = 1;
    const double g = acos(ds);
    const double Y = m->yz * perez_f(&m->pY, ct, g) / perez_f(&m->pY, 1.0, m->theta_s);
    const double x = m->xz * perez_f(&m->px, ct, g) / perez_f(&m->px, 1.0, m->theta_s);
    const double y = m->yzc * perez_f(&m->py, ct, g) / perez_f(&m->py, 1.0, m->theta_s);
    if (Y < 0 || y < 1e-6) return;
    xyz_to_rec709((x / y) * Y, Y, ((1.0 - x - y) / y) * Y, rgb);
}

/* EnvironmentSampler::Build (:150-246). */
static void env_build(const sky_model* m, env_table* t)
{
    memset(t, 0, sizeof(*t));
    if (!m->on) return;
    for (int v = 0; v < ENV_RES; ++v) {
        for (int u = 0; u < ENV_RES; ++u) {
            const double s = ((double)u + 0.5) / ENV_RES, tc = ((double)v + 0.5) / ENV_RES;
            const double phi = s * TWO_PI, th = tc * M_PI;
            double rgb[3];
            sky_rgb(m, mk(sin(th) * sin(phi), cos(th), sin(th) * cos(phi)), rgb);
            double lum = rgb[0];
            if (rgb[1] > lum) lum = rgb[1];
            if (rgb[2] > lum) lum = rgb[2];
            const double dx = s - 0.5, dy = -(tc - 0.5), r = sqrt(dx * dx + dy * dy), pt = r * TWO_PI;
            double w = pt < 1e-6 ? 1.0 : sin(pt) / pt;
            if (r > 0.5) w = 0.0;
            t->cond[v][u + 1] = t->cond[v][u] + lum * w;
        }
    }
    for (int v = 0; v < ENV_RES; ++v) t->marg[v + 1] = t->marg[v] + t->cond[v][ENV_RES];
    const double total = t->marg[ENV_RES];
    if (!(total > 0)) return;
    for (int v = 0; v < ENV_RES; ++v) {
        const double row = t->cond[v][ENV_RES];
        if (row > 0)
            for (int u = 0; u <= ENV_RES; ++u) t->cond[v][u] /= row;
        t->cond[v][ENV_RES] = 1.0;
    }
    for (int v = 0; v <= ENV_RES; ++v) t->marg[v] /= total;
    t->marg[ENV_RES] = 1.0;
    t->on = 1;
}

/* EnvironmentSampler::SampleCDF (:118-148). */
static unsigned env_search(const double* cdf, unsigned n, double xi, double* frac)
{
    unsigned lo = 0, hi = n;
    while (lo + 1 < hi) {
        const unsigned mid = (lo + hi) / 2;
        if (cdf[mid] <= xi) lo = mid;
        else hi = mid;
    }
    const double w = cdf[lo + 1] - cdf[lo];
    *frac = w > 0 ? (xi - cdf[lo]) / w : 0.0;
    return lo;
}

static double env_solid_angle(double pdf_st, double s, double tc)
{
    const double dx = s - 0.5, dy = -(tc - 0.5), r = sqrt(dx * dx + dy * dy), th = r * TWO_PI;
    if (th < 1e-6 || r > 0.5) return pdf_st / (4.0 * M_PI);
    const double st = sin(th);
    if (st < 1e-10) return 0.0;
    return pdf_st * th / ((4.0 * M_PI) * M_PI * st);
}

/* EnvironmentSampler::Sample (:248-312), identity map transform. */
static v3 env_sample(const env_table* t, double u1, double u2, double* pdf)
{
    double fv, fu;
    const unsigned v = env_search(t->marg, ENV_RES, u2, &fv);
    const unsigned u = env_search(t->cond[v], ENV_RES, u1, &fu);
    const double s = ((double)u + fu) / ENV_RES, tc = ((double)v + fv) / ENV_RES;
    const double dx = s - 0.5, dy = -(tc - 0.5), rt = sqrt(dx * dx + dy * dy);
    v3 d = mk(0, 0, -1);
    if (!(rt < 1e-12)) {
        const double th = rt * TWO_PI, sth = sin(th), cth = cos(th);
        d = mk(dx * sth / rt, dy * sth / rt, -cth);
    }
    const double pm = (t->marg[v + 1] - t->marg[v]) * ENV_RES;
    const double pc = (t->cond[v][u + 1] - t->cond[v][u]) * ENV_RES;
    *pdf = env_solid_angle(pm * pc, s, tc);
    return d;
}

/* EnvironmentSampler::Pdf (:314-355), identity map transform. */
static double env_pdf(const env_table* t, v3 d)
{
    double s = 0.5, tc = 0.5;
    const double den = sqrt(d.x * d.x + d.y * d.y);
    if (!(den < 1e-12)) {
        double c = -d.z;
        if (c < -1) c = -1;
        if (c > 1) c = 1;
        const double r = 0.159154943 * acos(c) / den;
        s = 0.5 + d.x * r;
        tc = 0.5 - d.y * r;
    }
    s = s < 0 ? 0 : (s > 0.999999 ? 0.999999 : s);
    tc = tc < 0 ? 0 : (tc > 0.999999 ? 0.999999 : tc);
    unsigned u = (unsigned)(s * ENV_RES), v = (unsigned)(tc * ENV_RES);
    if (u > ENV_RES - 1) u = ENV_RES - 1;
    if (v > ENV_RES - 1) v = ENV_RES - 1;
    const double pm = (t->marg[v + 1] - t->marg[v]) * ENV_RES;
    const double pc = (t->cond[v][u + 1] - t->cond[v][u]) * ENV_RES;
    return env_solid_angle(pm * pc, s, tc);
}

/* A directional light (Lights/DirectionalLight.cpp): unit direction toward it
 * and Rec.709 luminance of its colour times its energy. */
typedef struct {
    v3 d;
    double le;
} dir_light;

#define MAX_DIR_LIGHTS 4

/* ---- scene preparation --------------------------------------------------- */

typedef struct {
    int32_t object;
    double area;
    double weight;
    int64_t* tris;   /* triangle indices of this luminary */
    double* cdf;     /* area CDF over those triangles */
    int64_t n;
} luminary;

typedef struct {
    const ro_scene* s;
    double* v;       /* n_tris * 9 doubles */
    luminary* lums;
    int32_t n_lums;
    double lum_total;
    int32_t* lum_of_object; /* -1 if not a luminary */
    double y_norm;
    sky_model sky;
    env_table* env;  /* environment light sample over the sky (NULL: none) */
    dir_light dirs[MAX_DIR_LIGHTS];
    int n_dirs;
    int has_media;   /* some object has an interior medium (bSceneHasObjectMedia) */
    /* work distribution */
    int32_t s0, s1;
    double* xyz_sum;
    double* xyz_sq;
    int64_t next_pixel;
    pthread_mutex_t lock;
    uint64_t n_ext, n_shadow, n_paths;
} prep;

static void tri_verts(const prep* P, int64_t t, v3* a, v3* b, v3* c)
{
    const double* f = P->v + 9 * t;
    *a = mk(f[0], f[1], f[2]);
    *b = mk(f[3], f[4], f[5]);
    *c = mk(f[6], f[7], f[8]);
}

static double tri_area(const prep* P, int64_t t)
{
    v3 a, b, c;
    tri_verts(P, t, &a, &b, &c);
    return mag(cross(sub(b, a), sub(c, a))) * 0.5;
}

static int prepare(prep* P, const ro_scene* s)
{
    memset(P, 0, sizeof(*P));
    P->s = s;
    P->v = (double*)malloc(sizeof(double) * 9 * (size_t)(s->n_tris > 0 ? s->n_tris : 1));
    for (int64_t i = 0; i < 9 * s->n_tris; ++i) P->v[i] = s->tris[i];
    P->lum_of_object = (int32_t*)malloc(sizeof(int32_t) * (size_t)(s->n_objects > 0 ? s->n_objects : 1));
    P->lums = (luminary*)calloc((size_t)(s->n_objects > 0 ? s->n_objects : 1), sizeof(luminary));
    for (int o = 0; o < s->n_objects; ++o) {
        P->lum_of_object[o] = -1;
        const ro_material* m = &s->materials[s->object_material[o]];
        if (m->emit_spd < 0) continue;
        luminary L;
        memset(&L, 0, sizeof(L));
        L.object = o;
        int64_t cnt = 0;
        for (int64_t t = 0; t < s->n_tris; ++t) cnt += s->tri_object[t] == o;
        if (cnt == 0) continue;
        L.tris = (int64_t*)malloc(sizeof(int64_t) * (size_t)cnt);
        L.cdf = (double*)malloc(sizeof(double) * (size_t)cnt);
        /* ComputeAreas (TriangleMeshGeometryIndexed.cpp:404-442). */
        double total = 0;
        for (int64_t t = 0; t < s->n_tris; ++t)
            if (s->tri_object[t] == o) {
                L.tris[L.n] = t;
                const double a = tri_area(P, t);
                L.cdf[L.n] = a;
                total += a;
                L.n++;
            }
        if (!(total > 0)) continue;
        const double inv = 1.0 / total;
        double sum = 0;
        for (int64_t k = 0; k < L.n; ++k) {
            sum += L.cdf[k] * inv;
            L.cdf[k] = sum;
        }
        L.area = total;
        /* Selection weight: mean exitance of the SPD times scale times area
         * (LightSampler::Prepare, LightSampler.cpp:666-686, weights by
         * exitance * area). */
        const ro_spd* e = &s->spds[m->emit_spd];
        double mean = 0;
        if (e->n < 0) {
            /* sigmoid spectrum: mean over the 81 CIE sample wavelengths */
            for (int k = 0; k < 81; ++k) mean += spd_value(e, 380.0 + 5.0 * k);
            mean /= 81.0;
        } else {
            for (int k = 0; k < e->n; ++k) mean += e->amp[k];
            mean /= (double)e->n;
        }
        L.weight = mean * m->emit_scale * total;
        if (!(L.weight > 0)) continue;
        P->lum_of_object[o] = P->n_lums;
        P->lums[P->n_lums++] = L;
        P->lum_total += L.weight;
    }
    const double ky = rise_oracle_cie_y_integral(s->lambda_begin, s->lambda_end);
    P->y_norm = ky > NZ ? (s->lambda_end - s->lambda_begin) / ky : 1.0;
    sky_init(&P->sky, s);
    /* The environment sample's table over the sky (RayCaster.cpp:330-360; a
     * black sky disables it), and the directional lights: the scene's, then the
     * sky's matched sun (Job::AddHosekWilkieSkylight, Job.cpp:2734-2746). */
    if (P->sky.on) {
        P->env = (env_table*)malloc(sizeof(env_table));
        if (!P->env) return -1;
        env_build(&P->sky, P->env);
        if (!P->env->on) {
            free(P->env);
            P->env = NULL;
        }
    }
    for (int k = 0; k < s->n_dir_lights && P->n_dirs < MAX_DIR_LIGHTS; ++k) {
        const double* l = s->dir_lights + 7 * k;
        P->dirs[P->n_dirs].d = normalize(mk(l[0], l[1], l[2]));
        P->dirs[P->n_dirs].le = (0.2126 * l[3] + 0.7152 * l[4] + 0.0722 * l[5]) * l[6];
        ++P->n_dirs;
    }
    if (P->sky.on && s->sun_enabled && P->n_dirs < MAX_DIR_LIGHTS) {
        P->dirs[P->n_dirs].d = normalize(mk(P->sky.sun[0], P->sky.sun[1], P->sky.sun[2]));
        P->dirs[P->n_dirs].le = (0.2126 + 0.7152 + 0.0722) * s->sun_power;
        ++P->n_dirs;
    }
    /* LightSampler::Prepare's MediaScan (LightSampler.cpp:767-785). */
    if (s->object_medium)
        for (int o = 0; o < s->n_objects; ++o) P->has_media |= s->object_medium[o] >= 0;
    return 0;
}

static void release(prep* P)
{
    for (int i = 0; i < P->n_lums; ++i) {
        free(P->lums[i].tris);
        free(P->lums[i].cdf);
    }
    free(P->lums);
    free(P->lum_of_object);
    free(P->v);
    free(P->env);
}

/* ---- ray queries (rt_oracle semantics, inlined for speed) ---------------- */

/* The float pre-filter's record and ray (BVH.h:431-467, :1212-1225): float
 * vertex, float edges, float origin and direction. */
static int filter_tri(const double* f, const float fo[3], const float fd[3], float best)
{
    const float p0[3] = {(float)f[0], (float)f[1], (float)f[2]};
    const float e1[3] = {(float)f[3] - p0[0], (float)f[4] - p0[1], (float)f[5] - p0[2]};
    const float e2[3] = {(float)f[6] - p0[0], (float)f[7] - p0[1], (float)f[8] - p0[2]};
    return rise_oracle_filter_float(fo, fd, p0, e1, e2, best);
}

/* Closest hit for a unit-direction ray; returns tri or -1 and the MT t.
 * Each triangle passes the float pre-filter against the running best
 * (Bvh4Leaf_Geometric, BVH.h:1237-1254) before the double certifier. */
static int64_t closest(const prep* P, v3 o, v3 dn, double* t_out)
{
    double best = RISE_ORACLE_INFINITY;
    int64_t id = -1;
    const double oo[3] = {o.x, o.y, o.z}, dd[3] = {dn.x, dn.y, dn.z};
    const float fo[3] = {(float)o.x, (float)o.y, (float)o.z}, fd[3] = {(float)dn.x, (float)dn.y, (float)dn.z};
    for (int64_t k = 0; k < P->s->n_tris; ++k) {
        const double* f = P->v + 9 * k;
        const double eA[3] = {f[3] - f[0], f[4] - f[1], f[5] - f[2]};
        const double eB[3] = {f[6] - f[0], f[7] - f[1], f[8] - f[2]};
        double t;
        if (!filter_tri(f, fo, fd, (float)best)) continue;
        if (rise_oracle_ray_triangle(oo, dd, f, eA, eB, &t, NULL, NULL) && t < best) {
            best = t;
            id = k;
        }
    }
    *t_out = best;
    return id;
}

/* Any hit in (NEARZERO, dHowFar) after the Object layer's normalization
 * (Object.cpp:652-712), over the objects that cast shadows
 * (ObjectManager.cpp:119-124). */
static int occluded(const prep* P, v3 o, v3 d, double how_far)
{
    const uint8_t* casts = P->s->object_casts_shadows;
    const v3 dn = normalize(d);
    const double m = mag(d);
    const double factor = (m > NZ) ? m : 1.0;
    double how_far2 = how_far;
    if ((how_far != RISE_ORACLE_INFINITY) || (factor < 1.0)) how_far2 = factor * how_far;
    const double oo[3] = {o.x, o.y, o.z}, dd[3] = {dn.x, dn.y, dn.z};
    const float fo[3] = {(float)o.x, (float)o.y, (float)o.z}, fd[3] = {(float)dn.x, (float)dn.y, (float)dn.z};
    for (int64_t k = 0; k < P->s->n_tris; ++k) {
        const double* f = P->v + 9 * k;
        const double eA[3] = {f[3] - f[0], f[4] - f[1], f[5] - f[2]};
        const double eB[3] = {f[6] - f[0], f[7] - f[1], f[8] - f[2]};
        double t;
        if (casts && !casts[P->s->tri_object[k]]) continue;
        if (!filter_tri(f, fo, fd, (float)how_far2)) continue;
        if (rise_oracle_ray_triangle(oo, dd, f, eA, eB, &t, NULL, NULL) && t > NZ && t < how_far2) return 1;
    }
    return 0;
}

/* ---- shading ------------------------------------------------------------- */

typedef struct {
    v3 p;        /* ptIntersection */
    v3 dir;      /* incoming ray direction (ri.ray.Dir()) */
    v3 origin;   /* incoming ray origin */
    v3 n;        /* vNormal (shading) */
    v3 ng;       /* vGeomNormal */
    onb frame;
    int32_t object;
    const ro_material* mat;
} hit;

static void make_hit(const prep* P, v3 o, v3 dn, int64_t tri, double t, hit* h)
{
    const ro_scene* s = P->s;
    v3 a, b, c;
    tri_verts(P, tri, &a, &b, &c);
    const v3 fn = cross(sub(b, a), sub(c, a));
    h->object = s->tri_object[tri];
    h->mat = &s->materials[s->object_material[h->object]];
    h->dir = dn;
    h->origin = o;
    /* Object.cpp:641-643 */
    h->p = add(o, mul(dn, t - 1e-12));
    /* Face normal as both shading and geometric normal
     * (...Specializations.h:194-213), normalized by Object.cpp:543-549. */
    h->n = normalize(fn);
    h->ng = normalize(fn);
    if (s->object_double_sided[h->object]) {
        /* TriangleMeshGeometryIndexed.cpp:171-188 */
        if (dot(h->n, dn) > 0) h->n = neg(h->n);
        if (dot(h->ng, dn) > 0) h->ng = neg(h->ng);
    }
    h->frame = onb_from_w(h->n);
}

/* Geometric normal oriented against the incoming ray (LambertianBRDF.cpp:51-53). */
static v3 geom_against_ray(const hit* h, v3 fallback)
{
    const v3 raw = (dot(h->ng, h->ng) > 1e-12) ? h->ng : fallback;
    return (dot(raw, h->dir) < 0) ? raw : neg(raw);
}

/* ShouldReflect (LambertianBRDF.cpp:44-64). */
static int should_reflect(const hit* h, v3 vin)
{
    const v3 n = h->frame.w;
    const v3 v = normalize(vin);
    const v3 r = normalize(neg(h->dir));
    const double nr = dot(n, r), nv = dot(n, v);
    const v3 g = geom_against_ray(h, n);
    if ((nr <= -NZ && nv <= -NZ) || (nr >= NZ && nv >= NZ)) return dot(v, g) > 0 && dot(r, g) > 0;
    return 0;
}

static double lambert_value(const prep* P, const hit* h, v3 vin, double nm)
{
    if (!should_reflect(h, vin)) return 0.0;
    return spd_value(&P->s->spds[h->mat->reflect_spd], nm) * INV_PI;
}

/* LambertianSPF::Pdf (LambertianSPF.cpp:147-163). */
static double lambert_pdf(const hit* h, v3 wo)
{
    const int front = dot(h->dir, h->frame.w) <= NZ;
    const double c = front ? dot(wo, h->frame.w) : -dot(wo, h->frame.w);
    const v3 neff = front ? h->frame.w : neg(h->frame.w);
    const v3 g = geom_against_ray(h, neff);
    return (c > 0 && dot(wo, g) > 0) ? c * INV_PI : 0.0;
}

/* LambertianEmitter::emittedRadianceNM (LambertianEmitter.cpp:82-89). */
static double emitted(const prep* P, const ro_material* m, v3 out, v3 N, double nm)
{
    if (dot(out, N) <= 0) return 0.0;
    return spd_value(&P->s->spds[m->emit_spd], nm) * INV_PI * m->emit_scale;
}

/* ---- materials --------------------------------------------------------- */

enum { MAT_LAMBERT = 0, MAT_GGX = 1, MAT_DIELECTRIC = 2 };

static double spd_or(const prep* P, int32_t i, double nm, double fallback)
{
    return i >= 0 ? spd_value(&P->s->spds[i], nm) : fallback;
}

static rb_v3 rv(v3 a) { rb_v3 r = {a.x, a.y, a.z}; return r; }
static v3 vr(rb_v3 a) { return mk(a.x, a.y, a.z); }

static rb_ggx ggx_at(const prep* P, const hit* h, double nm, double amb)
{
    const ro_material* m = h->mat;
    rb_ggx g;
    g.dir = rv(h->dir);
    g.onb.u = rv(h->frame.u);
    g.onb.v = rv(h->frame.v);
    g.onb.w = rv(h->frame.w);
    g.ng = rv(h->ng);
    g.rd = spd_or(P, m->reflect_spd, nm, 0.0);
    g.rs = spd_or(P, m->spec_spd, nm, 0.0);
    g.alpha = m->alpha;
    g.ior = spd_or(P, m->ior_spd, nm, 1.0);
    g.ext = spd_or(P, m->ext_spd, nm, 0.0);
    g.amb = amb;
    return g;
}

/* IBSDF::valueNM of the hit's material (the dielectric has none). */
static double bsdf_value(const prep* P, const hit* h, v3 vin, double nm, double amb)
{
    if (h->mat->type == MAT_LAMBERT) return lambert_value(P, h, vin, nm);
    if (h->mat->type == MAT_GGX) {
        const rb_ggx g = ggx_at(P, h, nm, amb);
        return rb_ggx_value(&g, rv(vin));
    }
    return 0.0;
}

/* ISPF::PdfNM. */
static double bsdf_pdf(const prep* P, const hit* h, v3 wo, double nm, double amb)
{
    if (h->mat->type == MAT_LAMBERT) return lambert_pdf(h, wo);
    if (h->mat->type == MAT_GGX) {
        const rb_ggx g = ggx_at(P, h, nm, amb);
        return rb_ggx_pdf(&g, rv(wo));
    }
    return 0.0;
}

/* ISPF::ScatterNM for the non-delta materials: one ray or none.  u1, u2
 * direction, u3 lobe choice. */
static rb_sample bsdf_scatter(const prep* P, const hit* h, double nm, double amb, double u1, double u2, double u3)
{
    rb_sample out = {0, {0, 0, 0}, 0, 0, 0};
    if (h->mat->type == MAT_LAMBERT) {
        /* LambertianSPF::ScatterNM (LambertianSPF.cpp:90-131). */
        onb fr = h->frame;
        const int back = dot(h->dir, h->frame.w) > NZ;
        if (back) {
            fr.w = neg(fr.w);
            fr.u = neg(fr.u);
        }
        const double phi = TWO_PI * u1, st = sqrt(u2), ct = sqrt(1.0 - u2);
        const v3 nd = onb_xform(&fr, mk(cos(phi) * st, sin(phi) * st, ct));
        const v3 neff = back ? neg(h->frame.w) : h->frame.w;
        const v3 g = geom_against_ray(h, neff);
        if (!(dot(nd, g) > 0)) return out;
        out.ok = 1;
        out.dir = rv(nd);
        out.pdf = fabs(dot(nd, h->frame.w)) * INV_PI;
        out.kray = spd_value(&P->s->spds[h->mat->reflect_spd], nm);
        out.type = 0;
        return out;
    }
    if (h->mat->type == MAT_GGX) {
        const rb_ggx g = ggx_at(P, h, nm, amb);
        return rb_ggx_scatter(&g, u3, u1, u2);
    }
    return out;
}

/* ---- index-of-refraction stack (IORStack.h) ------------------------------- */

/* IORStack (Utilities/IORStack.h:95) is a std::stack over a vector: unbounded.
 * A path pushes at most one object per bounce, so air plus MAX_PATH_DEPTH + 1
 * entries never fill.  The HIP renderer holds 3 objects; the pushes it would
 * drop (a 4th nested object) are counted here so tests can compare. */
#define IOR_CAP (MAX_PATH_DEPTH + 2)
#define IOR_DEVICE_OBJECTS 3
static uint64_t g_ior_deep_pushes; /* of the last render (atomic adds) */

typedef struct {
    int n;
    int32_t obj[IOR_CAP];
    double ior[IOR_CAP];
} iorstack;

static void ior_init(iorstack* s)
{
    s->n = 1;
    s->obj[0] = -1;
    s->ior[0] = 1.0;
}

static double ior_top(const iorstack* s) { return s->ior[s->n - 1]; }

static int ior_contains(const iorstack* s, int32_t o)
{
    for (int i = s->n - 1; i >= 0; --i)
        if (s->obj[i] == o) return 1;
    return 0;
}

static void ior_push(iorstack* s, int32_t o, double ior)
{
    if (s->n < IOR_CAP) {
        s->obj[s->n] = o;
        s->ior[s->n] = ior;
        s->n++;
    }
}

static void ior_pop(iorstack* s, int32_t o)
{
    if (s->n <= 1) return;
    for (int i = s->n - 1; i >= 1; --i)
        if (s->obj[i] == o) {
            for (int k = i; k < s->n - 1; ++k) {
                s->obj[k] = s->obj[k + 1];
                s->ior[k] = s->ior[k + 1];
            }
            s->n--;
            return;
        }
}

/* DielectricSPF::ScatterNM + GenerateScatteredRay for a smooth dielectric
 * (DielectricSPF.cpp:120-287, 395-432), then ScatteredRayContainer::
 * RandomlySelect by krayNM (ScatteredRayContainer.cpp:43-92) with the
 * integrator's select probability (PathTracingIntegrator.cpp:2650-2670).
 * Returns 0 when no ray is produced; else the direction, the throughput
 * factor kray / selectProb, the lobe type and the stack the ray carries. */
static int dielectric_scatter(const prep* P, const hit* h, double nm, const iorstack* st, double u_select,
                              v3* dir_out, double* weight, int* type, iorstack* st_out)
{
    const ro_material* m = h->mat;
    const int inside = ior_contains(st, h->object);
    const double rindex = spd_or(P, m->ior_spd, nm, 1.33);
    double k_refr = 1.0;
    if (inside) {
        const double dist = mag(sub(h->p, h->origin));
        k_refr = pow(spd_or(P, m->ext_spd, nm, 1.0), dist);
    }
    const v3 w = h->frame.w;
    const v3 neff = inside ? neg(w) : w;
    const v3 graw = (dot(h->ng, h->ng) > 1e-12) ? h->ng : neff;
    const v3 gn = (dot(graw, neff) >= 0) ? graw : neg(graw);
    rb_v3 refr = rv(h->dir);
    double ref;
    iorstack st_refr = *st;
    if (inside) {
        iorstack exit_st = *st;
        ior_pop(&exit_st, h->object);
        const double exit_ior = ior_top(&exit_st);
        if (rb_refract(rv(neg(w)), rindex, exit_ior, &refr)) {
            ior_pop(&st_refr, h->object);
            ref = rb_dielectric_reflectance(rv(h->dir), refr, rv(neg(w)), rindex, exit_ior);
        } else {
            ref = 1.0;
        }
    } else {
        if (rb_refract(rv(w), ior_top(st), rindex, &refr)) {
            ref = rb_dielectric_reflectance(rv(h->dir), refr, rv(w), ior_top(st), rindex);
            ior_push(&st_refr, h->object, rindex);
        } else {
            ref = 1.0;
        }
    }
    int has_refl = 1, has_refr = 1;
    v3 refl = vr(rb_reflect(rv(h->dir), rv(inside ? w : neg(w))));
    if (dot(refl, gn) <= 0) {
        if (ref >= 1.0) refl = vr(rb_reflect(rv(h->dir), rv(gn)));
        else has_refl = 0;
    }
    const v3 rd = vr(refr);
    if (!inside && dot(rd, w) > -NZ) has_refr = 0;
    else if (inside && dot(rd, w) < NZ) has_refr = 0;
    /* Container order: the refracted ray first, then the reflected one. */
    double kr[2];
    v3 dr[2];
    int ty[2], cnt = 0;
    iorstack sts[2];
    if (has_refr && ref < 1.0) {
        kr[cnt] = k_refr * (1.0 - ref);
        dr[cnt] = rd;
        ty[cnt] = 2;
        sts[cnt] = st_refr;
        cnt++;
    }
    if (has_refl && ref != 0.0) {
        kr[cnt] = ref;
        dr[cnt] = refl;
        ty[cnt] = 1;
        sts[cnt] = *st;
        cnt++;
    }
    if (cnt == 0) return 0;
    int pick = 0;
    double select_prob = 1.0;
    if (cnt == 2) {
        const double total = kr[0] + kr[1];
        if (!(total > NZ)) return 0;
        pick = (u_select < kr[0] / total) ? 0 : 1;
        if (kr[pick] > NZ) select_prob = kr[pick] / total;
    }
    if (select_prob < NZ) return 0;
    *dir_out = dr[pick];
    *weight = kr[pick] / select_prob;
    *type = ty[pick];
    *st_out = sts[pick];
    return 1;
}

/* ---- participating media ------------------------------------------------ */

#define MAX_VOLUME_BOUNCE 64 /* StabilityConfig maxVolumeBounce */

typedef struct {
    double sigma_t, sigma_s;
} medium_coeff;

/* ColorMath::Luminance (ColorMath.h:124-129). */
static double luminance3(const double c[3]) { return 0.2126 * c[0] + 0.7152 * c[1] + 0.0722 * c[2]; }

/* HomogeneousMedium::GetCoefficientsNM without spectral functions
 * (HomogeneousMedium.cpp:95-115): the same at every wavelength. */
static medium_coeff medium_nm(const ro_medium* m)
{
    const double t[3] = {m->sigma_a[0] + m->sigma_s[0], m->sigma_a[1] + m->sigma_s[1], m->sigma_a[2] + m->sigma_s[2]};
    medium_coeff c;
    c.sigma_t = luminance3(t);
    c.sigma_s = luminance3(m->sigma_s);
    return c;
}

/* MediumTracking::GetCurrentMediumWithObject (MediumTracking.h): the top
 * object of the IOR stack if it has an interior medium (no global medium). */
static int32_t medium_object(const prep* P, const iorstack* st)
{
    const int32_t top = st->obj[st->n - 1];
    if (top < 0 || !P->s->object_medium || P->s->object_medium[top] < 0) return -1;
    return top;
}

double rise_oracle_hg_eval(double cos_theta, double g)
{
    /* HenyeyGreensteinPhaseFunction::EvaluateWithG (:20-33). */
    const double g2 = g * g;
    const double denom = 1.0 + g2 - 2.0 * g * cos_theta;
    if (denom <= 0.0) return 0.0;
    return (1.0 - g2) / (4.0 * M_PI * denom * sqrt(denom));
}

/* Vector3Ops::Perpendicular (VectorsOps.h:164-188). */
static v3 perpendicular(v3 v)
{
    const double x = fabs(v.x), y = fabs(v.y), z = fabs(v.z);
    if (x < y && x < z) return mk(0.0, v.z, -v.y);
    if (y < z) return mk(v.z, 0.0, -v.x);
    return mk(v.y, -v.x, 0.0);
}

/* GeometricUtilities::Perturb (GeometricUtilities.cpp:100-118): vec in the
 * basis U = vec, V = perpendicular, W = U x V (OrthonormalBasis3D::
 * CreateFromU), rotated by `down` about the basis Y axis, then by `around`
 * about its X axis (MatricesOps.h:274-303, row-vector transforms), and back. */
static v3 perturb(v3 vec, double down, double around)
{
    const v3 U = normalize(vec), V = normalize(perpendicular(vec)), W = cross(U, V);
    const double a0 = dot(U, vec), a1 = dot(V, vec), a2 = dot(W, vec);
    const double cd = cos(down), sd = sin(down), ca = cos(around), sa = sin(around);
    const double r0 = cd * a0 + sd * a2, r1 = a1, r2 = -sd * a0 + cd * a2;
    const double q0 = r0, q1 = ca * r1 - sa * r2, q2 = sa * r1 + ca * r2;
    return add(add(mul(U, q0), mul(V, q1)), mul(W, q2));
}

void rise_oracle_hg_sample(const double wo[3], double g, double u1, double u2, double out[3])
{
    /* HenyeyGreensteinPhaseFunction::SampleWithG (:35-56). */
    double cos_theta;
    if (fabs(g) < 1e-10) {
        cos_theta = 1.0 - 2.0 * u1;
    } else {
        const double inner = (1.0 - g * g) / (1.0 - g + 2.0 * g * u1);
        cos_theta = (1.0 / (2.0 * g)) * (1.0 + g * g - inner * inner);
    }
    cos_theta = fmax(-1.0, fmin(1.0, cos_theta));
    const v3 d = normalize(perturb(mk(wo[0], wo[1], wo[2]), acos(cos_theta), 2.0 * M_PI * u2));
    out[0] = d.x;
    out[1] = d.y;
    out[2] = d.z;
}

/* EvalShadowTransmittanceNM (LightSampler.cpp:369-470): walks the shadow
 * ray boundary by boundary (closest hits over every object from 1e-5 past
 * the last boundary, at most 16), keeping a stack of the media entered
 * (front face by the geometric normal: push; back face: remove) and
 * multiplying the innermost medium's transmittance over each segment. */
static double shadow_tr(const prep* P, v3 o, v3 d, double max_dist, int32_t origin_obj)
{
    const ro_scene* s = P->s;
    double tr = 1.0;
    if (max_dist <= 0 || !P->has_media) return tr;
    int32_t stk[4];
    int cnt = 0;
    if (origin_obj >= 0) stk[cnt++] = origin_obj;
    double seg = 0.0;
#define TOP_SIGMA_T (medium_nm(&s->media[s->object_medium[stk[cnt - 1]]]).sigma_t)
    for (int step = 0; step < 16 && seg < max_dist; ++step) {
        const double cs = seg + 1e-5;
        if (cs >= max_dist) break;
        const v3 co = add(o, mul(d, cs));
        const double cmax = max_dist - cs;
        double t;
        const int64_t tri = closest(P, co, d, &t);
        if (tri < 0 || t >= cmax) {
            const double rem = max_dist - seg;
            if (rem > 0 && cnt > 0) tr *= exp(-TOP_SIGMA_T * rem);
            seg = max_dist;
            break;
        }
        hit h;
        make_hit(P, co, d, tri, t, &h);
        const double bd = cs + t;
        const double len = bd - seg;
        if (len > 0 && cnt > 0) tr *= exp(-TOP_SIGMA_T * len);
        if (s->object_medium[h.object] >= 0) {
            if (dot(h.ng, d) < 0) {
                if (cnt < 4) stk[cnt++] = h.object;
            } else {
                for (int i = 0; i < cnt; ++i)
                    if (stk[i] == h.object) {
                        for (int j = i; j < cnt - 1; ++j) stk[j] = stk[j + 1];
                        --cnt;
                        break;
                    }
            }
        }
        seg = bd;
        if (tr < 1e-6) return 0.0;
    }
    if (seg < max_dist) {
        const double rem = max_dist - seg;
        if (rem > 0 && cnt > 0) tr *= exp(-TOP_SIGMA_T * rem);
    }
#undef TOP_SIGMA_T
    return tr;
}

/* Luminary selection by CDF inversion over the weights; pdf = w / total. */
static int select_luminary(const prep* P, double u, double* pdf)
{
    double acc = 0;
    for (int i = 0; i < P->n_lums; ++i) {
        acc += P->lums[i].weight;
        if (u * P->lum_total < acc || i == P->n_lums - 1) {
            *pdf = P->lums[i].weight / P->lum_total;
            return i;
        }
    }
    *pdf = 0;
    return -1;
}

static double lum_pdf_select(const prep* P, int lum)
{
    return P->lums[lum].weight / P->lum_total;
}

/* Sampler dimensions of one light sample: the luminary pick and point when
 * the scene has luminaries (an empty alias table draws nothing,
 * LightSampler.cpp:2053-2056), then the environment sample's two. */
static uint32_t light_dims(const prep* P) { return (P->n_lums > 0 ? NEE_DIMS : 0u) + (P->env ? 2u : 0u); }

/* EvaluateDirectLightingNM's luminary sample (LightSampler.cpp:2002-2269);
 * mobj: the object whose medium holds the shading point (-1: none). */
static double lum_nm(prep* P, const hit* h, double nm, double amb, int32_t mobj, const rstream* R, uint32_t dim0,
                     uint64_t* n_shadow)
{
    if (P->n_lums == 0) return 0.0;
    const int self = P->lum_of_object[h->object];
    double pdf_alias;
    const int li = select_luminary(P, u_sampler(R, dim0), &pdf_alias);
    if (li < 0 || li == self) return 0.0;
    const luminary* L = &P->lums[li];
    const double ux = u_sampler(R, dim0 + 1);
    const double uy = u_sampler(R, dim0 + 2);
    const double uz = u_sampler(R, dim0 + 3);
    /* UniformRandomPoint (TriangleMeshGeometryIndexed.cpp:207-218): first
     * CDF entry >= uz, then PointOnTriangle (GeometricUtilities.cpp:543-560). */
    int64_t k = 0;
    while (k < L->n - 1 && L->cdf[k] < uz) ++k;
    v3 a, b, c;
    tri_verts(P, L->tris[k], &a, &b, &c);
    const double sa = sqrt(1.0 - ux);
    const double alpha = 1.0 - sa, beta = sa * uy, gamma = 1.0 - alpha - beta;
    const v3 pl = mk(alpha * b.x + beta * c.x + gamma * a.x, alpha * b.y + beta * c.y + gamma * a.y,
                     alpha * b.z + beta * c.z + gamma * a.z);
    const v3 ln = normalize(cross(sub(b, a), sub(c, a)));
    v3 to = sub(pl, h->p);
    const double dist = mag(to);
    if (dist > 0.0) to = mul(to, 1.0 / dist);
    const double cos_s = dot(to, h->n);
    const double cos_l = dot(neg(to), ln);
    if (cos_l <= 0 || cos_s <= 0) return 0.0;
    ++*n_shadow;
    if (occluded(P, h->p, to, dist - 0.001)) return 0.0;
    const ro_material* lm = &P->s->materials[P->s->object_material[L->object]];
    const double le = emitted(P, lm, neg(to), ln, nm);
    const double geom = L->area * cos_l / (dist * dist);
    double contrib = le * cos_s * geom * bsdf_value(P, h, to, nm, amb);
    contrib *= shadow_tr(P, h->p, to, dist, mobj);
    if (L->area > 0 && cos_l > 0) {
        const double p_light = pdf_alias * (dist * dist) / (L->area * cos_l);
        const double p_bsdf = bsdf_pdf(P, h, to, nm, amb);
        if (p_bsdf > 0) contrib *= power_heuristic(p_light, p_bsdf);
    }
    return contrib / pdf_alias;
}

/* The environment sample (LightSampler.cpp:2278-2358): a direction from the
 * sky's importance table, the sky at nm, power heuristic against the BSDF's
 * pdf, a shadow ray to infinity and the transmittance walk.  Counted and
 * traced only when its value is positive. */
static double env_nm(prep* P, const hit* h, double nm, double amb, int32_t mobj, const rstream* R, uint32_t d0,
                     uint64_t* n_shadow)
{
    if (!P->env) return 0.0;
    double pdf;
    const v3 ed = env_sample(P->env, u_sampler(R, d0), u_sampler(R, d0 + 1), &pdf);
    const double c = dot(ed, h->n);
    if (!(c > 0 && pdf > 0)) return 0.0;
    double contrib = sky_radiance(&P->sky, ed, nm) * bsdf_value(P, h, ed, nm, amb) * c / pdf;
    const double pb = bsdf_pdf(P, h, ed, nm, amb);
    if (pb > 0) contrib *= power_heuristic(pdf, pb);
    if (!(contrib > 0)) return 0.0;
    ++*n_shadow;
    if (occluded(P, h->p, ed, RISE_ORACLE_INFINITY)) return 0.0;
    return contrib * shadow_tr(P, h->p, ed, RISE_ORACLE_INFINITY, mobj);
}

/* The directional lights (LightSampler.cpp:1968-2000 ->
 * DirectionalLight::ComputeDirectLightingNM, DirectionalLight.cpp:75-105):
 * luminance x power x BSDF x cos, a shadow ray to infinity, no transmittance.
 * `count`: count the shadow rays (a shared vertex traces one per light for its
 * four wavelengths). */
static double dirs_nm(prep* P, const hit* h, double nm, double amb, int count, uint64_t* n_shadow)
{
    double sum = 0.0;
    for (int k = 0; k < P->n_dirs; ++k) {
        const dir_light* L = &P->dirs[k];
        const double c = dot(L->d, h->n);
        if (!(c > 0)) continue;
        const double v = L->le * bsdf_value(P, h, L->d, nm, amb) * c;
        if (!(v > 0)) continue;
        if (count) ++*n_shadow;
        if (occluded(P, h->p, L->d, RISE_ORACLE_INFINITY)) continue;
        sum += v;
    }
    return sum;
}

/* The weight of a BSDF-sampled ray that escapes to the sky, against the
 * environment sample (PathTracingIntegrator.cpp:4259-4282): camera rays and
 * rays leaving a delta lobe (pdf 0) keep full weight. */
static double escape_mis(const prep* P, v3 dir, double bsdf_pdf)
{
    if (!P->env || !(bsdf_pdf > 0)) return 1.0;
    const double pe = env_pdf(P->env, dir);
    return pe > 0 ? power_heuristic(bsdf_pdf, pe) : 1.0;
}

/* EvaluateDirectLightingNM at a surface: the directional lights, one luminary
 * sample, one environment sample (their sum; the HIP renderer keeps the
 * three apart until the film). */
static double direct_nm(prep* P, const hit* h, double nm, double amb, int32_t mobj, const rstream* R, uint32_t dim0,
                        int count_dirs, uint64_t* n_shadow)
{
    const double lum = lum_nm(P, h, nm, amb, mobj, R, dim0, n_shadow);
    const double env = env_nm(P, h, nm, amb, mobj, R, dim0 + (P->n_lums > 0 ? NEE_DIMS : 0u), n_shadow);
    const double dirs = dirs_nm(P, h, nm, amb, count_dirs, n_shadow);
    return lum + env + dirs;
}

/* MediumTransport::EvaluateInScatteringNM (MediumTransport.cpp:139-174) ->
 * EvaluateDirectLightingNM with isVolumeScatter: cosSurface = 1, the phase
 * function as BSDF and as MIS pdf, no shading object -- hence no self
 * exclusion and no occlusion test (LightSampler.cpp:2200-2211 tests only for
 * a shading object) -- and the transmittance walk from inside mobj's medium. */
static double lum_volume(prep* P, v3 p, v3 wo, double g, int32_t mobj, double nm, const rstream* R,
                         uint32_t dim0, uint64_t* n_shadow)
{
    if (P->n_lums == 0) return 0.0;
    double pdf_alias;
    const int li = select_luminary(P, u_sampler(R, dim0), &pdf_alias);
    if (li < 0) return 0.0;
    const luminary* L = &P->lums[li];
    const double ux = u_sampler(R, dim0 + 1);
    const double uy = u_sampler(R, dim0 + 2);
    const double uz = u_sampler(R, dim0 + 3);
    int64_t k = 0;
    while (k < L->n - 1 && L->cdf[k] < uz) ++k;
    v3 a, b, c;
    tri_verts(P, L->tris[k], &a, &b, &c);
    const double sa = sqrt(1.0 - ux);
    const double alpha = 1.0 - sa, beta = sa * uy, gamma = 1.0 - alpha - beta;
    const v3 pl = mk(alpha * b.x + beta * c.x + gamma * a.x, alpha * b.y + beta * c.y + gamma * a.y,
                     alpha * b.z + beta * c.z + gamma * a.z);
    const v3 ln = normalize(cross(sub(b, a), sub(c, a)));
    v3 to = sub(pl, p);
    const double dist = mag(to);
    if (dist > 0.0) to = mul(to, 1.0 / dist);
    const double cos_l = dot(neg(to), ln);
    if (cos_l <= 0) return 0.0;
    ++*n_shadow;
    const ro_material* lm = &P->s->materials[P->s->object_material[L->object]];
    const double le = emitted(P, lm, neg(to), ln, nm);
    const double geom = L->area * cos_l / (dist * dist);
    const double phase = rise_oracle_hg_eval(dot(to, wo), g);
    double contrib = le * 1.0 * geom * phase;
    contrib *= shadow_tr(P, p, to, dist, mobj);
    if (L->area > 0 && cos_l > 0) {
        const double p_light = pdf_alias * (dist * dist) / (L->area * cos_l);
        if (phase > 0) contrib *= power_heuristic(p_light, phase);
    }
    return contrib / pdf_alias;
}

/* In-scattering NEE: the luminary sample above, the environment sample
 * (cos 1, the phase function as BSDF and MIS pdf, no occlusion test, the
 * walk), and the directional lights (ComputeDirectLightingNM with the
 * scatter point's normal set to the ray direction, MediumTransport.cpp:
 * 160-175: cos against it, the phase function, occlusion-tested, no walk). */
static double direct_volume(prep* P, v3 p, v3 wo, double g, int32_t mobj, double nm, const rstream* R,
                            uint32_t dim0, uint64_t* n_shadow)
{
    double sum = lum_volume(P, p, wo, g, mobj, nm, R, dim0, n_shadow);
    if (P->env) {
        double pdf;
        const uint32_t d0 = dim0 + (P->n_lums > 0 ? NEE_DIMS : 0u);
        const v3 ed = env_sample(P->env, u_sampler(R, d0), u_sampler(R, d0 + 1), &pdf);
        if (pdf > 0) {
            const double ph = rise_oracle_hg_eval(dot(ed, wo), g);
            double contrib = sky_radiance(&P->sky, ed, nm) * ph / pdf;
            if (ph > 0) contrib *= power_heuristic(pdf, ph);
            if (contrib > 0) {
                ++*n_shadow;
                sum += contrib * shadow_tr(P, p, ed, RISE_ORACLE_INFINITY, mobj);
            }
        }
    }
    for (int k = 0; k < P->n_dirs; ++k) {
        const dir_light* L = &P->dirs[k];
        const double c = dot(L->d, wo);
        if (!(c > 0)) continue;
        const double v = L->le * rise_oracle_hg_eval(dot(L->d, wo), g) * c;
        if (!(v > 0)) continue;
        ++*n_shadow;
        if (!occluded(P, p, L->d, RISE_ORACLE_INFINITY)) sum += v;
    }
    return sum;
}

/* Emission at a hit, weighted against NEE by the power heuristic when the
 * ray came from a non-delta scatter (PathTracingIntegrator.cpp:4422-4515). */
static double emission_mis(const prep* P, const hit* h, double nm, double bsdf_pdf)
{
    if (h->mat->emit_spd < 0) return 0.0;
    double e = emitted(P, h->mat, neg(h->dir), h->ng, nm);
    const int li = P->lum_of_object[h->object];
    if (bsdf_pdf > 0 && li >= 0) {
        const double area = P->lums[li].area;
        const double cos_l = fabs(dot(h->dir, h->ng));
        if (area > 0 && cos_l > 0) {
            const double dist = mag(sub(h->p, h->origin));
            double ps = lum_pdf_select(P, li);
            if (ps <= 0) ps = 1.0;
            const double p_nee = ps * (dist * dist) / (area * cos_l);
            e *= power_heuristic(bsdf_pdf, p_nee);
        }
    }
    return e;
}

/* PropagateBounceLimits (PathTransportUtilities.h:170-210): 1 if the scatter
 * type's cap is exceeded. */
static int over_cap(const ro_scene* s, int type, int* diffuse, int* glossy, int* transmission)
{
    if (type == 0) return ++*diffuse > s->max_diffuse_bounce;
    if (type == 1) return ++*glossy > s->max_glossy_bounce;
    return ++*transmission > s->max_transmission_bounce;
}

/* Key offset of single-wavelength lane w's independent draws (free flight):
 * (w + 1) << 20, so lanes do not share them. */
#define NM_LANE_DIMS (1u << 20)

/* The BSDF's sampler draws at a vertex whose BSDF dimensions start at s0;
 * returns the dimension of its Russian-roulette draw. */
static uint32_t bsdf_draws(const rstream* R, int type, uint32_t s0, double* u1, double* u2, double* u3)
{
    if (type == MAT_GGX) {
        *u3 = u_sampler(R, s0);
        *u1 = u_sampler(R, s0 + 1);
        *u2 = u_sampler(R, s0 + 2);
        return s0 + 4;
    }
    *u1 = u_sampler(R, s0);
    *u2 = u_sampler(R, s0 + 1);
    *u3 = 0.0;
    return s0 + 3;
}

/* IntegrateFromHitNM: the single-wavelength loop a path continues with after
 * a dielectric hit, starting at that hit (PathTracingIntegrator.cpp:1561-3377).
 * Returns the radiance of lane `lane` (wavelength nm) relative to the
 * throughput at the hit. */
static double nm_path(prep* P, const hit* h0, double nm, int lane, int depth0, double importance, double bsdf_pdf,
                      int diffuse, int glossy, int transmission, const rstream* R, uint64_t* n_ext,
                      uint64_t* n_shadow)
{
    int vb = 0; /* volume bounces (none before the branch: the shared loop has no medium) */
    const ro_scene* s = P->s;
    double result = 0.0, thr = 1.0;
    iorstack st;
    ior_init(&st);
    hit h = *h0;
    v3 org = h0->origin, dir = h0->dir;
    for (int depth = depth0; depth < MAX_PATH_DEPTH; ++depth) {
        if (!isfinite(thr) || fabs(thr) > 1e6) break;
        const uint32_t base = STREAM_BASE(depth);
        const uint32_t lane_key = base + NM_LANE_DIMS * (uint32_t)(lane + 1);
        if (depth != depth0) {
            double t;
            ++*n_ext;
            const int64_t tri = closest(P, org, dir, &t);
            /* Medium transport (:1686-1925): free flight at this wavelength
             * (SampleDistanceWithEquiangularMIS_NM without positional lights,
             * :364-386 -> HomogeneousMedium::SampleDistanceNM, :150-172). */
            const int32_t mobj = medium_object(P, &st);
            if (mobj >= 0) {
                const ro_medium* M = &s->media[s->object_medium[mobj]];
                const medium_coeff mc = medium_nm(M);
                const double max_dist = tri >= 0 ? t : RISE_ORACLE_INFINITY;
                int scattered = 0;
                double tm = max_dist;
                if (mc.sigma_t > 0.0) {
                    const double xi = u_indep(R, lane_key + 20);
                    const double tt = -log(fmax(1.0 - xi, 1e-30)) / mc.sigma_t;
                    if (tt < max_dist) {
                        scattered = 1;
                        tm = tt;
                    }
                }
                if (scattered && vb < MAX_VOLUME_BOUNCE) {
                    /* Volume scatter: weight Tr sigma_s / (sigma_t Tr), in-scattering
                     * NEE, phase sampling, Russian roulette at depth + volume
                     * bounces; the ray goes on from the scatter point. */
                    const v3 sp = add(org, mul(dir, tm));
                    const double tr = exp(-mc.sigma_t * tm);
                    double mw = 0.0;
                    if (mc.sigma_t > 0 && tr > 0) mw = tr * mc.sigma_s / (mc.sigma_t * tr);
                    if (!(mw > 0)) break;
                    thr *= mw;
                    const double ld = direct_volume(P, sp, dir, M->g, mobj, nm, R, base, n_shadow);
                    if (ld > 0) result += thr * ld;
                    const double wo[3] = {dir.x, dir.y, dir.z};
                    double wi[3];
                    const uint32_t ls = light_dims(P);
                    rise_oracle_hg_sample(wo, M->g, u_sampler(R, base + ls), u_sampler(R, base + ls + 1), wi);
                    const v3 nd = mk(wi[0], wi[1], wi[2]);
                    const double pp = rise_oracle_hg_eval(dot(dir, nd), M->g);
                    if (pp <= NZ) break;
                    thr *= rise_oracle_hg_eval(dot(dir, nd), M->g) / pp;
                    const rr_result rr = russian_roulette(depth + vb, s->rr_min_depth, s->rr_threshold, fabs(thr),
                                                          importance, u_sampler(R, base + ls + 2));
                    if (rr.terminate) break;
                    if (rr.survival < 1.0) thr /= rr.survival;
                    org = sp;
                    dir = nd;
                    bsdf_pdf = pp;
                    ++vb;
                    continue;
                } else if (!scattered) {
                    /* Survival: Tr / pSurvival, the no-scatter pdf being the
                     * transmittance itself floored at 1e-30 (IMedium.h:235-251). */
                    const double tr = exp(-mc.sigma_t * max_dist);
                    const double ps = fmax(exp(-mc.sigma_t * max_dist), 1e-30);
                    thr *= tr / ps;
                }
            }
            if (tri < 0) {
                /* Escape: the radiance map at this wavelength, weighed against
                 * the environment sample. */
                result += thr * (sky_radiance(&P->sky, dir, nm) * escape_mis(P, dir, bsdf_pdf));
                break;
            }
            make_hit(P, org, dir, tri, t, &h);
        }
        const double amb = ior_top(&st);
        result += thr * emission_mis(P, &h, nm, bsdf_pdf);
        if (h.mat->type == MAT_DIELECTRIC) {
            v3 nd;
            double wgt;
            int type;
            iorstack nst;
            if (!dielectric_scatter(P, &h, nm, &st, u_sampler(R, base + 2), &nd, &wgt, &type, &nst))
                break;
            if (over_cap(s, type, &diffuse, &glossy, &transmission)) break;
            if (nst.n > st.n && nst.n - 1 > IOR_DEVICE_OBJECTS)
                __atomic_add_fetch(&g_ior_deep_pushes, 1, __ATOMIC_RELAXED);
            importance *= fabs(wgt);
            thr *= wgt;
            bsdf_pdf = 0.0;
            st = nst;
            org = add(h.p, mul(nd, 1e-8));
            dir = nd;
            continue;
        }
        result += thr * direct_nm(P, &h, nm, amb, medium_object(P, &st), R, base, 1, n_shadow);
        double u1, u2, u3;
        const uint32_t rr_dim = bsdf_draws(R, h.mat->type, base + light_dims(P), &u1, &u2, &u3);
        const rb_sample sc = bsdf_scatter(P, &h, nm, amb, u1, u2, u3);
        if (!sc.ok) break;
        double scat = sc.kray;
        if (fabs(scat) <= NZ) break;
        const rr_result rr = russian_roulette(depth, s->rr_min_depth, s->rr_threshold, importance * fabs(scat),
                                              importance, u_sampler(R, rr_dim));
        if (rr.terminate) break;
        if (rr.survival < 1.0) scat /= rr.survival;
        importance *= fabs(scat);
        if (over_cap(s, sc.type, &diffuse, &glossy, &transmission)) break;
        thr *= scat;
        bsdf_pdf = sc.pdf;
        const v3 nd = vr(sc.dir);
        org = add(h.p, mul(nd, 1e-8));
        dir = nd;
    }
    return result;
}

/* One camera sample: IntegratePixelSpectral + IntegrateRayHWSS +
 * IntegrateFromHitHWSS.  Returns the sample XYZ. */
static void sample_xyz(prep* P, int32_t px, int32_t py, int32_t si, double xyz_out[3])
{
    const ro_scene* s = P->s;
    const rstream RS = make_stream(s, (uint32_t)px, (uint32_t)py, (uint32_t)si);
    const rstream* R = &RS;
    /* Camera (stream 0): jitter, then hero wavelength. */
    /* Pixel position: SobolSampling2D::GenerateSamplePoints (Sampling/SobolSampling2D.cpp),
     * Owen-scrambled Sobol points of the pixel's samples under one seed per
     * pixel (the reference draws that seed from its Mersenne Twister). */
    const uint32_t pix_seed = rng_u32(s->seed, (uint32_t)((uint64_t)py * s->width + px), 0xffffffffu);
    const double jx = rise_oracle_sobol((uint32_t)si, 0, pix_seed), jy = rise_oracle_sobol((uint32_t)si, 1, pix_seed);
    const double ul = u_indep(R, 2);
    const v3 cp = mk(s->cam_pos[0], s->cam_pos[1], s->cam_pos[2]);
    const v3 f = normalize(sub(mk(s->cam_look[0], s->cam_look[1], s->cam_look[2]), cp));
    const v3 r = normalize(cross(f, mk(s->cam_up[0], s->cam_up[1], s->cam_up[2])));
    const v3 u = cross(r, f);
    const double hh = 2.0 * tan(s->cam_fov_deg * (M_PI / 180.0) * 0.5);
    const double sx = px + jx, sy = py + jy;
    const double cx = (sx - 0.5 * s->width) * hh / s->height;
    const double cy = -(sy - 0.5 * s->height) * hh / s->height;
    v3 dir = normalize(add(add(f, mul(r, cx)), mul(u, cy)));
    v3 org = cp;

    double lam[HWSS_N];
    sample_wavelengths(ul, s->lambda_begin, s->lambda_end, lam);

    double result[HWSS_N] = {0, 0, 0, 0};
    double thr[HWSS_N] = {1, 1, 1, 1};
    double bsdf_pdf = 0.0, importance = 1.0;
    int diffuse = 0, glossy = 0, transmission = 0;
    uint64_t n_ext = 0, n_shadow = 0;

    for (int depth = 0; depth < MAX_PATH_DEPTH; ++depth) {
        /* Runaway guard (PathTracingIntegrator.cpp:4096-4110). */
        double mx = 0;
        int bad = 0;
        for (int w = 0; w < HWSS_N; ++w) {
            if (!isfinite(thr[w])) bad = 1;
            if (fabs(thr[w]) > mx) mx = fabs(thr[w]);
        }
        if (bad || mx > 1e6) break;
        const uint32_t base = STREAM_BASE(depth);

        double t;
        ++n_ext;
        const int64_t tri = closest(P, org, dir, &t);
        if (tri < 0) {
            /* Escape (PathTracingIntegrator.cpp:4230-4290, 5070-5110): the sky at each
             * wavelength; a camera ray sees it only when the map is the background. */
            if (depth > 0 || s->sky_background) {
                const double wmis = escape_mis(P, dir, bsdf_pdf);
                for (int w = 0; w < HWSS_N; ++w) result[w] += thr[w] * (sky_radiance(&P->sky, dir, lam[w]) * wmis);
            }
            break;
        }
        hit h;
        make_hit(P, org, dir, tri, t, &h);

        /* A material without a BSDF (the dielectric): every wavelength goes on
         * as its own single-wavelength path (PathTracingIntegrator.cpp:4348-4365). */
        if (h.mat->type == MAT_DIELECTRIC) {
            for (int w = 0; w < HWSS_N; ++w)
                result[w] += thr[w] * nm_path(P, &h, lam[w], w, depth, importance, bsdf_pdf, diffuse, glossy,
                                              transmission, R, &n_ext, &n_shadow);
            break;
        }

        /* Emission with MIS (PathTracingIntegrator.cpp:4422-4515). */
        for (int w = 0; w < HWSS_N; ++w) result[w] += thr[w] * emission_mis(P, &h, lam[w], bsdf_pdf);
        /* Next-event estimation, one light sample per wavelength
         * (PathTracingIntegrator.cpp:4521-4552). */
        for (int w = 0; w < HWSS_N; ++w)
            result[w] += thr[w] * direct_nm(P, &h, lam[w], 1.0, -1, R, base + light_dims(P) * (uint32_t)w, w == 0,
                                            &n_shadow);

        /* Hero ScatterNM, then companions at the hero's direction (:4620-4718). */
        double u1, u2, u3;
        const uint32_t rr_dim = bsdf_draws(R, h.mat->type, base + HWSS_N * light_dims(P), &u1, &u2, &u3);
        const rb_sample hs = bsdf_scatter(P, &h, lam[0], 1.0, u1, u2, u3);
        if (!hs.ok) break;
        const v3 nd = vr(hs.dir);
        double sc[HWSS_N];
        sc[0] = hs.kray;
        for (int w = 1; w < HWSS_N; ++w) {
            double cw = bsdf_value(P, &h, nd, lam[w], 1.0);
            cw *= fabs(dot(nd, h.n));
            if (hs.pdf > 0) cw /= hs.pdf;
            sc[w] = cw > 0 ? cw : 0;
        }
        /* Russian roulette over the max across wavelengths (:4721-4768). */
        double mprev = fabs(thr[0]), mcurr = fabs(thr[0] * sc[0]);
        for (int w = 1; w < HWSS_N; ++w) {
            if (fabs(thr[w]) > mprev) mprev = fabs(thr[w]);
            if (fabs(thr[w] * sc[w]) > mcurr) mcurr = fabs(thr[w] * sc[w]);
        }
        if (mcurr <= NZ) break;
        const rr_result rr = russian_roulette(depth, s->rr_min_depth, s->rr_threshold, mcurr, mprev,
                                              u_sampler(R, rr_dim));
        if (rr.terminate) break;
        if (rr.survival < 1.0)
            for (int w = 0; w < HWSS_N; ++w) sc[w] *= 1.0 / rr.survival;
        importance *= fabs(sc[0]);
        /* Bounce limits (PathTransportUtilities.h:170-210). */
        if (over_cap(s, hs.type, &diffuse, &glossy, &transmission)) break;
        for (int w = 0; w < HWSS_N; ++w) thr[w] *= sc[w];
        bsdf_pdf = hs.pdf;
        /* currentRay = scattered ray from ptIntersection; Advance(1e-8). */
        org = add(h.p, mul(nd, 1e-8));
        dir = nd;
    }

    double acc[3] = {0, 0, 0};
    for (int w = 0; w < HWSS_N; ++w) {
        double c[3];
        if (rise_oracle_xyz_from_nm(lam[w], c)) {
            acc[0] += c[0] * result[w];
            acc[1] += c[1] * result[w];
            acc[2] += c[2] * result[w];
        }
    }
    const double k = P->y_norm / HWSS_N;
    xyz_out[0] = acc[0] * k;
    xyz_out[1] = acc[1] * k;
    xyz_out[2] = acc[2] * k;
    pthread_mutex_lock(&P->lock);
    P->n_ext += n_ext;
    P->n_shadow += n_shadow;
    P->n_paths += 1;
    pthread_mutex_unlock(&P->lock);
}

static void* render_worker(void* arg)
{
    prep* P = (prep*)arg;
    const ro_scene* s = P->s;
    const int64_t npx = (int64_t)s->width * s->height;
    for (;;) {
        pthread_mutex_lock(&P->lock);
        const int64_t i = P->next_pixel++;
        pthread_mutex_unlock(&P->lock);
        if (i >= npx) break;
        const int32_t px = (int32_t)(i % s->width), py = (int32_t)(i / s->width);
        double sum[3] = {0, 0, 0}, sq[3] = {0, 0, 0};
        for (int32_t si = P->s0; si < P->s1; ++si) {
            double x[3];
            sample_xyz(P, px, py, si, x);
            for (int c = 0; c < 3; ++c) {
                sum[c] += x[c];
                sq[c] += x[c] * x[c];
            }
        }
        for (int c = 0; c < 3; ++c) {
            P->xyz_sum[3 * i + c] = sum[c];
            if (P->xyz_sq) P->xyz_sq[3 * i + c] = sq[c];
        }
    }
    return NULL;
}

uint64_t rise_oracle_last_ior_deep_pushes(void) { return __atomic_load_n(&g_ior_deep_pushes, __ATOMIC_RELAXED); }

int rise_oracle_render(const ro_scene* s, int32_t s0, int32_t s1, double* xyz_sum, double* xyz_sq,
                       uint64_t* counters, int n_threads)
{
    __atomic_store_n(&g_ior_deep_pushes, 0, __ATOMIC_RELAXED);
    if (!s || !xyz_sum || s->width <= 0 || s->height <= 0 || s0 < 0 || s1 < s0 || s1 > s->spp_total) return -1;
    for (int o = 0; o < s->n_objects; ++o) {
        const int32_t mi = s->object_material[o];
        if (mi < 0 || mi >= s->n_materials) return -1;
        const ro_material* m = &s->materials[mi];
        if (m->type < 0 || m->type > 2 || m->reflect_spd >= s->n_spds || m->emit_spd >= s->n_spds ||
            m->spec_spd >= s->n_spds || m->ior_spd >= s->n_spds || m->ext_spd >= s->n_spds)
            return -1;
        if (m->type == 0 && m->reflect_spd < 0) return -1;
        if (s->object_medium && (s->object_medium[o] >= s->n_media || (s->object_medium[o] >= 0 && !s->media)))
            return -1;
    }
    prep P;
    prepare(&P, s);
    P.s0 = s0;
    P.s1 = s1;
    P.xyz_sum = xyz_sum;
    P.xyz_sq = xyz_sq;
    pthread_mutex_init(&P.lock, NULL);
    if (n_threads <= 0) {
        const long c = sysconf(_SC_NPROCESSORS_ONLN);
        n_threads = c > 0 ? (int)c : 1;
    }
    if (n_threads > 256) n_threads = 256;
    pthread_t th[256];
    for (int i = 0; i < n_threads; ++i) pthread_create(&th[i], NULL, render_worker, &P);
    for (int i = 0; i < n_threads; ++i) pthread_join(th[i], NULL);
    if (counters) {
        counters[0] = P.n_ext;
        counters[1] = P.n_shadow;
        counters[2] = P.n_paths;
    }
    pthread_mutex_destroy(&P.lock);
    release(&P);
    return 0;
}

double rise_oracle_shadow_transmittance(const ro_scene* scene, const double o[3], const double d[3], double max_dist,
                                        int32_t origin_object, double nm)
{
    (void)nm; /* RGB media have the same coefficients at every wavelength */
    prep P;
    prepare(&P, scene);
    const double tr = shadow_tr(&P, mk(o[0], o[1], o[2]), mk(d[0], d[1], d[2]), max_dist, origin_object);
    release(&P);
    return tr;
}

/* ---- entry points for the value-for-value comparison with the reference -- */

void rise_oracle_russian_roulette(int depth, int min_depth, double threshold, double cur_max, double prev_max,
                                  double u, double out[2])
{
    const rr_result r = russian_roulette(depth, min_depth, threshold, cur_max, prev_max, u);
    out[0] = r.terminate ? 1.0 : 0.0;
    out[1] = r.survival;
}

double rise_oracle_power_heuristic(double pa, double pb) { return power_heuristic(pa, pb); }

void rise_oracle_sample_wavelengths(double u, double a, double b, double out[4]) { sample_wavelengths(u, a, b, out); }

void rise_oracle_onb_from_w(const double w[3], double out[9])
{
    const onb o = onb_from_w(mk(w[0], w[1], w[2]));
    const v3 k[3] = {o.u, o.v, o.w};
    for (int i = 0; i < 3; ++i) {
        out[3 * i] = k[i].x;
        out[3 * i + 1] = k[i].y;
        out[3 * i + 2] = k[i].z;
    }
}
