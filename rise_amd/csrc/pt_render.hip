// pt_render.hip -- wavefront spectral path tracer: renders of one request
// (rise_hip_render) or of many (rise_hip_batch_*) on one stream.
//
// Work is cut into chunks (a block of pixels x a block of samples of one
// request) whose layout depends only on that request, and chunks are packed
// into passes holding at most one chunk per request.  Per pass, kernels run
// in this order:
//   generate            camera rays, hero wavelength, path state
//   for each bounce d:  extend(d)   closest hit for the active queue
//                       shade(d)    emission + MIS, one NEE sample per
//                                   wavelength (shadow queue), hero BSDF
//                                   sample, companions, RR, bounce caps
//                                   (continuation queue)
//                       connect(d)  any-hit for the shadow queue; adds the
//                                   unoccluded contributions
//   film                per-pixel FP32 Kahan sums over the chunk's samples
// and one resolve per request at the end.  Every path carries its request
// index, so paths of different scenes and image sizes share the queues.
// Queues are compacted with wave ballots and one atomic per workgroup; every
// count lives in device memory and every kernel is a grid-stride loop over
// it, so a batch is a fixed sequence of kernel launches with no host
// synchronization (capturable in a hipGraph).  A pixel's samples are summed
// in sample order whatever the chunking, so a request's film is bit-identical
// whether it is rendered alone or in a batch.
//
// The bounce loop follows PathTracingIntegrator::IntegrateFromHitHWSS
// (Shaders/PathTracingIntegrator.cpp:3939-4822) for Lambertian materials and
// Lambertian area luminaries; each step cites the lines it follows.
#include "rise_hip_render.h"

#include "api_common.h"
#include "cie_table.h"
#include "pt_ggx.h"
#include "env_sampler.h"
#include "pt_internal.h"
#include "wave_fetch.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <chrono>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <tuple>
#include <string>
#include <vector>

using namespace rise_hip;

namespace {

__constant__ float c_cie_x[81] = {RISE_CIE_X_VALUES};
__constant__ float c_cie_y[81] = {RISE_CIE_Y_VALUES};
__constant__ float c_cie_z[81] = {RISE_CIE_Z_VALUES};
const double h_cie_y[81] = {RISE_CIE_Y_VALUES};

constexpr int kTraceBlock = 128;
constexpr int kShadeBlock = 512;  // shading / binning workgroups: 8 waves share one queue atomic
#ifndef RISE_SHADE_GRID
#define RISE_SHADE_GRID 2048      // workgroups of the grid-stride shading / binning launches (at most)
#endif
constexpr int64_t kChunkPaths = int64_t(1) << 20;   // paths per chunk (at most)
#ifndef RISE_MAX_PATHS_LOG2
#define RISE_MAX_PATHS_LOG2 25
#endif
constexpr int64_t kMaxPaths = int64_t(1) << RISE_MAX_PATHS_LOG2;  // paths in flight per pass (at most)
// Per pass: extension rays, shadow rays and 4 binned-hit counts per depth,
// then the index-of-refraction pushes dropped for lack of room.
constexpr int kCounters = 6 * (kMaxPathDepth + 1) + 1;
constexpr int kOverflowCounter = 6 * (kMaxPathDepth + 1);
constexpr int kBins = 4;                            // hit bins: the 3 material kinds, then volume scatters
constexpr int kMatVolume = 3;                       // bin of the lanes that scatter inside a medium
constexpr int kMaxVolumeBounce = 64;                // StabilityConfig maxVolumeBounce (StabilityConfig.h:112)

// ColorUtils::CIE_Y_Integral (Utilities/Color/ColorUtils.cpp:327-350).
double cie_y_integral(double a, double b)
{
    const double lo = a < 380.0 ? 380.0 : a, hi = b > 780.0 ? 780.0 : b;
    if (hi <= lo) return 0.0;
    const int first = (int)((lo - 380.0) / 5.0), last = (int)((hi - 380.0) / 5.0);
    double s = 0.0;
    for (int i = first; i < last; ++i) s += 0.5 * (h_cie_y[i] + h_cie_y[i + 1]) * 5.0;
    return s;
}

// XYZFromNM (ColorUtils.cpp:270-325): linear interpolation of the 5 nm table.
__device__ inline bool xyz_from_nm(float nm, float& X, float& Y, float& Z)
{
    if (nm < 380.0f || nm > 780.0f) return false;
    const float x = (nm - 380.0f) / 5.0f;
    const int lo = (int)floorf(x);
    const int hi = lo + 1 > 80 ? 80 : lo + 1;
    const float w = x - (float)lo;
    X = c_cie_x[lo] + (c_cie_x[hi] - c_cie_x[lo]) * w;
    Y = c_cie_y[lo] + (c_cie_y[hi] - c_cie_y[lo]) * w;
    Z = c_cie_z[lo] + (c_cie_z[hi] - c_cie_z[lo]) * w;
    return true;
}

// HosekWilkieSkyModel::SampleRadiance (Utilities/HosekWilkieSkyModel.cpp:
// 172-256) times the map's scale: Perez-weighted xyY in direction d (unit),
// turned into a radiance at nm by the colour-matching functions.
__device__ inline float sky_radiance(const DevSky& k, f3 d, float nm)
{
    if (!k.on || nm < 380.0f || nm > 780.0f) return 0.0f;
    d = norm3(d);
    if (d.y <= 0.0f) return 0.0f;
    const float ds = fmaxf(-1.0f, fminf(1.0f, d.x * k.sun[0] + d.y * k.sun[1] + d.z * k.sun[2]));
    const float g = acosf(ds), cg = cosf(g);
    const float sc = fmaxf(1e-3f, d.y);
    float v[3];
    for (int c = 0; c < 3; ++c) {
        const float* p = k.pz[c];
        v[c] = k.zen[c] * (1.0f + p[0] * expf(p[1] / sc)) * (1.0f + p[2] * expf(p[3] * g) + p[4] * cg * cg);
    }
    const float Y = v[0], x = v[1], y = v[2];
    if (Y < 0.0f || y < 1e-6f) return 0.0f;
    const float X = (x / y) * Y, Z = ((1.0f - x - y) / y) * Y;
    float ox, oy, oz;
    if (!xyz_from_nm(nm, ox, oy, oz)) return 0.0f;
    const float den = ox + oy + oz;
    if (den < 1e-9f) return 0.0f;
    return (X * ox + Y * oy + Z * oz) / den * k.scale;
}

// One request: scene tables, render parameters, film accumulators, output.
struct ReqInfo {
    PtSceneView sc;
    int width, height, spp, spp_total;
    int max_diffuse, max_glossy, max_transmission, rr_min_depth;
    float rr_threshold, lambda_begin, lambda_end;
    float xyz_scale;  // mYNormalization / N (PathTracingSpectralRasterizer.cpp:168-173)
    int sky_background;  // camera rays that miss see the sky
    unsigned long long seed;
    // Sampler (sample_key): Z-order index, log2 of spp rounded up to a power
    // of two, the shuffle mask (0: unshuffled), the render seed folded to 32 bits.
    int sob_zorder;
    unsigned sob_l2, sob_mask, seed32;
    float cam_pos[3], cam_f[3], cam_r[3], cam_u[3];
    float cam_h;      // 2 tan(fov / 2)
    float *fsum, *fcomp, *fsq, *fsqc;  // 3 * width * height each
    float *out, *out_sq;
};

// One chunk of a pass: pixels [p0, p0 + np) x samples [s0, s0 + ns) of
// request `req`, in pool slots [slot0, slot0 + np * ns) (pixel-major), and
// film threads [film0, film0 + nf).  A pass may hold several chunks of one
// request (contiguous, in chunk order); a chunk whose pixel block an earlier
// chunk of the pass already covers has nf = 0, and that chunk's film
// threads add both, in sample order.
struct PassEntry {
    int req, p0, np, s0, ns, slot0, film0, nf;
};

// Path state, structure of arrays.  A path is one slot; its rays are kept per
// lane (4 * cap entries, index lane * cap + slot): a path whose wavelengths
// share one ray uses lane 0, and after a dielectric hit lane w carries
// wavelength w on its own.
struct Paths {
    int cap;
    float *ox, *oy, *oz, *dx, *dy, *dz;  // 4 * cap
    float* thr;      // 4 * cap, per wavelength; a single-wavelength lane's is relative to its branch
    float* thr0;     // 4 * cap: a single-wavelength lane's throughput at the dielectric hit it branched at
    float* res;      // 4 * cap, per wavelength
    float* lambda0;  // cap
    float* pdf;      // 4 * cap: bsdfPdf of the ray being traced (0: camera or delta)
    float* imp;      // 4 * cap: importance (product of scatter weights)
    unsigned* state; // 4 * cap: bounce counts (bits 0-7 diffuse, 8-15 glossy, 16-23 transmission), IOR-stack depth (24-25)
    int* ior_obj;    // 3 * 4 * cap: objects on the IOR stack above the ambient air
    unsigned* flags; // cap: kModeNm once the wavelengths go on as separate lanes
    unsigned* path_id;
    unsigned* req;
    unsigned* hit_leaf;  // 4 * cap
    float* hit_t;        // 4 * cap
    unsigned* queue[2];  // 4 * cap entries, slot * 4 + lane
    unsigned* mqueue[kBins]; // hits binned by material kind (and volume scatters), 4 * cap entries each
    // shadow records, 4 * cap: origin, unit direction, distance to the light point, value
    float *sox, *soy, *soz, *sdx, *sdy, *sdz, *stmax, *sval;
    unsigned* sslot;     // record code: slot * 4 + wavelength, kind, all-lanes flag (kRec*)
    // Result arrays of the environment sample and of each directional light
    // (4 * cap each; null when no scene of the pass has them), and a shared
    // ray's four directional-light values awaiting their one shadow ray.
    float* res_env;
    float* res_dir[kMaxDirLights];
    float* dir_val[kMaxDirLights];
    // Media only (null when no scene of the pass holds a medium):
    unsigned char* vb;   // 4 * cap: volume bounces of the lane
    unsigned* sinfo;     // per shadow record: kShadowInScatter | (object whose medium holds the origin) + 1
    unsigned* svis;      // per shadow record: 1 once found unoccluded, awaiting the transmittance walk
    unsigned* cnt_ext;   // [kMaxPathDepth + 1] of the current pass
    unsigned* cnt_sh;    // [kMaxPathDepth + 1]
    unsigned* cnt_mat;   // [kBins][kMaxPathDepth + 1]: binned hits per material kind, then volume scatters
    unsigned* cnt_ovf;   // index-of-refraction pushes dropped (stack full)
    unsigned long long* fetch;  // [2 * (kMaxPathDepth + 1)]: work counters of the persistent traversal kernels
};

// Workgroup-wide compaction: a ballot per wave and one atomic per workgroup,
// so a queue counter sees one atomic per BLOCK threads (same-address atomics
// serialize in L2, and one per wave made them the shading kernels' bottleneck).  Every thread of the block calls it
// at the same point: the grid-stride loops run one trip count per block.
// Consecutive calls alternate between two LDS slots, so each needs two
// barriers.  Positions within a queue carry no meaning (each path's result is
// summed per slot and per pixel in sample order), so the order of
// reservations does not change any film.
template <int BLOCK, int NB>
struct BlockPush {
    static constexpr int W = BLOCK / 64;
    unsigned* lds;  // 2 * NB * (W + 1) entries
    int flip;
    // Reserves one entry in queue `kind` (0..NB-1; -1: none) for this thread;
    // counters[k] counts queue k.  Returns the index, or 0xffffffff.
    __device__ inline unsigned operator()(unsigned* const* counters, int kind)
    {
        unsigned* b = lds + flip * NB * (W + 1);
        flip ^= 1;
        const unsigned wave = threadIdx.x >> 6;
        unsigned rank = 0;
        for (int k = 0; k < NB; ++k) {
            const unsigned long long m = __ballot(kind == k);
            if (kind == k)
                rank = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
            if (__lane_id() == 0) b[k * (W + 1) + wave] = (unsigned)__popcll(m);
        }
        __syncthreads();
        if (threadIdx.x < NB) {
            unsigned* c = b + threadIdx.x * (W + 1);
            unsigned total = 0;
            for (int w = 0; w < W; ++w) {
                const unsigned n = c[w];
                c[w] = total;
                total += n;
            }
            c[W] = total ? atomicAdd(counters[threadIdx.x], total) : 0u;
        }
        __syncthreads();
        return kind >= 0 ? b[kind * (W + 1) + W] + b[kind * (W + 1) + wave] + rank : 0xffffffffu;
    }
    __device__ inline unsigned operator()(unsigned* counter, bool pred)
    {
        unsigned* const cs[1] = {counter};
        return (*this)(cs, pred ? 0 : -1);
    }
};

// Index of the last entry whose field KEY is <= x (entries sorted by KEY,
// the first one 0).
template <int PassEntry::*KEY>
__device__ inline int find_entry(const PassEntry* __restrict__ ent, int n, int x)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (ent[mid].*KEY <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// The sampler key of a path (pixel index * spp_total + sample):
// PathTracingSpectralRasterizer::IntegratePixel (:229-345) with the Z-order
// (ZSobolSampler) or plain (SobolSampler) index and per-pixel seed; the render
// seed, when not 0, folded into the pixel seed with HashCombine.  The oracle's
// make_stream (oracle/pt_oracle.c) computes the same key.
__device__ inline SampleKey sample_key(const ReqInfo& R, unsigned path)
{
    const unsigned st = (unsigned)R.spp_total, w = (unsigned)R.width;
    const unsigned pix = path / st, si = path - pix * st;
    const unsigned py = pix / w, px = pix - py * w;
    unsigned pseed = hash_combine(px, py);
    SampleKey k{si, 0u, R.sob_mask};
    if (R.sob_zorder && px <= 0xffffu && py <= 0xffffu) {
        const unsigned mi = morton_part1by1(px) | (morton_part1by1(py) << 1);
        if (R.sob_l2 < 32u && ((unsigned long long)mi << R.sob_l2) < (1ull << 32)) {
            k.index = (mi << R.sob_l2) | si;
            pseed = hash_combine(mi, 0u);
        }
    }
    k.scramble = R.seed32 ? hash_combine(pseed, R.seed32) : pseed;
    return k;
}

__global__ void k_fill(float* p, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0f;
}

__global__ void k_zero_u32(unsigned* p, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

// Camera samples: PathTracingSpectralRasterizer::IntegratePixel (box filter,
// :310-338) and IntegratePixelSpectral (:121-150).
__global__ void k_generate(const ReqInfo* __restrict__ reqs, const PassEntry* __restrict__ ent, int n_ent,
                           int n_paths, Paths S)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) S.cnt_ext[0] = (unsigned)n_paths;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n_paths; i += gridDim.x * blockDim.x) {
        const PassEntry E = ent[find_entry<&PassEntry::slot0>(ent, n_ent, i)];
        const ReqInfo& R = reqs[E.req];
        const int local = i - E.slot0;
        const int pix = E.p0 + local / E.ns;
        const int s = E.s0 + local % E.ns;
        const int px = pix % R.width, py = pix / R.width;
        const unsigned path = (unsigned)((unsigned long long)pix * (unsigned long long)R.spp_total + (unsigned)s);
        // Pixel position: SobolSampling2D::GenerateSamplePoints (Sampling/SobolSampling2D.cpp),
        // Owen-scrambled Sobol points of the pixel's samples under one seed per pixel.
        const unsigned pix_seed = rng_u32(R.seed, (unsigned)pix, 0xffffffffu);
        const float jx = sobol_f((unsigned)s, 0u, pix_seed), jy = sobol_f((unsigned)s, 1u, pix_seed);
        const float ul = rng_u(R.seed, path, 2);
        const float sx = (float)px + jx, sy = (float)py + jy;
        const float cx = (sx - 0.5f * (float)R.width) * R.cam_h / (float)R.height;
        const float cy = -(sy - 0.5f * (float)R.height) * R.cam_h / (float)R.height;
        const f3 f = mk3(R.cam_f[0], R.cam_f[1], R.cam_f[2]);
        const f3 r = mk3(R.cam_r[0], R.cam_r[1], R.cam_r[2]);
        const f3 u = mk3(R.cam_u[0], R.cam_u[1], R.cam_u[2]);
        const f3 d = norm3(f + r * cx + u * cy);
        S.ox[i] = R.cam_pos[0];
        S.oy[i] = R.cam_pos[1];
        S.oz[i] = R.cam_pos[2];
        S.dx[i] = d.x;
        S.dy[i] = d.y;
        S.dz[i] = d.z;
        for (int w = 0; w < kHwss; ++w) {
            S.thr[w * S.cap + i] = 1.0f;
            S.res[w * S.cap + i] = 0.0f;
            if (S.res_env) S.res_env[w * S.cap + i] = 0.0f;
            for (int l = 0; l < kMaxDirLights; ++l)
                if (S.res_dir[l]) S.res_dir[l][w * S.cap + i] = 0.0f;
        }
        S.lambda0[i] = R.lambda_begin + ul * (R.lambda_end - R.lambda_begin);
        S.pdf[i] = 0.0f;
        S.imp[i] = 1.0f;
        S.state[i] = 0u;
        S.flags[i] = 0u;
        S.path_id[i] = path;
        S.req[i] = (unsigned)E.req;
        if (S.vb) S.vb[i] = 0;
        S.queue[0][i] = (unsigned)i * 4u;
    }
}

// Closest hit for the active queue (ObjectManager::IntersectRay).
// PERSIST: persistent waves take queue entries with a wave-wide fetch
// (wave_fetch.h) and step one traversal iteration at a time, so lanes that
// finish early are refilled; otherwise one thread traces one ray to the end
// (grid-stride), which costs less per ray when trees are shallow.
template <int STACK, int BLOCK, bool PERSIST>
__global__ __launch_bounds__(BLOCK) void k_extend(const ReqInfo* __restrict__ reqs, Paths S, int depth)
{
    __shared__ uint2 lds[STACK > 0 ? STACK * BLOCK : 1];
    LdsStack<STACK> lst{lds + threadIdx.x, BLOCK, 0};
    PrivStack pst;
    pst.sp = 0;
    const unsigned n = S.cnt_ext[depth];
    const unsigned* q = S.queue[depth & 1];
    if constexpr (!PERSIST) {
        for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
            const unsigned code = q[i];
            const unsigned slot = (code & 3u) * (unsigned)S.cap + (code >> 2);
            const ReqInfo& R = reqs[S.req[code >> 2]];
            RayState r;
            init_ray(r, S.ox[slot], S.oy[slot], S.oz[slot], S.dx[slot], S.dy[slot], S.dz[slot]);
            double best;
            unsigned leaf, nn = 0, nt = 0;
            lst.sp = 0;
            pst.sp = 0;
            const int id = STACK > 0 ? trace_closest<false>(r, R.sc.nodes, R.sc.tris, lst, best, leaf, nn, nt)
                                     : trace_closest<false>(r, R.sc.nodes, R.sc.tris, pst, best, leaf, nn, nt);
            S.hit_leaf[slot] = id >= 0 ? leaf : 0xffffffffu;
            S.hit_t[slot] = id >= 0 ? (float)best : 0.0f;
        }
        return;
    }
    unsigned long long* next = S.fetch + depth;
    RayState r;
    ClosestState s;
    const NodeSlot* nodes = nullptr;
    const Tri48* tris = nullptr;
    unsigned slot = 0, nn = 0, nt = 0;  // slot: lane * cap + path slot of the current item
    long long item = -1;  // -1 idle, -2 exhausted
    const int64_t grid_threads = (int64_t)gridDim.x * blockDim.x;
    for (bool first_pass = true;; first_pass = false) {
        // The first item of every thread is assigned statically; later ones come from the counter.
        const long long got = first_pass ? first_item(n) : refill(item == -1, item >= 0, next, n, grid_threads);
        if (item == -1 && got != -1) {
            item = got;
            if (item >= 0) {
                const unsigned code = q[item];
                slot = (code & 3u) * (unsigned)S.cap + (code >> 2);
                const ReqInfo& R = reqs[S.req[code >> 2]];
                nodes = R.sc.nodes;
                tris = R.sc.tris;
                init_ray(r, S.ox[slot], S.oy[slot], S.oz[slot], S.dx[slot], S.dy[slot], S.dz[slot]);
                closest_begin(r, s);
                lst.sp = 0;
                pst.sp = 0;
            }
        }
        if (__ballot(item >= 0) == 0ull && __ballot(item == -1) == 0ull) break;
        if (item >= 0) {
            const bool more = STACK > 0 ? closest_step<false>(r, nodes, tris, lst, s, nn, nt)
                                        : closest_step<false>(r, nodes, tris, pst, s, nn, nt);
            if (!more) {
                S.hit_leaf[slot] = s.best_id >= 0 ? s.best_leaf : 0xffffffffu;
                S.hit_t[slot] = s.best_id >= 0 ? (float)s.best_t : 0.0f;
                item = -1;
            }
        }
    }
}

// ---- materials at one wavelength ------------------------------------------------

constexpr GgxTables make_ggx_tables()
{
    GgxTables t{};
    constexpr double eavg[kLut] = {RISE_GGX_EAVG_VALUES};
    constexpr double glx[21] = {RISE_GL21_NODES};
    constexpr double glw[21] = {RISE_GL21_WEIGHTS};
    const std::array<float, kLut> z = make_row_z();
    for (int i = 0; i < kLut * kLut; ++i) t.ess[i] = (float)kEssTable[i];
    for (int i = 0; i < kLut; ++i) {
        t.eavg[i] = (float)eavg[i];
        t.row_z[i] = z[i];
    }
    for (int i = 0; i < 21; ++i) {
        t.gl_x[i] = (float)glx[i];
        t.gl_w[i] = (float)glw[i];
    }
    return t;
}

__constant__ GgxTables c_ggx = make_ggx_tables();

// Painter defaults of the scene-file chunks when a spectrum is absent:
// ggx_material ior 2.45, extinction 3.45; dielectric_material ior 1.33, tau
// none (ChunkParserRegistry.cpp:3206-3207, 3835-3836).
__device__ inline float spd_or(const PtSceneView& sc, int spd, float nm, float fallback)
{
    return spd >= 0 ? spd_value(sc, spd, nm) : fallback;
}

__device__ inline GgxParams ggx_params(const PtSceneView& sc, const DevMaterial& m, float nm, float amb)
{
    return GgxParams{spd_or(sc, m.refl_spd, nm, 0.0f), spd_or(sc, m.spec_spd, nm, 0.0f), m.alpha,
                     spd_or(sc, m.ior_spd, nm, 2.45f), spd_or(sc, m.ext_spd, nm, 3.45f), amb};
}

// IBSDF::valueNM of a non-delta material of kind MAT (the dielectric has no BSDF).
template <int MAT>
__device__ inline float bsdf_value(const PtSceneView& sc, const DevMaterial& m, const Surf& h, f3 v, float nm,
                                   float amb)
{
    if (MAT == kMatLambert) return lambert_value(sc, h, m.refl_spd, v, nm);
    if (MAT == kMatGgx) return ggx_value(c_ggx, h, ggx_params(sc, m, nm, amb), v);
    return 0.0f;
}

// ISPF::PdfNM.
template <int MAT>
__device__ inline float bsdf_pdf(const PtSceneView& sc, const DevMaterial& m, const Surf& h, f3 wo, float nm,
                                 float amb)
{
    if (MAT == kMatLambert) return lambert_pdf(h, wo);
    if (MAT == kMatGgx) return ggx_pdf(c_ggx, h, ggx_params(sc, m, nm, amb), wo);
    return 0.0f;
}

// ISPF::ScatterNM of a non-delta material: one ray or none.
template <int MAT>
__device__ inline ScatterOut bsdf_scatter(const PtSceneView& sc, const DevMaterial& m, const Surf& h, float nm,
                                          float amb, float u1, float u2, float u3)
{
    if (MAT == kMatGgx) return ggx_scatter(c_ggx, h, ggx_params(sc, m, nm, amb), u3, u1, u2);
    // LambertianSPF::ScatterNM (Materials/LambertianSPF.cpp:90-131).
    ScatterOut out{false, mk3(0, 0, 0), 0.0f, 0.0f, 0};
    Onb fr = h.frame;
    const bool back = dot3(h.dir, h.frame.w) > kNearZeroF;
    if (back) {
        fr.w = -fr.w;
        fr.u = -fr.u;
    }
    const float phi = kTwoPi * u1, st = sqrtf(u2), ct = sqrtf(1.0f - u2);
    const f3 nd = onb_xform(fr, mk3(cosf(phi) * st, sinf(phi) * st, ct));
    const f3 g = geom_against_ray(h, back ? -h.frame.w : h.frame.w);
    if (dot3(nd, g) > 0.0f)
        out = ScatterOut{true, nd, spd_value(sc, m.refl_spd, nm), fabsf(dot3(nd, h.frame.w)) * kInvPi, 0};
    return out;
}

// The BSDF's sampler draws at a vertex whose BSDF dimensions start at s0, in
// the reference's order (GGX: the lobe, then two, GGXSPF.cpp:479-519;
// Lambertian: two, LambertianSPF.cpp:90-99), after them the lobe choice;
// returns the dimension of the Russian-roulette draw.
template <int MAT>
__device__ inline unsigned bsdf_draws(const SampleKey& k, unsigned s0, float& u1, float& u2, float& u3)
{
    if (MAT == kMatGgx) {
        u3 = sample_u(k, s0);
        u1 = sample_u(k, s0 + 1);
        u2 = sample_u(k, s0 + 2);
        return s0 + 4;
    }
    u1 = sample_u(k, s0);
    u2 = sample_u(k, s0 + 1);
    u3 = 0.0f;
    return s0 + 3;
}

// Index-of-refraction stack of one single-wavelength path (IORStack.h):
// objects above the ambient air; the index of an entry is its dielectric's
// index at the path's wavelength.  At most 3 nested objects (the reference's
// stack is unbounded, IORStack.h:95): a deeper push is dropped and counted,
// and the count is reported as rise_hip_render_stats::ior_overflows.
struct IorStack {
    int n;
    int obj[3];
};

__device__ inline float ior_of(const PtSceneView& sc, int obj, float nm)
{
    return spd_or(sc, sc.materials[sc.objects[obj].material].ior_spd, nm, 1.33f);
}

__device__ inline float ior_top(const PtSceneView& sc, const IorStack& s, float nm)
{
    return s.n > 0 ? ior_of(sc, s.obj[s.n - 1], nm) : 1.0f;
}

__device__ inline bool ior_contains(const IorStack& s, int o)
{
    bool c = false;
    for (int i = 0; i < 3; ++i) c = c || (i < s.n && s.obj[i] == o);
    return c;
}

// Returns false when the stack is full and the push is dropped.
__device__ inline bool ior_push(IorStack& s, int o)
{
    if (s.n >= 3) return false;
    s.obj[s.n++] = o;
    return true;
}

__device__ inline void ior_pop(IorStack& s, int o)
{
    int at = -1;
    for (int i = 0; i < 3; ++i)
        if (i < s.n && s.obj[i] == o) at = i;
    if (at < 0) return;
    for (int i = at; i < 2; ++i) s.obj[i] = s.obj[i + 1];
    s.n--;
}

// DielectricSPF::ScatterNM + GenerateScatteredRay, smooth (DielectricSPF.cpp:
// 120-287, 395-432), then ScatteredRayContainer::RandomlySelect by krayNM
// (ScatteredRayContainer.cpp:43-92) and the integrator's division by the
// select probability (PathTracingIntegrator.cpp:2650-2670).  `from` is the
// origin of the ray that hit (for absorption over the segment inside).
__device__ inline bool dielectric_scatter(const PtSceneView& sc, const DevMaterial& m, const Surf& h, f3 from,
                                          float nm, const IorStack& st, float u_select, f3& dir, float& weight,
                                          int& type, IorStack& st_out, bool& dropped)
{
    bool push_dropped = false;
    const bool inside = ior_contains(st, h.object);
    const float rindex = spd_or(sc, m.ior_spd, nm, 1.33f);
    const float k_refr = inside ? powf(spd_or(sc, m.ext_spd, nm, 1.0f), len3(h.p - from)) : 1.0f;
    const f3 w = h.frame.w;
    const f3 neff = inside ? -w : w;
    const f3 graw = dot3(h.ng, h.ng) > 1e-12f ? h.ng : neff;
    const f3 gn = dot3(graw, neff) >= 0.0f ? graw : -graw;
    f3 refr = h.dir;
    float ref;
    IorStack st_refr = st;
    if (inside) {
        IorStack exit_st = st;
        ior_pop(exit_st, h.object);
        const float exit_ior = ior_top(sc, exit_st, nm);
        if (refract3(-w, rindex, exit_ior, refr)) {
            ior_pop(st_refr, h.object);
            ref = dielectric_reflectance(h.dir, refr, -w, rindex, exit_ior);
        } else {
            ref = 1.0f;
        }
    } else {
        const float top = ior_top(sc, st, nm);
        if (refract3(w, top, rindex, refr)) {
            ref = dielectric_reflectance(h.dir, refr, w, top, rindex);
            push_dropped = !ior_push(st_refr, h.object);
        } else {
            ref = 1.0f;
        }
    }
    bool has_refl = true, has_refr = true;
    f3 refl = reflect3(h.dir, inside ? w : -w);
    if (dot3(refl, gn) <= 0.0f) {
        if (ref >= 1.0f) refl = reflect3(h.dir, gn);
        else has_refl = false;
    }
    if (!inside && dot3(refr, w) > -kNearZeroF) has_refr = false;
    else if (inside && dot3(refr, w) < kNearZeroF) has_refr = false;
    const bool a = has_refr && ref < 1.0f;   // the refracted ray is added first
    const bool b = has_refl && ref != 0.0f;
    if (!a && !b) return false;
    const float ka = k_refr * (1.0f - ref), kb = ref;
    bool pick_a = a;
    float select = 1.0f;
    if (a && b) {
        const float total = ka + kb;
        if (!(total > kNearZeroF)) return false;
        pick_a = u_select < ka / total;
        const float kp = pick_a ? ka : kb;
        if (kp > kNearZeroF) select = kp / total;
    }
    if (select < kNearZeroF) return false;
    dropped = false;
    if (pick_a) {
        dir = refr;
        weight = ka / select;
        type = 2;
        st_out = st_refr;
        dropped = push_dropped;
    } else {
        dir = refl;
        weight = kb / select;
        type = 1;
        st_out = st;
    }
    return true;
}

// PropagateBounceLimits (PathTransportUtilities.h:170-210) on the packed
// counters (bits 0-7 diffuse, 8-15 glossy, 16-23 transmission): false when
// the scatter type's cap is exceeded.
__device__ inline bool bump_bounces(unsigned& st, int type, const ReqInfo& R)
{
    const int shift = 8 * type;
    const unsigned c = ((st >> shift) & 0xffu) + 1u;
    const int cap = type == 0 ? R.max_diffuse : (type == 1 ? R.max_glossy : R.max_transmission);
    if ((int)c > cap) return false;
    st = (st & ~(0xffu << shift)) | (c << shift);
    return true;
}

constexpr unsigned kLaneDims = 1u << 20;  // key offset of lane w's independent draws (free flight): (w + 1) << 20
constexpr unsigned kModeNm = 1u;          // flags bit: the path continues as single-wavelength lanes

constexpr unsigned kShadowInScatter = 0x80000000u;  // shadow record of an in-scattering NEE sample
constexpr unsigned kShadowNoWalk = 0x40000000u;     // shadow record added without the transmittance walk

// Shadow-record codes: slot * 4 + wavelength in the low 27 bits (slots < 2^25),
// the kind of contribution above them -- each kind adds into its own result
// array, so no two records of one bounce add into the same place -- and, for
// a directional light seen by a shared ray, a flag for all four wavelengths.
constexpr unsigned kRecSlotMask = (1u << 27) - 1u;
constexpr unsigned kRecKindShift = 27;
constexpr unsigned kRecLum = 0u, kRecEnv = 1u, kRecDir = 2u;  // kRecDir + l: directional light l
constexpr unsigned kRecAllLanes = 1u << 30;

// Sampler dimensions of one light sample: the luminary pick and point when the
// scene has luminaries (an empty alias table draws nothing,
// LightSampler.cpp:2053-2056), then the environment sample's two.
__device__ inline unsigned light_sample_dims(const PtSceneView& sc)
{
    return (sc.n_lums > 0 ? kNeeDims : 0u) + (sc.env ? 2u : 0u);
}

__device__ inline void write_shadow(Paths& S, unsigned si, f3 o, f3 d, float tmax, float val, unsigned code,
                                    unsigned info)
{
    S.sox[si] = o.x;
    S.soy[si] = o.y;
    S.soz[si] = o.z;
    S.sdx[si] = d.x;
    S.sdy[si] = d.y;
    S.sdz[si] = d.z;
    S.stmax[si] = tmax;
    S.sval[si] = val;
    S.sslot[si] = code;
    if (S.sinfo) {
        S.sinfo[si] = info;
        S.svis[si] = 0u;
    }
}

// Adds an unoccluded record's value, times the transmittance tr, to its result array.
__device__ inline void add_shadow(Paths& S, unsigned code, float val, float tr)
{
    const unsigned cap = (unsigned)S.cap, base = code & kRecSlotMask, slot = base >> 2, w = base & 3u;
    const unsigned kind = (code >> kRecKindShift) & 7u;
    float* dst = kind == kRecLum ? S.res : (kind == kRecEnv ? S.res_env : S.res_dir[kind - kRecDir]);
    if (code & kRecAllLanes) {
        const float* v = S.dir_val[kind - kRecDir];
        for (int k = 0; k < kHwss; ++k) dst[k * cap + slot] += v[k * cap + slot] * tr;
    } else {
        dst[w * cap + slot] += val * tr;
    }
}

// The object on top of a lane's IOR stack when it holds a medium, else -1
// (MediumTracking::GetCurrentMediumWithObject, Utilities/MediumTracking.h;
// no global medium).
__device__ inline int lane_medium_object(const PtSceneView& sc, const Paths& S, unsigned li)
{
    if (!sc.has_media) return -1;
    const int n = (int)((S.state[li] >> 24) & 3u);
    if (n == 0) return -1;
    const int top = S.ior_obj[(n - 1) * 4 * S.cap + li];
    return sc.objects[top].medium >= 0 ? top : -1;
}

// Bins the extension queue's hits by material kind (misses end their path),
// so each shading kernel runs one material's code over a dense queue.  A
// single-wavelength lane inside a medium first samples a free-flight
// distance (PathTracingIntegrator.cpp:1686-1712 ->
// SampleDistanceWithEquiangularMIS_NM without positional lights, :364-386,
// -> HomogeneousMedium::SampleDistanceNM, HomogeneousMedium.cpp:150-172): a
// scatter before the hit goes to the volume bin with its distance in hit_t;
// otherwise the lane takes the survival weight Tr / pSurvival (:1895-1910,
// IMedium.h:235-251) and goes on to the surface or the escape.
__global__ __launch_bounds__(kShadeBlock) void k_bin(const ReqInfo* __restrict__ reqs, Paths S, int depth)
{
    __shared__ unsigned lds[2 * kBins * (kShadeBlock / 64 + 1)];
    BlockPush<kShadeBlock, kBins> push{lds, 0};
    const unsigned n = S.cnt_ext[depth];
    const unsigned* q = S.queue[depth & 1];
    unsigned* const cnt[kBins] = {S.cnt_mat + depth, S.cnt_mat + (kMaxPathDepth + 1) + depth,
                                  S.cnt_mat + 2 * (kMaxPathDepth + 1) + depth, S.cnt_mat + 3 * (kMaxPathDepth + 1) + depth};
    const unsigned stride = gridDim.x * blockDim.x;
    for (unsigned b = blockIdx.x * blockDim.x; b < n; b += stride) {
        const unsigned i = b + threadIdx.x;
        int kind = -1;
        unsigned code = 0;
        if (i < n) {
            code = q[i];
            const unsigned slot = code >> 2, lane = code & 3u, li = lane * (unsigned)S.cap + slot;
            const unsigned leaf = S.hit_leaf[li];
            const ReqInfo& R = reqs[S.req[slot]];
            const PtSceneView& sc = R.sc;
            const bool nm = S.flags[slot] & kModeNm;
            bool volume = false;
            const int mobj = (nm && S.vb) ? lane_medium_object(sc, S, li) : -1;
            if (mobj >= 0) {
                const DevMedium M = sc.media[sc.objects[mobj].medium];
                const bool hit = leaf != 0xffffffffu;
                const float maxd = hit ? S.hit_t[li] : INFINITY;
                bool scattered = false;
                float tm = 0.0f;
                if (M.sigma_t > 0.0f) {
                    const unsigned dim = 32u * (unsigned)(16 + depth) + kLaneDims * (lane + 1u) + 20u;
                    const float xi = rng_u(R.seed, S.path_id[slot], dim);
                    const float tt = -logf(fmaxf(1.0f - xi, 1e-30f)) / M.sigma_t;
                    if (tt < maxd) {
                        scattered = true;
                        tm = tt;
                    }
                }
                if (scattered && (int)S.vb[li] < kMaxVolumeBounce) {
                    volume = true;
                    S.hit_t[li] = tm;
                } else if (!scattered && M.sigma_t > 0.0f) {  // sigma_t = 0: Tr = pSurvival = 1
                    const float tr = expf(-M.sigma_t * maxd);
                    S.thr[li] *= tr / fmaxf(tr, 1e-30f);
                }
            }
            if (volume) {
                kind = kMatVolume;
            } else if (leaf != 0xffffffffu) {
                const int obj = float_to_bits(reinterpret_cast<const float4*>(sc.tris + leaf)[1].w);
                kind = sc.materials[sc.objects[obj].material].type;
            } else if (sc.sky.on && (depth > 0 || R.sky_background)) {
                // Escape (PathTracingIntegrator.cpp:4230-4290, 5070-5110 and the
                // single-wavelength twin): the sky at each wavelength the ray
                // carries.
                float lam[kHwss];
                wavelengths(S.lambda0[slot], R.lambda_begin, R.lambda_end, lam);
                const f3 d = mk3(S.dx[li], S.dy[li], S.dz[li]);
                // A BSDF-sampled ray is weighed against the environment light
                // sample by the power heuristic (:4259-4282); camera rays and
                // rays leaving a delta lobe (pdf 0) keep full weight.
                float wmis = 1.0f;
                const float pin = S.pdf[li];
                if (sc.env && pin > 0.0f) {
                    const double dd[3] = {d.x, d.y, d.z};
                    const float pe = (float)env_pdf(*sc.env, dd);
                    if (pe > 0.0f) wmis = power_heuristic(pin, pe);
                }
                for (int w = 0; w < kHwss; ++w)
                    if (!nm) S.res[w * S.cap + slot] += S.thr[w * S.cap + slot] * (sky_radiance(sc.sky, d, lam[w]) * wmis);
                if (nm) S.res[li] += S.thr0[li] * (S.thr[li] * (sky_radiance(sc.sky, d, lam[lane]) * wmis));
            }
        }
        const unsigned at = push(cnt, kind);  // kind kMatVolume only when the pass has media
        if (kind >= 0) S.mqueue[kind][at] = code;
    }
}


// Origin of a ray leaving a surface: offset off it toward the ray's side.
__device__ inline f3 surface_origin(const Surf& h, f3 dir)
{
    return offset_ray(h.p, dot3(dir, h.ng) > 0.0f ? h.ng : -h.ng);
}

// Writes lane w's next ray from `no` and appends it to the next extension
// queue at `at`.
__device__ inline void write_continuation(Paths& S, f3 no, unsigned slot, unsigned w, unsigned* at, f3 dir, float pdf,
                                          float imp, unsigned bounces, const IorStack& st, unsigned vbn)
{
    const unsigned cap = (unsigned)S.cap, lw = w * cap + slot;
    *at = slot * 4u + w;
    if (S.vb) S.vb[lw] = (unsigned char)vbn;
    S.ox[lw] = no.x;
    S.oy[lw] = no.y;
    S.oz[lw] = no.z;
    S.dx[lw] = dir.x;
    S.dy[lw] = dir.y;
    S.dz[lw] = dir.z;
    S.pdf[lw] = pdf;
    S.imp[lw] = imp;
    S.state[lw] = (bounces & 0x00ffffffu) | ((unsigned)st.n << 24);
    for (int k = 0; k < 3; ++k)
        if (k < st.n) S.ior_obj[k * 4 * cap + lw] = st.obj[k];
}

// One bounce for the active queue: IntegrateFromHitHWSS
// (PathTracingIntegrator.cpp:3939-4822) for paths still sharing one ray, and
// the single-wavelength loop IntegrateFromHitNM (:1561-3377) for paths that
// hit a dielectric, whose wavelengths then continue as separate lanes.  Every
// thread of a wave runs the same loop trips (the queue bound is uniform per
// block) so the compaction ballots see all lanes; each shadow record is
// written as soon as it is made, which keeps one record live at a time.
template <int MAT>
__global__ __launch_bounds__(kShadeBlock) void k_shade(const ReqInfo* __restrict__ reqs, Paths S, int depth)
{
    __shared__ unsigned lds[2 * (kShadeBlock / 64 + 1)];
    BlockPush<kShadeBlock, 1> push{lds, 0};
    const unsigned n = S.cnt_mat[MAT * (kMaxPathDepth + 1) + depth];
    const unsigned* qin = S.mqueue[MAT];
    unsigned* qout = S.queue[(depth + 1) & 1];
    unsigned* cnt_next = S.cnt_ext + depth + 1;
    unsigned* cnt_sh = S.cnt_sh + depth;
    const unsigned base_dim = 32u * (unsigned)(16 + depth);
    const unsigned stride = gridDim.x * blockDim.x;
    const unsigned cap = (unsigned)S.cap;
    for (unsigned b = blockIdx.x * blockDim.x; b < n; b += stride) {
        const unsigned i = b + threadIdx.x;
        unsigned slot = 0, lane = 0, li = 0, leaf = 0xffffffffu;
        if (i < n) {
            const unsigned code = qin[i];
            slot = code >> 2;
            lane = code & 3u;
            li = lane * cap + slot;
            leaf = S.hit_leaf[li];
        }
        const bool valid = i < n;  // binned entries are hits on a material of kind MAT
        const ReqInfo& R = reqs[valid ? S.req[slot] : 0u];
        const PtSceneView& sc = R.sc;
        const bool nm = valid && (S.flags[slot] & kModeNm);

        // Hit record (TriangleMeshGeometryIndexed + Object layers).
        Surf h;
        DevObject ob{};
        DevMaterial mat{};
        SampleKey key{0u, 0u, 0u};
        f3 o = mk3(0, 0, 0);
        float lam[kHwss] = {0, 0, 0, 0};
        float pdf_in = 0.0f, imp = 1.0f;
        unsigned bst = 0, vbn = 0;
        int mobj = -1;  // object whose medium holds the hit point (single-wavelength lanes)
        IorStack ist{0, {0, 0, 0}};
        if (valid) {
            key = sample_key(R, S.path_id[slot]);
            o = mk3(S.ox[li], S.oy[li], S.oz[li]);
            const f3 d = mk3(S.dx[li], S.dy[li], S.dz[li]);
            wavelengths(S.lambda0[slot], R.lambda_begin, R.lambda_end, lam);
            pdf_in = S.pdf[li];
            imp = S.imp[li];
            bst = S.state[li];
            ist.n = (int)((bst >> 24) & 3u);
            for (int k = 0; k < 3; ++k) ist.obj[k] = k < ist.n ? S.ior_obj[k * 4 * cap + li] : 0;
            if (nm && S.vb) {
                vbn = S.vb[li];
                mobj = lane_medium_object(sc, S, li);
            }
            const float4* tp = reinterpret_cast<const float4*>(sc.tris + leaf);
            const float4 ta = tp[0], tb = tp[1], tc = tp[2];
            const f3 v0 = mk3(ta.x, ta.y, ta.z), v1 = mk3(tb.x, tb.y, tb.z), v2 = mk3(tc.x, tc.y, tc.z);
            h.object = float_to_bits(tb.w);
            ob = sc.objects[h.object];
            h.material = ob.material;
            mat = sc.materials[ob.material];
            h.p = o + d * S.hit_t[li];
            h.dir = d;
            const f3 fn = norm3(cross3(v1 - v0, v2 - v0));
            h.n = fn;
            h.ng = fn;
            if (ob.double_sided) {  // TriangleMeshGeometryIndexed.cpp:171-188
                if (dot3(h.n, d) > 0.0f) h.n = -h.n;
                if (dot3(h.ng, d) > 0.0f) h.ng = -h.ng;
            }
            h.frame = onb_from_w(h.n);
        }
        // A material without a BSDF hit by a shared ray: every wavelength goes
        // on as its own lane from this hit (:4348-4365).
        const bool branch = valid && !nm && MAT == kMatDielectric;
        const bool surface = valid && MAT != kMatDielectric;  // emission, NEE, BSDF sampling
        const float amb = nm ? ior_top(sc, ist, lam[lane]) : 1.0f;

        // Emission with MIS against NEE (:4422-4515).
        if (MAT == kMatLambert && valid && mat.emit_spd >= 0) {
            float mis = 1.0f;
            if (pdf_in > 0.0f && ob.luminary >= 0) {
                const DevLuminary L = sc.lums[ob.luminary];
                const float cos_l = fabsf(dot3(h.dir, h.ng));
                if (L.area > 0.0f && cos_l > 0.0f) {
                    const float dist = len3(h.p - o);
                    const float ps = L.pdf_select > 0.0f ? L.pdf_select : 1.0f;
                    mis = power_heuristic(pdf_in, ps * (dist * dist) / (L.area * cos_l));
                }
            }
            for (int w = 0; w < kHwss; ++w)
                if (!nm) S.res[w * cap + slot] += S.thr[w * cap + slot] * (emitted(sc, mat, -h.dir, h.ng, lam[w]) * mis);
            if (nm) S.res[li] += S.thr0[li] * (S.thr[li] * (emitted(sc, mat, -h.dir, h.ng, lam[lane]) * mis));
        }

        // Next-event estimation (:4521-4552 and the NM twin ->
        // LightSampler::EvaluateDirectLightingNM, Lights/LightSampler.cpp:1942-2362):
        // per wavelength of a shared ray (one for a single-wavelength lane) one
        // luminary sample, one environment sample when the sky is sampled, and
        // every directional light deterministically.  Each kind of contribution
        // has its own shadow record and result array, so no two records of a
        // bounce add into the same place.
        const unsigned ls_dims = light_sample_dims(sc);
        const bool blk_env = __syncthreads_or(surface && sc.env != nullptr);
        const bool blk_dir = __syncthreads_or(surface && sc.n_dir > 0);
        for (int w = 0; w < kHwss; ++w) {
            const bool mine = surface && (!nm || (unsigned)w == lane);
            const unsigned dim = nm ? base_dim : base_dim + ls_dims * (unsigned)w;
            bool has = false;
            float v = 0.0f, stm = 0.0f;
            f3 sorg = mk3(0, 0, 0), to = mk3(0, 0, 0);
            if (mine && sc.n_lums > 0) {
                const float usel = sample_u(key, dim);
                int li_ = 0;
                while (li_ < sc.n_lums - 1 && !(usel < sc.lums[li_].cdf_select)) ++li_;
                if (li_ != ob.luminary) {  // never the shading object itself (:2061)
                    const DevLuminary L = sc.lums[li_];
                    const float ux = sample_u(key, dim + 1), uy = sample_u(key, dim + 2), uz = sample_u(key, dim + 3);
                    // UniformRandomPoint: first CDF entry >= uz, then PointOnTriangle
                    // (TriangleMeshGeometryIndexed.cpp:207-218, GeometricUtilities.cpp:543-560).
                    int k = 0;
                    while (k < L.tri_count - 1 && sc.lum_cdf[L.tri_first + k] < uz) ++k;
                    const float4* lp = reinterpret_cast<const float4*>(sc.tris + sc.lum_tris[L.tri_first + k]);
                    const float4 la = lp[0], lb = lp[1], lc = lp[2];
                    const f3 a0 = mk3(la.x, la.y, la.z), a1 = mk3(lb.x, lb.y, lb.z), a2 = mk3(lc.x, lc.y, lc.z);
                    const float sa = sqrtf(1.0f - ux);
                    const float al = 1.0f - sa, be = sa * uy, ga = 1.0f - al - be;
                    const f3 pl = a1 * al + a2 * be + a0 * ga;
                    const f3 ln = norm3(cross3(a1 - a0, a2 - a0));
                    to = pl - h.p;
                    const float dist = len3(to);
                    if (dist > 0.0f) to = to * (1.0f / dist);
                    const float cos_s = dot3(to, h.n);
                    const float cos_l = dot3(-to, ln);
                    if (cos_l > 0.0f && cos_s > 0.0f) {
                        const DevMaterial lm = sc.materials[L.material];
                        const float le = emitted(sc, lm, -to, ln, lam[w]);
                        const float geom = L.area * cos_l / (dist * dist);
                        float contrib = le * cos_s * geom * bsdf_value<MAT>(sc, mat, h, to, lam[w], amb);
                        const float p_light = L.pdf_select * (dist * dist) / (L.area * cos_l);
                        const float p_bsdf = bsdf_pdf<MAT>(sc, mat, h, to, lam[w], amb);
                        if (p_bsdf > 0.0f) contrib *= power_heuristic(p_light, p_bsdf);
                        contrib /= L.pdf_select;
                        v = nm ? S.thr0[li] * (S.thr[li] * contrib) : S.thr[w * cap + slot] * contrib;
                        if (v > 0.0f) {
                            has = true;
                            sorg = offset_ray(h.p, dot3(to, h.ng) > 0.0f ? h.ng : -h.ng);
                            stm = dist;
                        }
                    }
                }
            }
            const unsigned si = push(cnt_sh, has);
            if (has) write_shadow(S, si, sorg, to, stm, v, slot * 4u + (unsigned)w, (unsigned)(mobj + 1));

            // Environment sample (:2278-2358): a direction from the sky's
            // importance table, the sky at this wavelength, power heuristic
            // against the BSDF's pdf, a shadow ray to infinity.
            if (blk_env) {
                bool has_e = false;
                float ve = 0.0f;
                f3 ed = mk3(0, 0, 0), eo = mk3(0, 0, 0);
                if (mine && sc.env) {
                    const unsigned ed0 = dim + (sc.n_lums > 0 ? kNeeDims : 0u);
                    double dd[3], epdf;
                    env_sample(*sc.env, sample_u(key, ed0), sample_u(key, ed0 + 1), dd, epdf);
                    ed = mk3((float)dd[0], (float)dd[1], (float)dd[2]);
                    const float cos_e = dot3(ed, h.n);
                    if (cos_e > 0.0f && epdf > 0.0) {
                        const float pe = (float)epdf;
                        float contrib = sky_radiance(sc.sky, ed, lam[w]) *
                                        bsdf_value<MAT>(sc, mat, h, ed, lam[w], amb) * cos_e / pe;
                        const float p_bsdf = bsdf_pdf<MAT>(sc, mat, h, ed, lam[w], amb);
                        if (p_bsdf > 0.0f) contrib *= power_heuristic(pe, p_bsdf);
                        ve = nm ? S.thr0[li] * (S.thr[li] * contrib) : S.thr[w * cap + slot] * contrib;
                        if (ve > 0.0f) {
                            has_e = true;
                            eo = offset_ray(h.p, dot3(ed, h.ng) > 0.0f ? h.ng : -h.ng);
                        }
                    }
                }
                const unsigned se = push(cnt_sh, has_e);
                if (has_e)
                    write_shadow(S, se, eo, ed, INFINITY, ve, (slot * 4u + (unsigned)w) | (kRecEnv << kRecKindShift),
                                 (unsigned)(mobj + 1));
            }
        }
        // Directional lights (:1968-2000 -> DirectionalLight::ComputeDirectLightingNM,
        // Lights/DirectionalLight.cpp:75-105): luminance x power x BSDF x cos along
        // the light's direction, shadow ray to infinity, no transmittance walk.  A
        // shared ray's four wavelengths share one shadow ray: their values wait in
        // dir_val and the record adds all four.
        if (blk_dir) {
            for (int l = 0; l < kMaxDirLights; ++l) {
                if (!__syncthreads_or(surface && l < sc.n_dir)) break;
                bool has_d = false;
                float vd = 0.0f;
                f3 dd = mk3(0, 0, 0), dorg = mk3(0, 0, 0);
                if (surface && l < sc.n_dir) {
                    const DevDirLight L = sc.dir[l];
                    dd = mk3(L.d[0], L.d[1], L.d[2]);
                    const float cosd = dot3(dd, h.n);
                    if (cosd > 0.0f) {
                        if (nm) {
                            vd = S.thr0[li] * (S.thr[li] * (L.le * bsdf_value<MAT>(sc, mat, h, dd, lam[lane], amb) * cosd));
                            has_d = vd > 0.0f;
                        } else {
                            for (int w = 0; w < kHwss; ++w) {
                                const float x = S.thr[w * cap + slot] *
                                                (L.le * bsdf_value<MAT>(sc, mat, h, dd, lam[w], 1.0f) * cosd);
                                S.dir_val[l][w * cap + slot] = x;
                                has_d = has_d || x > 0.0f;
                            }
                        }
                        if (has_d) dorg = offset_ray(h.p, dot3(dd, h.ng) > 0.0f ? h.ng : -h.ng);
                    }
                }
                const unsigned sd = push(cnt_sh, has_d);
                if (has_d)
                    write_shadow(S, sd, dorg, dd, INFINITY, vd,
                                 (slot * 4u + (nm ? lane : 0u)) | ((kRecDir + (unsigned)l) << kRecKindShift) |
                                     (nm ? 0u : kRecAllLanes),
                                 kShadowNoWalk | (unsigned)(mobj + 1));
            }
        }

        // Scattering and the continuation: a shared ray goes on as lane 0, a
        // single-wavelength lane as itself; a dielectric continues each lane.
        if constexpr (MAT != kMatDielectric) {
            bool cont = false;
            f3 cdir = mk3(0, 0, 0);
            float cpdf = 0.0f, cimp = 0.0f;
            unsigned cst = 0u;
            const unsigned cw = nm ? lane : 0u;
            if (surface && !nm) {
                // Hero ScatterNM, companions at the hero's direction (:4620-4718).
                float u1, u2, u3;
                const unsigned rr_dim = bsdf_draws<MAT>(key, base_dim + kHwss * ls_dims, u1, u2, u3);
                const ScatterOut hs = bsdf_scatter<MAT>(sc, mat, h, lam[0], 1.0f, u1, u2, u3);
                if (hs.ok) {
                    float thr[kHwss], sc4[kHwss];
                    for (int w = 0; w < kHwss; ++w) thr[w] = S.thr[w * cap + slot];
                    sc4[0] = hs.kray;
                    for (int w = 1; w < kHwss; ++w) {
                        float cwt = bsdf_value<MAT>(sc, mat, h, hs.dir, lam[w], 1.0f) * fabsf(dot3(hs.dir, h.n));
                        if (hs.pdf > 0.0f) cwt /= hs.pdf;
                        sc4[w] = cwt > 0.0f ? cwt : 0.0f;
                    }
                    // Russian roulette over the max across wavelengths (:4721-4768).
                    float mprev = fabsf(thr[0]), mcurr = fabsf(thr[0] * sc4[0]);
                    for (int w = 1; w < kHwss; ++w) {
                        mprev = fmaxf(mprev, fabsf(thr[w]));
                        mcurr = fmaxf(mcurr, fabsf(thr[w] * sc4[w]));
                    }
                    bool go = mcurr > kNearZeroF;
                    if (go && depth >= R.rr_min_depth) {
                        const float p = fminf(1.0f, mcurr / fmaxf(mprev, R.rr_threshold));
                        if (sample_u(key, rr_dim) >= p) {
                            go = false;
                        } else if (p > 0.0f && p < 1.0f) {
                            const float inv = 1.0f / p;
                            for (int w = 0; w < kHwss; ++w) sc4[w] *= inv;
                        }
                    }
                    unsigned nb = bst;
                    if (go && !bump_bounces(nb, hs.type, R)) go = false;
                    if (go) {
                        // Runaway guard for the next bounce (:4096-4110).
                        float mx = 0.0f;
                        bool bad = false;
                        for (int w = 0; w < kHwss; ++w) {
                            const float t = thr[w] * sc4[w];
                            if (!(fabsf(t) <= 3.0e38f)) bad = true;
                            mx = fmaxf(mx, fabsf(t));
                            S.thr[w * cap + slot] = t;
                        }
                        if (!bad && mx <= 1e6f) {
                            cont = true;
                            cdir = hs.dir;
                            cpdf = hs.pdf;
                            cimp = imp * fabsf(sc4[0]);
                            cst = nb;
                        }
                    }
                }
            } else if (surface && nm) {
                // IntegrateFromHitNM PART 3 at the lane's wavelength (:2905-3290).
                float u1, u2, u3;
                const unsigned rr_dim = bsdf_draws<MAT>(key, base_dim + ls_dims, u1, u2, u3);
                const ScatterOut so = bsdf_scatter<MAT>(sc, mat, h, lam[lane], amb, u1, u2, u3);
                float scat = so.kray;
                bool go = so.ok && fabsf(scat) > kNearZeroF;
                if (go && depth >= R.rr_min_depth) {
                    const float p = fminf(1.0f, imp * fabsf(scat) / fmaxf(imp, R.rr_threshold));
                    if (sample_u(key, rr_dim) >= p) go = false;
                    else if (p > 0.0f && p < 1.0f) scat /= p;
                }
                unsigned nb = bst;
                if (go && !bump_bounces(nb, so.type, R)) go = false;
                if (go) {
                    const float t = S.thr[lane * cap + slot] * scat;
                    S.thr[lane * cap + slot] = t;
                    if (fabsf(t) <= 1e6f) {
                        cont = true;
                        cdir = so.dir;
                        cpdf = so.pdf;
                        cimp = imp * fabsf(scat);
                        cst = nb;
                    }
                }
            }
            const unsigned qi = push(cnt_next, cont);
            if (cont) write_continuation(S, surface_origin(h, cdir), slot, cw, qout + qi, cdir, cpdf, cimp, cst, ist, vbn);
        } else {
            // A dielectric: for a lane already on its own, and for each wavelength of
            // a shared ray that branches here (:2634-2760; no NEE or RR at this vertex).
            for (int w = 0; w < kHwss; ++w) {
                bool cont = false;
                f3 nd = mk3(0, 0, 0);
                float wgt = 0.0f;
                int type = 0;
                IorStack nst = ist;
                unsigned nb = bst;
                bool dropped = false;
                if (valid && (!nm || (unsigned)w == lane)) {
                    if (dielectric_scatter(sc, mat, h, o, lam[w], ist, sample_u(key, base_dim + 2u), nd, wgt, type,
                                           nst, dropped) &&
                        bump_bounces(nb, type, R)) {
                        // A shared ray's wavelength starts its own loop at throughput 1
                        // (PathTracingIntegrator.cpp:1593), its shared throughput kept aside.
                        if (!nm) S.thr0[w * cap + slot] = S.thr[w * cap + slot];
                        const float t = (nm ? S.thr[w * cap + slot] : 1.0f) * wgt;
                        S.thr[w * cap + slot] = t;
                        cont = fabsf(t) <= 1e6f;
                    }
                }
                if (cont && dropped) atomicAdd(S.cnt_ovf, 1u);
                const unsigned qi = push(cnt_next, cont);
                if (cont)
                    write_continuation(S, surface_origin(h, nd), slot, (unsigned)w, qout + qi, nd, 0.0f, imp * fabsf(wgt),
                                       nb, nst, vbn);
            }
            if (branch) S.flags[slot] |= kModeNm;
        }
    }
}


// ---- participating media --------------------------------------------------

// HenyeyGreensteinPhaseFunction::EvaluateWithG (Materials/
// HenyeyGreensteinPhaseFunction.cpp:20-33); Pdf is the same value.
__device__ inline float hg_eval(float c, float g)
{
    const float g2 = g * g;
    const float den = 1.0f + g2 - 2.0f * g * c;
    if (den <= 0.0f) return 0.0f;
    return (1.0f - g2) / (2.0f * kTwoPi * den * sqrtf(den));
}

// Vector3Ops::Perpendicular (Utilities/Math3D/VectorsOps.h:164-188).
__device__ inline f3 perpendicular3(f3 v)
{
    const float x = fabsf(v.x), y = fabsf(v.y), z = fabsf(v.z);
    if (x < y && x < z) return mk3(0.0f, v.z, -v.y);
    if (y < z) return mk3(v.z, 0.0f, -v.x);
    return mk3(v.y, -v.x, 0.0f);
}

// HenyeyGreensteinPhaseFunction::SampleWithG (:35-56): cos(theta) by inverting
// the HG CDF, then GeometricUtilities::Perturb (GeometricUtilities.cpp:100-118)
// -- wo in the basis U = wo, V = Perpendicular(wo), W = U x V, rotated by
// theta about the basis Y axis and by phi about its X axis (row-vector
// matrices, Utilities/Math3D/MatricesOps.h:274-303) -- and back.
__device__ inline f3 hg_sample(f3 wo, float g, float u1, float u2)
{
    float ct;
    if (fabsf(g) < 1e-10f) {
        ct = 1.0f - 2.0f * u1;
    } else {
        const float inner = (1.0f - g * g) / (1.0f - g + 2.0f * g * u1);
        ct = (1.0f / (2.0f * g)) * (1.0f + g * g - inner * inner);
    }
    ct = fmaxf(-1.0f, fminf(1.0f, ct));
    const float th = acosf(ct), ph = kTwoPi * u2;
    const f3 U = norm3(wo), V = norm3(perpendicular3(wo)), W = cross3(U, V);
    const float a0 = dot3(U, wo), a1 = dot3(V, wo), a2 = dot3(W, wo);
    const float cd = cosf(th), sd = sinf(th), ca = cosf(ph), sa = sinf(ph);
    const float r0 = cd * a0 + sd * a2, r1 = a1, r2 = -sd * a0 + cd * a2;
    return norm3(U * r0 + V * (ca * r1 - sa * r2) + W * (sa * r1 + ca * r2));
}

// A volume scatter of a single-wavelength lane (PathTracingIntegrator.cpp:
// 1714-1889): weight Tr sigma_s / (sigma_t Tr); in-scattering NEE
// (MediumTransport::EvaluateInScatteringNM, Utilities/MediumTransport.cpp:
// 139-174, -> EvaluateDirectLightingNM with isVolumeScatter: cosSurface 1,
// the phase function as BSDF and MIS pdf, no shading object, so no
// self-exclusion and no occlusion test -- its record only takes the
// transmittance walk); HG continuation from the scatter point with
// bsdfPdf = the phase pdf; Russian roulette at depth + volume bounces
// (:1862-1881), which leaves the importance unchanged.
__global__ __launch_bounds__(kShadeBlock) void k_volume(const ReqInfo* __restrict__ reqs, Paths S, int depth)
{
    __shared__ unsigned lds[2 * (kShadeBlock / 64 + 1)];
    BlockPush<kShadeBlock, 1> push{lds, 0};
    const unsigned n = S.cnt_mat[kMatVolume * (kMaxPathDepth + 1) + depth];
    const unsigned* qin = S.mqueue[kMatVolume];
    unsigned* qout = S.queue[(depth + 1) & 1];
    unsigned* cnt_next = S.cnt_ext + depth + 1;
    unsigned* cnt_sh = S.cnt_sh + depth;
    const unsigned base_dim = 32u * (unsigned)(16 + depth);
    const unsigned stride = gridDim.x * blockDim.x;
    const unsigned cap = (unsigned)S.cap;
    for (unsigned b = blockIdx.x * blockDim.x; b < n; b += stride) {
        const unsigned i = b + threadIdx.x;
        const bool valid = i < n;
        unsigned slot = 0, lane = 0, li = 0;
        if (valid) {
            const unsigned code = qin[i];
            slot = code >> 2;
            lane = code & 3u;
            li = lane * cap + slot;
        }
        const ReqInfo& R = reqs[valid ? S.req[slot] : 0u];
        const PtSceneView& sc = R.sc;
        bool alive = false;
        f3 p = mk3(0, 0, 0), d = mk3(0, 0, 0);
        float nm = 0.0f, thr = 0.0f, imp = 1.0f;
        unsigned bst = 0, vbn = 0;
        SampleKey key{0u, 0u, 0u};
        int mobj = -1;
        IorStack ist{0, {0, 0, 0}};
        DevMedium M{};
        if (valid) {
            key = sample_key(R, S.path_id[slot]);
            d = mk3(S.dx[li], S.dy[li], S.dz[li]);
            const float tm = S.hit_t[li];
            p = mk3(S.ox[li], S.oy[li], S.oz[li]) + d * tm;
            float lam[kHwss];
            wavelengths(S.lambda0[slot], R.lambda_begin, R.lambda_end, lam);
            nm = lam[lane];
            imp = S.imp[li];
            bst = S.state[li];
            vbn = S.vb[li];
            ist.n = (int)((bst >> 24) & 3u);
            for (int k = 0; k < 3; ++k) ist.obj[k] = k < ist.n ? S.ior_obj[k * 4 * cap + li] : 0;
            mobj = lane_medium_object(sc, S, li);  // binned here only with a medium on top
            M = sc.media[sc.objects[mobj].medium];
            const float tr = expf(-M.sigma_t * tm);
            float mw = 0.0f;
            if (M.sigma_t > 0.0f && tr > 0.0f) mw = tr * M.sigma_s / (M.sigma_t * tr);
            if (mw > 0.0f) {
                thr = S.thr[li] * mw;
                S.thr[li] = thr;
                alive = true;
            }
        }
        // In-scattering NEE at the lane's wavelength (sampler dimensions as at a
        // single-wavelength surface vertex: the light sample, then the phase
        // function's two, then Russian roulette; PathTracingIntegrator.cpp:1760-1870).
        bool has = false;
        float v = 0.0f, dist = 0.0f;
        f3 to = mk3(0, 0, 0);
        if (alive && sc.n_lums > 0) {
            const float usel = sample_u(key, base_dim);
            int l = 0;
            while (l < sc.n_lums - 1 && !(usel < sc.lums[l].cdf_select)) ++l;
            const DevLuminary L = sc.lums[l];
            const float ux = sample_u(key, base_dim + 1), uy = sample_u(key, base_dim + 2),
                        uz = sample_u(key, base_dim + 3);
            int k = 0;
            while (k < L.tri_count - 1 && sc.lum_cdf[L.tri_first + k] < uz) ++k;
            const float4* lp = reinterpret_cast<const float4*>(sc.tris + sc.lum_tris[L.tri_first + k]);
            const float4 la = lp[0], lb = lp[1], lc = lp[2];
            const f3 a0 = mk3(la.x, la.y, la.z), a1 = mk3(lb.x, lb.y, lb.z), a2 = mk3(lc.x, lc.y, lc.z);
            const float sa = sqrtf(1.0f - ux);
            const float al = 1.0f - sa, be = sa * uy, ga = 1.0f - al - be;
            const f3 pl = a1 * al + a2 * be + a0 * ga;
            const f3 ln = norm3(cross3(a1 - a0, a2 - a0));
            to = pl - p;
            dist = len3(to);
            if (dist > 0.0f) to = to * (1.0f / dist);
            const float cos_l = dot3(-to, ln);
            if (cos_l > 0.0f) {
                const float le = emitted(sc, sc.materials[L.material], -to, ln, nm);
                const float ph = hg_eval(dot3(to, d), M.g);
                float contrib = le * 1.0f * (L.area * cos_l / (dist * dist)) * ph;
                const float p_light = L.pdf_select * (dist * dist) / (L.area * cos_l);
                if (ph > 0.0f) contrib *= power_heuristic(p_light, ph);
                contrib /= L.pdf_select;
                v = S.thr0[li] * (thr * contrib);
                has = v > 0.0f;
            }
        }
        const unsigned si = push(cnt_sh, has);
        if (has) {
            write_shadow(S, si, p, to, dist, v, slot * 4u + lane, kShadowInScatter | (unsigned)(mobj + 1));
            S.svis[si] = 1u;
        }
        // Environment sample from the scatter point (:2278-2358 with
        // isVolumeScatter: cos 1, the phase function as BSDF and MIS pdf, no
        // occlusion test without a shading object, the transmittance walk).
        const unsigned ls_dims = light_sample_dims(sc);
        if (__syncthreads_or(alive && sc.env != nullptr)) {
            bool has_e = false;
            float ve = 0.0f;
            f3 ed = mk3(0, 0, 0);
            if (alive && sc.env) {
                const unsigned ed0 = base_dim + (sc.n_lums > 0 ? kNeeDims : 0u);
                double dd[3], epdf;
                env_sample(*sc.env, sample_u(key, ed0), sample_u(key, ed0 + 1), dd, epdf);
                ed = mk3((float)dd[0], (float)dd[1], (float)dd[2]);
                if (epdf > 0.0) {
                    const float pe = (float)epdf;
                    const float ph = hg_eval(dot3(ed, d), M.g);
                    float contrib = sky_radiance(sc.sky, ed, nm) * ph / pe;
                    if (ph > 0.0f) contrib *= power_heuristic(pe, ph);
                    ve = S.thr0[li] * (thr * contrib);
                    has_e = ve > 0.0f;
                }
            }
            const unsigned se = push(cnt_sh, has_e);
            if (has_e) {
                write_shadow(S, se, p, ed, INFINITY, ve, (slot * 4u + lane) | (kRecEnv << kRecKindShift),
                             kShadowInScatter | (unsigned)(mobj + 1));
                S.svis[se] = 1u;
            }
        }
        // Directional lights (ComputeDirectLightingNM with the scatter point's
        // normal set to the ray direction, MediumTransport.cpp:160-175): cos
        // against the ray direction, the phase function as BSDF, a shadow ray,
        // no transmittance walk.
        for (int l = 0; l < kMaxDirLights; ++l) {
            if (!__syncthreads_or(alive && l < sc.n_dir)) break;
            bool has_d = false;
            float vd = 0.0f;
            f3 dd = mk3(0, 0, 0);
            if (alive && l < sc.n_dir) {
                const DevDirLight L = sc.dir[l];
                dd = mk3(L.d[0], L.d[1], L.d[2]);
                const float cosd = dot3(dd, d);
                if (cosd > 0.0f) {
                    vd = S.thr0[li] * (thr * (L.le * hg_eval(dot3(dd, d), M.g) * cosd));
                    has_d = vd > 0.0f;
                }
            }
            const unsigned sd = push(cnt_sh, has_d);
            if (has_d)
                write_shadow(S, sd, p, dd, INFINITY, vd, (slot * 4u + lane) | ((kRecDir + (unsigned)l) << kRecKindShift),
                             kShadowNoWalk | (unsigned)(mobj + 1));
        }

        // Phase sampling, Russian roulette, continuation.
        bool cont = false;
        f3 nd = mk3(0, 0, 0);
        float pp = 0.0f;
        if (alive) {
            nd = hg_sample(d, M.g, sample_u(key, base_dim + ls_dims), sample_u(key, base_dim + ls_dims + 1));
            pp = hg_eval(dot3(d, nd), M.g);
            if (pp > kNearZeroF) {
                float t = thr * (hg_eval(dot3(d, nd), M.g) / pp);
                bool go = true;
                if (depth + (int)vbn >= R.rr_min_depth) {
                    const float q = fminf(1.0f, fabsf(t) / fmaxf(imp, R.rr_threshold));
                    if (sample_u(key, base_dim + ls_dims + 2) >= q) go = false;
                    else if (q > 0.0f && q < 1.0f) t /= q;
                }
                if (go) {
                    S.thr[li] = t;
                    cont = fabsf(t) <= 1e6f;  // runaway guard of the next bounce (:1660-1683)
                }
            }
        }
        const unsigned qi = push(cnt_next, cont);
        if (cont) write_continuation(S, p, slot, lane, qout + qi, nd, pp, imp, bst, ist, vbn + 1u);
    }
}

// EvalShadowTransmittanceNM (Lights/LightSampler.cpp:369-470): the shadow
// ray from o (unit d) over max_dist, starting inside the medium of object
// `origin` (-1: none), walks boundary by boundary -- closest hits from 1e-5
// past the last boundary, at most 16 -- keeping a stack of at most 4 media
// (front face by the geometric normal: enter; back face: leave) and
// multiplying the innermost medium's transmittance over each segment.  The
// walk here intersects only the shells of objects holding a medium: in the
// reference, crossings of other objects only split segments (the same
// transmittance product) and count toward the 16 steps.
__device__ inline float shadow_walk(const PtSceneView& sc, f3 o, f3 d, float max_dist, int origin)
{
    float tr = 1.0f;
    if (!(max_dist > 0.0f)) return tr;
    int stk[4] = {origin, -1, -1, -1};
    int cnt = origin >= 0 ? 1 : 0;
    float seg = 0.0f;
    PrivStack st;
    for (int step = 0; step < 16 && seg < max_dist; ++step) {
        const float cs = seg + 1e-5f;
        if (cs >= max_dist) break;
        const f3 co = o + d * cs;
        const float cmax = max_dist - cs;
        RayState r;
        init_ray(r, co.x, co.y, co.z, d.x, d.y, d.z);
        st.sp = 0;
        double best;
        unsigned leaf, nn = 0, nt = 0;
        const int id = trace_closest<false>(r, sc.mnodes, sc.mtris, st, best, leaf, nn, nt);
        const float sig = cnt > 0 ? sc.media[sc.objects[stk[cnt - 1]].medium].sigma_t : 0.0f;
        const float t = (float)best;
        if (id < 0 || t >= cmax) {
            const float rem = max_dist - seg;
            if (rem > 0.0f && cnt > 0) tr *= expf(-sig * rem);
            seg = max_dist;
            break;
        }
        const float4* tp = reinterpret_cast<const float4*>(sc.mtris + leaf);
        const float4 a = tp[0], b = tp[1], c = tp[2];
        const int obj = float_to_bits(b.w);
        f3 ng = cross3(mk3(b.x - a.x, b.y - a.y, b.z - a.z), mk3(c.x - a.x, c.y - a.y, c.z - a.z));
        if (sc.objects[obj].double_sided && dot3(ng, d) > 0.0f) ng = -ng;  // TriangleMeshGeometryIndexed.cpp:171-188
        const float bd = cs + t;
        const float len = bd - seg;
        if (len > 0.0f && cnt > 0) tr *= expf(-sig * len);
        if (dot3(ng, d) < 0.0f) {
            if (cnt < 4) stk[cnt++] = obj;
        } else {
            int at = -1;
            for (int k = 0; k < 4; ++k)
                if (at < 0 && k < cnt && stk[k] == obj) at = k;
            if (at >= 0) {
                for (int k = 0; k < 3; ++k)
                    if (k >= at) stk[k] = stk[k + 1];
                --cnt;
            }
        }
        seg = bd;
        if (tr < 1e-6f) return 0.0f;
    }
    if (seg < max_dist && cnt > 0) {
        const float rem = max_dist - seg;
        if (rem > 0.0f) tr *= expf(-sc.media[sc.objects[stk[cnt - 1]].medium].sigma_t * rem);
    }
    return tr;
}

// Adds the unoccluded shadow records of scenes with media, times the
// transmittance of the walk (LightSampler.cpp:2222-2228).
__global__ void k_walk(const ReqInfo* __restrict__ reqs, Paths S, int depth)
{
    const unsigned n = S.cnt_sh[depth];
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (!S.svis[i]) continue;
        const unsigned code = S.sslot[i];
        const ReqInfo& R = reqs[S.req[(code & kRecSlotMask) >> 2]];
        const unsigned info = S.sinfo[i];
        const int origin = (int)(info & ~(kShadowInScatter | kShadowNoWalk)) - 1;
        const float tr = (info & kShadowNoWalk)
                             ? 1.0f
                             : shadow_walk(R.sc, mk3(S.sox[i], S.soy[i], S.soz[i]), mk3(S.sdx[i], S.sdy[i], S.sdz[i]),
                                           S.stmax[i], origin);
        add_shadow(S, code, S.sval[i], tr);
    }
}

// Waves per SIMD k_connect's registers are budgeted for: 5, i.e. at most 96
// VGPRs (it needs 98 unbudgeted, which leaves 4).
constexpr int kConnectWaves = 5;

// Shadow rays (LightSampler::ShadowOccludedNM -> ObjectManager::IntersectShadowRay)
// to 0.001 short of the light point (LightSampler.cpp:2206); adds each
// unoccluded contribution to its path and wavelength, or, in a scene with
// media, marks it for the transmittance walk.  In-scattering records take no
// occlusion test.  There is one record per (path, wavelength) and bounce, so
// the additions never collide.  The same two traversal schedules as k_extend.
template <int STACK, int BLOCK, bool PERSIST>
__global__ __launch_bounds__(BLOCK, kConnectWaves) void k_connect(const ReqInfo* __restrict__ reqs, Paths S, int depth)
{
    __shared__ uint2 lds[STACK > 0 ? STACK * BLOCK : 1];
    LdsStack<STACK> lst{lds + threadIdx.x, BLOCK, 0};
    PrivStack pst;
    pst.sp = 0;
    const unsigned n = S.cnt_sh[depth];
    if constexpr (!PERSIST) {
        for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
            if (S.sinfo && (S.sinfo[i] & kShadowInScatter)) continue;
            const unsigned code = S.sslot[i];
            const ReqInfo& R = reqs[S.req[(code & kRecSlotMask) >> 2]];
            RayState r;
            init_ray(r, S.sox[i], S.soy[i], S.soz[i], S.sdx[i], S.sdy[i], S.sdz[i]);
            unsigned nn = 0, nt = 0;
            lst.sp = 0;
            pst.sp = 0;
            const double how_far2 = scale_how_far(r, (double)(S.stmax[i] - 0.001f));
            const bool occ = STACK > 0 ? trace_any<false>(r, R.sc.nodes, R.sc.tris, lst, how_far2, nn, nt)
                                       : trace_any<false>(r, R.sc.nodes, R.sc.tris, pst, how_far2, nn, nt);
            if (!occ) {
                if (R.sc.has_media) S.svis[i] = 1u;
                else add_shadow(S, code, S.sval[i], 1.0f);
            }
        }
        return;
    }
    unsigned long long* next = S.fetch + (kMaxPathDepth + 1) + depth;
    RayState r;
    AnyState s;
    const NodeSlot* nodes = nullptr;
    const Tri48* tris = nullptr;
    bool media = false;
    unsigned nn = 0, nt = 0;
    long long item = -1;
    const int64_t grid_threads = (int64_t)gridDim.x * blockDim.x;
    for (bool first_pass = true;; first_pass = false) {
        // The first item of every thread is assigned statically; later ones come from the counter.
        const long long got = first_pass ? first_item(n) : refill(item == -1, item >= 0, next, n, grid_threads);
        if (item == -1 && got != -1) {
            item = got;
            if (item >= 0 && S.sinfo && (S.sinfo[item] & kShadowInScatter)) {
                item = -1;  // no occlusion test; the walk adds it
            } else if (item >= 0) {
                const unsigned slot = (S.sslot[item] & kRecSlotMask) >> 2;
                const ReqInfo& R = reqs[S.req[slot]];
                nodes = R.sc.nodes;
                tris = R.sc.tris;
                media = R.sc.has_media;
                init_ray(r, S.sox[item], S.soy[item], S.soz[item], S.sdx[item], S.sdy[item], S.sdz[item]);
                any_begin(r, scale_how_far(r, (double)(S.stmax[item] - 0.001f)), s);
                lst.sp = 0;
                pst.sp = 0;
            }
        }
        if (__ballot(item >= 0) == 0ull && __ballot(item == -1) == 0ull) break;
        if (item >= 0) {
            const bool more = STACK > 0 ? any_step<false>(r, nodes, tris, lst, s, nn, nt)
                                        : any_step<false>(r, nodes, tris, pst, s, nn, nt);
            if (!more) {
                if (!s.hit) {
                    if (media) S.svis[item] = 1u;
                    else add_shadow(S, S.sslot[item], S.sval[item], 1.0f);
                }
                item = -1;
            }
        }
    }
}

// Kernel tables over the traversal variants: stack (LDS 8, LDS 16, private),
// workgroup size (64, 128, 256), schedule (persistent or one ray per thread).
using TraceKernel = void (*)(const ReqInfo*, Paths, int);

template <int S, int B, bool P>
struct ExtendOf {
    static constexpr TraceKernel fn = k_extend<S, B, P>;
};
template <int S, int B, bool P>
struct ConnectOf {
    static constexpr TraceKernel fn = k_connect<S, B, P>;
};

template <template <int, int, bool> class K, bool P>
TraceKernel pick_block(int stack, int block)
{
    if (stack == 8) return block == 64 ? K<8, 64, P>::fn : block == 256 ? K<8, 256, P>::fn : K<8, 128, P>::fn;
    if (stack == 16) return block == 64 ? K<16, 64, P>::fn : block == 256 ? K<16, 256, P>::fn : K<16, 128, P>::fn;
    return block == 64 ? K<0, 64, P>::fn : block == 256 ? K<0, 256, P>::fn : K<0, 128, P>::fn;
}

template <template <int, int, bool> class K>
TraceKernel pick_trace(int stack, int block, bool persist)
{
    return persist ? pick_block<K, true>(stack, block) : pick_block<K, false>(stack, block);
}

// Per-pixel FP32 sums with a Kahan compensation term, samples in index order.
__device__ inline void kahan_add(float& sum, float& comp, float x)
{
    const float y = x - comp;
    const float t = sum + y;
    comp = (t - sum) - y;
    sum = t;
}

__global__ void k_film(const ReqInfo* __restrict__ reqs, const PassEntry* __restrict__ ent, int n_ent, int n_film,
                       Paths S)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n_film; i += gridDim.x * blockDim.x) {
        const int e0 = find_entry<&PassEntry::film0>(ent, n_ent, i);
        const PassEntry E = ent[e0];
        const ReqInfo& R = reqs[E.req];
        const int local = i - E.film0;
        const int pix = E.p0 + local;
        float s[3], c[3], q[3], qc[3];
        for (int k = 0; k < 3; ++k) {
            s[k] = R.fsum[3 * pix + k];
            c[k] = R.fcomp[3 * pix + k];
            q[k] = R.fsq[3 * pix + k];
            qc[k] = R.fsqc[3 * pix + k];
        }
        // This chunk and the later chunks of the same pixel block in this pass,
        // in chunk (= sample) order.
        for (int e = e0; e < n_ent && ent[e].req == E.req; ++e) {
          const PassEntry F = ent[e];
          if (F.p0 != E.p0) continue;
          for (int j = 0; j < F.ns; ++j) {
            const int slot = F.slot0 + local * F.ns + j;
            float lam[kHwss];
            wavelengths(S.lambda0[slot], R.lambda_begin, R.lambda_end, lam);
            float X = 0.0f, Y = 0.0f, Z = 0.0f;
            for (int w = 0; w < kHwss; ++w) {
                float x, y, z;
                if (xyz_from_nm(lam[w], x, y, z)) {
                    // Luminaries, emission and escapes; the environment sample;
                    // each directional light: always in this order.
                    float r = S.res[w * S.cap + slot];
                    if (S.res_env) r += S.res_env[w * S.cap + slot];
                    for (int l = 0; l < kMaxDirLights; ++l)
                        if (S.res_dir[l]) r += S.res_dir[l][w * S.cap + slot];
                    X += x * r;
                    Y += y * r;
                    Z += z * r;
                }
            }
            const float v[3] = {X * R.xyz_scale, Y * R.xyz_scale, Z * R.xyz_scale};
            for (int k = 0; k < 3; ++k) {
                kahan_add(s[k], c[k], v[k]);
                kahan_add(q[k], qc[k], v[k] * v[k]);
            }
          }
        }
        for (int k = 0; k < 3; ++k) {
            R.fsum[3 * pix + k] = s[k];
            R.fcomp[3 * pix + k] = c[k];
            R.fsq[3 * pix + k] = q[k];
            R.fsqc[3 * pix + k] = qc[k];
        }
    }
}

__global__ void k_resolve(const ReqInfo* __restrict__ reqs, int req)
{
    const ReqInfo& R = reqs[req];
    const int n = 3 * R.width * R.height;
    const float inv_n = 1.0f / (float)R.spp;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        R.out[i] = R.fsum[i] * inv_n;
        if (R.out_sq) R.out_sq[i] = R.fsq[i] * inv_n;
    }
}

size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

// Carves the path pool for `cap` paths plus `passes` counter blocks from
// `base` (NULL: sizes only); returns the size.  `env` and `n_dir`: the
// environment sample's and n_dir directional lights' result arrays, and room
// for their shadow records (4 per path and bounce for each kind).
size_t carve(char* base, int64_t cap, int passes, bool media, bool env, int n_dir, Paths* P, unsigned** counters)
{
    size_t off = 0;
    auto take = [&](size_t bytes) -> char* {
        char* p = base ? base + off : nullptr;
        off += align_up(bytes);
        return p;
    };
    const size_t c = (size_t)cap, c4 = 4 * c;
    const size_t csh = c4 * (size_t)(1 + (env ? 1 : 0) + n_dir);  // shadow records
    Paths S{};
    S.cap = (int)cap;
    S.ox = (float*)take(c4 * 4);
    S.oy = (float*)take(c4 * 4);
    S.oz = (float*)take(c4 * 4);
    S.dx = (float*)take(c4 * 4);
    S.dy = (float*)take(c4 * 4);
    S.dz = (float*)take(c4 * 4);
    S.thr = (float*)take(c4 * 4);
    S.thr0 = (float*)take(c4 * 4);
    S.res = (float*)take(c4 * 4);
    S.lambda0 = (float*)take(c * 4);
    S.pdf = (float*)take(c4 * 4);
    S.imp = (float*)take(c4 * 4);
    S.state = (unsigned*)take(c4 * 4);
    S.ior_obj = (int*)take(3 * c4 * 4);
    S.flags = (unsigned*)take(c * 4);
    S.path_id = (unsigned*)take(c * 4);
    S.req = (unsigned*)take(c * 4);
    S.hit_leaf = (unsigned*)take(c4 * 4);
    S.hit_t = (float*)take(c4 * 4);
    S.queue[0] = (unsigned*)take(c4 * 4);
    S.queue[1] = (unsigned*)take(c4 * 4);
    for (int m = 0; m < 3; ++m) S.mqueue[m] = (unsigned*)take(c4 * 4);
    if (media) {
        S.mqueue[kMatVolume] = (unsigned*)take(c4 * 4);
        S.vb = (unsigned char*)take(c4);
        S.sinfo = (unsigned*)take(csh * 4);
        S.svis = (unsigned*)take(csh * 4);
    }
    S.sox = (float*)take(csh * 4);
    S.soy = (float*)take(csh * 4);
    S.soz = (float*)take(csh * 4);
    S.sdx = (float*)take(csh * 4);
    S.sdy = (float*)take(csh * 4);
    S.sdz = (float*)take(csh * 4);
    S.stmax = (float*)take(csh * 4);
    S.sval = (float*)take(csh * 4);
    S.sslot = (unsigned*)take(csh * 4);
    if (env) S.res_env = (float*)take(c4 * 4);
    for (int l = 0; l < n_dir; ++l) {
        S.res_dir[l] = (float*)take(c4 * 4);
        S.dir_val[l] = (float*)take(c4 * 4);
    }
    unsigned* cnt = (unsigned*)take(sizeof(unsigned) * (size_t)kCounters * (size_t)std::max(1, passes));
    S.fetch = (unsigned long long*)take(sizeof(unsigned long long) * 2 * (kMaxPathDepth + 1));
    if (P) *P = S;
    if (counters) *counters = cnt;
    return off;
}

// Workgroups of `kernel` resident at once on the current device.
template <class K>
unsigned resident_grid_uncached(K kernel, int block)
{
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess || per_cu <= 0)
        per_cu = 4;
    return (unsigned)(cus * per_cu);
}

template <class K>
unsigned resident_grid(K kernel, int block)
{
    // Occupancy per (kernel, block size, device), queried once.
    static std::mutex mu;
    static std::vector<std::tuple<const void*, int, int, unsigned>> cache;
    int cur = 0;
    (void)hipGetDevice(&cur);
    {
        std::lock_guard<std::mutex> lk(mu);
        for (const auto& c : cache)
            if (std::get<0>(c) == (const void*)kernel && std::get<1>(c) == block && std::get<2>(c) == cur)
                return std::get<3>(c);
    }
    const unsigned g = resident_grid_uncached(kernel, block);
    std::lock_guard<std::mutex> lk(mu);
    cache.emplace_back((const void*)kernel, block, cur, g);
    return g;
}


unsigned grid_for(int64_t n, int block, unsigned max_blocks)
{
    const int64_t g = (n + block - 1) / block;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, max_blocks));
}

// Traversal launch shape: the node-group stack holds at most one entry per
// level below the root, and the smallest LDS stack that fits keeps occupancy
// highest; persistent waves pay off on deep trees, one ray per thread on
// shallow ones.
void choose_traversal(const rise_hip_batch_tuning& tn, int depth, int64_t max_tris, int64_t cap, int& stack,
                      int& block, bool& persist, TraceKernel& ext_fn, TraceKernel& con_fn, unsigned& ge, unsigned& gc)
{
    const int need_stack = depth - 1;
    const int auto_stack = need_stack <= 8 ? 8 : (need_stack <= 16 ? 16 : 0);
    if (tn.stack == 8 && need_stack <= 8) stack = 8;
    else if (tn.stack == 16 && need_stack <= 16) stack = 16;
    else if (tn.stack == -1) stack = 0;
    else stack = auto_stack;
    block = tn.trace_block ? tn.trace_block : 128;
    persist = tn.persistent == 1 || (tn.persistent == 0 && max_tris >= 4096);
    ext_fn = pick_trace<ExtendOf>(stack, block, persist);
    con_fn = pick_trace<ConnectOf>(stack, block, persist);
    if (persist) {
        // Never more workgroups than queue entries of a full pass.
        ge = std::min(resident_grid(ext_fn, block), grid_for(4 * cap, block, 1u << 30));
        gc = std::min(resident_grid(con_fn, block), grid_for(4 * cap, block, 1u << 30));
    } else {
        ge = grid_for(4 * cap, block, 8192);
        gc = grid_for(4 * cap, block, 8192);
    }
}

rise_hip_status validate(const rise_hip_render_request& q, size_t k)
{
    const std::string who = "request " + std::to_string(k) + ": ";
    if (!q.scene || !q.film_xyz) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, who + "NULL scene or film");
    const rise_hip_render_desc& D = q.desc;
    if (D.width <= 0 || D.height <= 0 || D.spp <= 0 || D.spp_offset < 0 || D.spp_total < D.spp_offset + D.spp)
        return fail(RISE_HIP_ERR_INVALID_ARGUMENT, who + "bad image size or sample range");
    const uint64_t npx = (uint64_t)D.width * (uint64_t)D.height;
    if (npx * (uint64_t)D.spp_total > 0xffffffffull)
        return fail(RISE_HIP_ERR_INVALID_ARGUMENT, who + "width * height * spp_total exceeds 2^32 (path id space)");
    if (!(D.lambda_end > D.lambda_begin) || D.lambda_begin < 0)
        return fail(RISE_HIP_ERR_INVALID_ARGUMENT, who + "bad wavelength range");
    if (D.max_diffuse_bounce < 0 || D.max_glossy_bounce < 0 || D.max_transmission_bounce < 0 || D.rr_min_depth < 0)
        return fail(RISE_HIP_ERR_INVALID_ARGUMENT, who + "negative bounce cap or RR depth");
    if (!(D.camera.fov_deg > 0 && D.camera.fov_deg < 180))
        return fail(RISE_HIP_ERR_INVALID_ARGUMENT, who + "bad field of view");
    if (D.sampler < RISE_HIP_SAMPLER_ZSOBOL || D.sampler > RISE_HIP_SAMPLER_ZSOBOL_UNSHUFFLED)
        return fail(RISE_HIP_ERR_INVALID_ARGUMENT, who + "unknown sampler");
    return RISE_HIP_OK;
}

rise_hip_status make_info(const rise_hip_render_request& q, ReqInfo& I)
{
    const rise_hip_render_desc& D = q.desc;
    I = ReqInfo{};
    I.sc = q.scene->view;
    I.width = D.width;
    I.height = D.height;
    I.spp = D.spp;
    I.spp_total = D.spp_total;
    I.max_diffuse = D.max_diffuse_bounce;
    I.max_glossy = D.max_glossy_bounce;
    I.max_transmission = D.max_transmission_bounce;
    I.rr_min_depth = D.rr_min_depth;
    I.rr_threshold = D.rr_threshold;
    I.lambda_begin = D.lambda_begin;
    I.lambda_end = D.lambda_end;
    const double ky = cie_y_integral(D.lambda_begin, D.lambda_end);
    const double ynorm = ky > 1e-12 ? ((double)D.lambda_end - D.lambda_begin) / ky : 1.0;
    I.xyz_scale = (float)(ynorm / kHwss);
    I.seed = D.seed;
    I.sky_background = D.sky_background;
    unsigned l2 = 0;
    while (l2 < 32u && (1ull << l2) < (unsigned long long)std::max(1, D.spp_total)) ++l2;
    I.sob_zorder = D.sampler != RISE_HIP_SAMPLER_SOBOL;
    I.sob_l2 = l2;
    I.sob_mask = D.sampler == RISE_HIP_SAMPLER_ZSOBOL_UNSHUFFLED ? 0u : (l2 >= 32u ? 0xffffffffu : (1u << l2) - 1u);
    I.seed32 = (unsigned)D.seed ^ (unsigned)(D.seed >> 32);
    const double px = D.camera.position[0], py = D.camera.position[1], pz = D.camera.position[2];
    double f[3] = {D.camera.look_at[0] - px, D.camera.look_at[1] - py, D.camera.look_at[2] - pz};
    const double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    if (!(fl > 0)) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "camera position equals look_at");
    for (double& x : f) x /= fl;
    const double up[3] = {D.camera.up[0], D.camera.up[1], D.camera.up[2]};
    double r[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
    const double rl = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    if (!(rl > 0)) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "camera up is parallel to the view direction");
    for (double& x : r) x /= rl;
    const double u[3] = {r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2], r[0] * f[1] - r[1] * f[0]};
    for (int k = 0; k < 3; ++k) {
        I.cam_pos[k] = D.camera.position[k];
        I.cam_f[k] = (float)f[k];
        I.cam_r[k] = (float)r[k];
        I.cam_u[k] = (float)u[k];
    }
    I.cam_h = (float)(2.0 * std::tan(D.camera.fov_deg * (M_PI / 180.0) * 0.5));
    I.out = q.film_xyz;
    I.out_sq = q.film_xyz_sq;
    return RISE_HIP_OK;
}

// A request's chunks, in the order they are rendered: all samples of a block
// of pixels, or a block of the samples of every pixel, at most kChunkPaths
// paths each.  The layout depends on the request only.
void chunks_of(const rise_hip_render_desc& D, int req, std::vector<PassEntry>& out)
{
    const int64_t npx = (int64_t)D.width * D.height;
    const int ns = (int)std::min<int64_t>(D.spp, std::max<int64_t>(1, kChunkPaths / npx));
    const int64_t np = std::min<int64_t>(npx, kChunkPaths / ns);
    for (int s0 = D.spp_offset; s0 < D.spp_offset + D.spp; s0 += ns)
        for (int64_t p0 = 0; p0 < npx; p0 += np) {
            PassEntry e{};
            e.req = req;
            e.p0 = (int)p0;
            e.np = (int)std::min<int64_t>(np, npx - p0);
            e.s0 = s0;
            e.ns = std::min(ns, D.spp_offset + D.spp - s0);
            out.push_back(e);
        }
}

// The chunks still to render of one request.
struct PassSource {
    const std::vector<PassEntry>* chunks;
    size_t* next;
};

// Fills one pass from the sources: round-robin, each round taking the next
// chunk of every source in turn while it fits in `cap` paths and the pass has
// fewer than `max_entries` chunks, so every source gets one chunk before any
// gets a second.  A source's chunks are emitted together in chunk order;
// film threads go to the first chunk of each pixel block (PassEntry).
// Returns the paths and film threads of the pass; `taken[k]` = chunks of
// source k in it.
void fill_pass(const std::vector<PassSource>& src, int64_t cap, size_t max_entries, std::vector<PassEntry>& out,
               int64_t& paths, int64_t& film, std::vector<int>& taken)
{
    std::vector<std::vector<PassEntry>> per(src.size());
    taken.assign(src.size(), 0);
    paths = 0;
    size_t count = 0;
    for (bool added = true; added;) {
        added = false;
        for (size_t k = 0; k < src.size() && count < max_entries; ++k) {
            const std::vector<PassEntry>& ch = *src[k].chunks;
            size_t& nx = *src[k].next;
            if (nx >= ch.size()) continue;
            const int64_t pp = (int64_t)ch[nx].np * ch[nx].ns;
            if (paths + pp > cap) continue;
            per[k].push_back(ch[nx]);
            paths += pp;
            ++nx;
            ++count;
            ++taken[k];
            added = true;
        }
    }
    int64_t slot = 0;
    film = 0;
    for (size_t k = 0; k < src.size(); ++k) {
        const size_t first = out.size();
        for (PassEntry e : per[k]) {
            bool covered = false;
            for (size_t j = first; j < out.size(); ++j) covered = covered || out[j].p0 == e.p0;
            e.slot0 = (int)slot;
            e.film0 = (int)film;
            e.nf = covered ? 0 : e.np;
            slot += (int64_t)e.np * e.ns;
            film += e.nf;
            out.push_back(e);
        }
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Batches
// ---------------------------------------------------------------------------

struct rise_hip_batch {
    int device = 0;
    int n_req = 0;
    int bounces = 1;                   // bounce iterations per pass
    int stack = 0;                     // LDS stack entries per lane (0: private array)
    int trace_block = 128;             // workgroup size of the traversal kernels
    bool persist = true;               // persistent traversal waves (else one ray per thread)
    TraceKernel ext_fn = nullptr, con_fn = nullptr;
    unsigned grid_ext = 1, grid_con = 1;  // workgroups of the traversal kernels
    unsigned mat_mask = 0;                // material kinds present in the batch's scenes
    int64_t cap = 0;                   // path pool size
    std::vector<rise_hip_render_request> requests;
    std::vector<PassEntry> entries;    // all passes, concatenated
    std::vector<int> pass_first;       // n_pass + 1 offsets into entries
    std::vector<int> pass_paths, pass_film;
    std::vector<int64_t> req_px;       // pixels per request
    hipStream_t last_stream = nullptr;
    bool launched = false;
    // device
    void* work = nullptr;              // path pool + counters
    size_t work_bytes = 0;
    Paths S{};
    unsigned* counters = nullptr;      // kCounters per pass
    ReqInfo* d_info = nullptr;
    PassEntry* d_entries = nullptr;
    float* d_film = nullptr;           // film accumulators of every request
    size_t film_floats = 0;
    size_t info_cap = 0, entries_cap = 0, film_cap = 0;
};

namespace {

void free_batch(rise_hip_batch* b)
{
    if (!b) return;
    (void)hipFree(b->work);
    (void)hipFree(b->d_info);
    (void)hipFree(b->d_entries);
    (void)hipFree(b->d_film);
    delete b;
}

template <class T>
hipError_t ensure(T** p, size_t& cap_elems, size_t need)
{
    need = std::max<size_t>(1, need);
    if (*p && need <= cap_elems) return hipSuccess;
    (void)hipFree(*p);
    *p = nullptr;
    cap_elems = 0;
    const hipError_t e = hipMalloc((void**)p, need * sizeof(T));
    if (e == hipSuccess) cap_elems = need;
    return e;
}

bool same_requests(const rise_hip_batch* b, const rise_hip_render_request* reqs, size_t n)
{
    return b->requests.size() == n && std::memcmp(b->requests.data(), reqs, n * sizeof(*reqs)) == 0;
}

// Plans the passes and uploads the tables, reusing b's device buffers when
// they are large enough.  The caller has made sure no launch of b is in flight.
rise_hip_status plan(rise_hip_batch* b, const rise_hip_render_request* reqs, size_t n,
                     const rise_hip_batch_tuning* tuning)
{
    rise_hip_batch_tuning tn{};
    if (tuning) tn = *tuning;
    if (tn.max_paths < 0 || (tn.trace_block != 0 && tn.trace_block != 64 && tn.trace_block != 128 &&
                             tn.trace_block != 256) ||
        (tn.stack != 0 && tn.stack != 8 && tn.stack != 16 && tn.stack != -1) || tn.persistent < 0 || tn.persistent > 2)
        return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "bad tuning values");
    if (n == 0 || !reqs) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "empty batch");
    if (n > (size_t)0x7fffffff) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "too many requests");
    b->requests.clear();
    std::vector<ReqInfo> info(n);
    std::vector<std::vector<PassEntry>> chunks(n);
    int64_t total_paths = 0, max_chunk = 0;
    int bounces = 1, depth = 1;
    int64_t max_tris = 0;
    unsigned mat_mask = 0;
    bool has_media = false, has_env = false;
    int n_dir = 0;
    std::vector<int64_t> req_px(n, 0);
    size_t film_floats = 0;
    for (size_t k = 0; k < n; ++k) {
        rise_hip_status st = validate(reqs[k], k);
        if (st != RISE_HIP_OK) return st;
        st = make_info(reqs[k], info[k]);
        if (st != RISE_HIP_OK) return st;
        const rise_hip_render_desc& D = reqs[k].desc;
        chunks_of(D, (int)k, chunks[k]);
        for (const PassEntry& e : chunks[k]) max_chunk = std::max<int64_t>(max_chunk, (int64_t)e.np * e.ns);
        req_px[k] = (int64_t)D.width * D.height;
        total_paths += req_px[k] * D.spp;
        // A path makes at most 1 + (sum of the caps of the lobe kinds its scene has) hits.
        const rise_hip_pt_scene* sc = reqs[k].scene;
        // Each volume scatter takes one more bounce, up to kMaxVolumeBounce.
        const int hits = 1 + D.max_diffuse_bounce + (sc->has_glossy ? D.max_glossy_bounce : 0) +
                         (sc->has_transmission ? D.max_transmission_bounce : 0) +
                         (sc->has_media ? kMaxVolumeBounce : 0);
        bounces = std::max(bounces, std::min(kMaxPathDepth, hits));
        depth = std::max(depth, reqs[k].scene->view.depth);
        mat_mask |= reqs[k].scene->material_kinds | (sc->has_media ? 1u << kMatVolume : 0u);
        has_media = has_media || sc->has_media;
        has_env = has_env || sc->view.env != nullptr;
        n_dir = std::max(n_dir, sc->view.n_dir);
        max_tris = std::max<int64_t>(max_tris, reqs[k].scene->n_tris);
        film_floats += 12 * (size_t)req_px[k];
    }
    const int64_t want = tn.max_paths > 0 ? std::min<int64_t>(tn.max_paths, kMaxPaths) : kMaxPaths;
    const int64_t cap = std::max<int64_t>(max_chunk, std::min<int64_t>(want, total_paths));

    // Passes: round-robin over the requests (fill_pass) until every chunk is placed.
    std::vector<PassEntry> entries;
    std::vector<int> pass_first(1, 0), pass_paths, pass_film;
    std::vector<size_t> next(n, 0);
    std::vector<PassSource> src(n);
    for (size_t k = 0; k < n; ++k) src[k] = PassSource{&chunks[k], &next[k]};
    size_t left = 0;
    for (const auto& c : chunks) left += c.size();
    std::vector<int> taken;
    while (left > 0) {
        int64_t paths = 0, film = 0;
        const size_t first = entries.size();
        fill_pass(src, cap, SIZE_MAX, entries, paths, film, taken);
        if (entries.size() == first) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "internal error: empty pass");
        left -= entries.size() - first;
        pass_first.push_back((int)entries.size());
        pass_paths.push_back((int)paths);
        pass_film.push_back((int)film);
    }
    const int passes = (int)pass_paths.size();

    // Device buffers.
    hipError_t e = hipGetDevice(&b->device);
    if (e != hipSuccess) return fail_hip(e, "hipGetDevice");
    const size_t need = carve(nullptr, cap, passes, has_media, has_env, n_dir, nullptr, nullptr);
    if (need > b->work_bytes || !b->work) {
        (void)hipFree(b->work);
        b->work = nullptr;
        b->work_bytes = 0;
        e = hipMalloc(&b->work, need);
        if (e != hipSuccess) return fail_hip(e, "hipMalloc (path pool)");
        b->work_bytes = need;
    }
    e = ensure(&b->d_film, b->film_cap, film_floats);
    if (e != hipSuccess) return fail_hip(e, "hipMalloc (film accumulators)");
    size_t off = 0;
    for (size_t k = 0; k < n; ++k) {
        const size_t f = 3 * (size_t)req_px[k];
        info[k].fsum = b->d_film + off;
        info[k].fcomp = b->d_film + off + f;
        info[k].fsq = b->d_film + off + 2 * f;
        info[k].fsqc = b->d_film + off + 3 * f;
        off += 4 * f;
    }
    e = ensure(&b->d_info, b->info_cap, n);
    if (e == hipSuccess) e = ensure(&b->d_entries, b->entries_cap, entries.size());
    if (e == hipSuccess) e = hipMemcpy(b->d_info, info.data(), n * sizeof(ReqInfo), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(b->d_entries, entries.data(), entries.size() * sizeof(PassEntry), hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail_hip(e, "batch tables");

    b->cap = cap;
    carve((char*)b->work, cap, passes, has_media, has_env, n_dir, &b->S, &b->counters);
    b->film_floats = film_floats;
    b->entries.swap(entries);
    b->pass_first.swap(pass_first);
    b->pass_paths.swap(pass_paths);
    b->pass_film.swap(pass_film);
    b->req_px.swap(req_px);
    choose_traversal(tn, depth, max_tris, cap, b->stack, b->trace_block, b->persist, b->ext_fn, b->con_fn,
                     b->grid_ext, b->grid_con);
    b->bounces = bounces;
    b->mat_mask = mat_mask;
    b->n_req = (int)n;
    b->requests.assign(reqs, reqs + n);
    b->launched = false;
    return RISE_HIP_OK;
}

// Kernel launches of a planned batch (no host synchronization, no copies).
// Launch shape of one pass: kernel variants and grids.
struct PassShape {
    TraceKernel ext_fn, con_fn;
    unsigned ge, gc, block;
    int64_t cap;
    int bounces;
    unsigned mat_mask;
};

// Kernels of one pass over `n_ent` chunks at `ent` (paths and film threads
// given), counters of the pass in S.cnt_*: generate, the bounce loop, film.
void launch_pass(const ReqInfo* d_info, Paths S, const PassEntry* ent, int n_ent, int paths, int film,
                 const PassShape& sh, hipStream_t stream)
{
    const unsigned gs = grid_for(4 * sh.cap, kShadeBlock, RISE_SHADE_GRID);
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(256), 0, stream, S.cnt_ext, kCounters);
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(256), 0, stream, reinterpret_cast<unsigned*>(S.fetch),
                       4 * (kMaxPathDepth + 1));
    hipLaunchKernelGGL(k_generate, dim3(grid_for(paths, 256, 4096)), dim3(256), 0, stream, d_info, ent, n_ent, paths,
                       S);
    for (int d = 0; d < sh.bounces; ++d) {
        void* args[] = {(void*)&d_info, &S, &d};
        hipLaunchKernel((const void*)sh.ext_fn, dim3(sh.ge), dim3(sh.block), args, 0, stream);
        hipLaunchKernelGGL(k_bin, dim3(gs), dim3(kShadeBlock), 0, stream, d_info, S, d);
        if (sh.mat_mask & (1u << kMatLambert))
            hipLaunchKernelGGL(k_shade<kMatLambert>, dim3(gs), dim3(kShadeBlock), 0, stream, d_info, S, d);
        if (sh.mat_mask & (1u << kMatGgx))
            hipLaunchKernelGGL(k_shade<kMatGgx>, dim3(gs), dim3(kShadeBlock), 0, stream, d_info, S, d);
        if (sh.mat_mask & (1u << kMatDielectric))
            hipLaunchKernelGGL(k_shade<kMatDielectric>, dim3(gs), dim3(kShadeBlock), 0, stream, d_info, S, d);
        if (sh.mat_mask & (1u << kMatVolume))
            hipLaunchKernelGGL(k_volume, dim3(gs), dim3(kShadeBlock), 0, stream, d_info, S, d);
        hipLaunchKernel((const void*)sh.con_fn, dim3(sh.gc), dim3(sh.block), args, 0, stream);
        if (sh.mat_mask & (1u << kMatVolume))
            hipLaunchKernelGGL(k_walk, dim3(grid_for(4 * sh.cap, 256, 2048)), dim3(256), 0, stream, d_info, S, d);
    }
    hipLaunchKernelGGL(k_film, dim3(grid_for(film, 256, 4096)), dim3(256), 0, stream, d_info, ent, n_ent, film, S);
}

Paths pass_counters(Paths S, unsigned* block)
{
    S.cnt_ext = block;
    S.cnt_sh = block + (kMaxPathDepth + 1);
    S.cnt_mat = block + 2 * (kMaxPathDepth + 1);
    S.cnt_ovf = block + kOverflowCounter;
    return S;
}

// Kernel launches of a planned batch (no host synchronization, no copies).
hipError_t launch(rise_hip_batch* b, hipStream_t stream)
{
    hipLaunchKernelGGL(k_fill, dim3(grid_for((int64_t)b->film_floats, 256, 4096)), dim3(256), 0, stream, b->d_film,
                       b->film_floats);
    const PassShape sh{b->ext_fn, b->con_fn, b->grid_ext, b->grid_con, (unsigned)b->trace_block,
                       b->cap,    b->bounces, b->mat_mask};
    const int passes = (int)b->pass_paths.size();
    for (int p = 0; p < passes; ++p) {
        const Paths S = pass_counters(b->S, b->counters + (size_t)p * kCounters);
        launch_pass(b->d_info, S, b->d_entries + b->pass_first[p], b->pass_first[p + 1] - b->pass_first[p],
                    b->pass_paths[p], b->pass_film[p], sh, stream);
    }
    for (int k = 0; k < b->n_req; ++k)
        hipLaunchKernelGGL(k_resolve, dim3(grid_for(3 * b->req_px[k], 256, 2048)), dim3(256), 0, stream, b->d_info, k);
    b->last_stream = stream;
    b->launched = true;
    return hipGetLastError();
}

rise_hip_status read_stats(rise_hip_batch* b, hipStream_t stream, rise_hip_render_stats* out)
{
    const int passes = (int)b->pass_paths.size();
    std::vector<unsigned> hc((size_t)kCounters * passes);
    hipError_t e = hipMemcpyAsync(hc.data(), b->counters, hc.size() * sizeof(unsigned), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return fail_hip(e, "render statistics");
    rise_hip_render_stats s{};
    for (int p = 0; p < passes; ++p) {
        const unsigned* c = hc.data() + (size_t)p * kCounters;
        for (int d = 0; d < b->bounces; ++d) {
            s.extension_rays += c[d];
            s.shadow_rays += c[kMaxPathDepth + 1 + d];
            if (c[d]) ++s.bounces;
        }
        s.ior_overflows += c[kOverflowCounter];
        s.paths += b->pass_paths[p];
    }
    *out = s;
    return RISE_HIP_OK;
}

}  // namespace

void release_solo_batch(rise_hip_pt_scene* s)
{
    if (s && s->solo) {
        if (s->solo->launched) (void)hipStreamSynchronize(s->solo->last_stream);
        free_batch(s->solo);
        s->solo = nullptr;
    }
}

extern "C" rise_hip_status rise_hip_batch_create_tuned(const rise_hip_render_request* requests, size_t n_requests,
                                                       const rise_hip_batch_tuning* tuning, rise_hip_batch** out_batch)
{
    clear_error();
    if (!out_batch) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "out_batch is NULL");
    *out_batch = nullptr;
    rise_hip_batch* b = new (std::nothrow) rise_hip_batch;
    if (!b) return fail(RISE_HIP_ERR_OUT_OF_MEMORY, "host allocation failed");
    const rise_hip_status st = plan(b, requests, n_requests, tuning);
    if (st != RISE_HIP_OK) {
        free_batch(b);
        return st;
    }
    *out_batch = b;
    return RISE_HIP_OK;
}

extern "C" rise_hip_status rise_hip_batch_create(const rise_hip_render_request* requests, size_t n_requests,
                                                 rise_hip_batch** out_batch)
{
    return rise_hip_batch_create_tuned(requests, n_requests, nullptr, out_batch);
}

extern "C" rise_hip_status rise_hip_batch_get_tuning(const rise_hip_batch* batch, rise_hip_batch_tuning* out)
{
    clear_error();
    if (!batch || !out) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    out->max_paths = (int32_t)batch->cap;
    out->trace_block = batch->trace_block;
    out->stack = batch->stack > 0 ? batch->stack : -1;
    out->persistent = batch->persist ? 1 : 2;
    return RISE_HIP_OK;
}

extern "C" rise_hip_status rise_hip_batch_launch(rise_hip_batch* batch, hipStream_t stream)
{
    clear_error();
    if (!batch) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "batch is NULL");
    const hipError_t e = launch(batch, stream);
    if (e != hipSuccess) return fail_hip(e, "batch launch");
    return RISE_HIP_OK;
}

extern "C" rise_hip_status rise_hip_batch_get_stats(rise_hip_batch* batch, hipStream_t stream,
                                                    rise_hip_render_stats* out_stats)
{
    clear_error();
    if (!batch || !out_stats) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!batch->launched) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "batch has not been launched");
    return read_stats(batch, stream, out_stats);
}

extern "C" rise_hip_status rise_hip_batch_release(rise_hip_batch* batch)
{
    clear_error();
    free_batch(batch);
    return RISE_HIP_OK;
}

extern "C" rise_hip_status rise_hip_render(rise_hip_pt_scene* scene, const rise_hip_render_desc* desc, float* film_xyz,
                                           float* film_xyz_sq, rise_hip_render_stats* stats, hipStream_t stream)
{
    clear_error();
    if (!scene || !desc || !film_xyz) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    rise_hip_render_request q;
    std::memset(&q, 0, sizeof(q));
    q.scene = scene;
    q.desc = *desc;
    q.film_xyz = film_xyz;
    q.film_xyz_sq = film_xyz_sq;
    if (!scene->solo) {
        scene->solo = new (std::nothrow) rise_hip_batch;
        if (!scene->solo) return fail(RISE_HIP_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
    rise_hip_batch* b = scene->solo;
    if (!same_requests(b, &q, 1)) {
        if (b->launched) {
            const hipError_t e = hipStreamSynchronize(b->last_stream);
            if (e != hipSuccess) return fail_hip(e, "previous render");
        }
        const rise_hip_status st = plan(b, &q, 1, nullptr);
        if (st != RISE_HIP_OK) {
            b->requests.clear();
            return st;
        }
    }
    const hipError_t e = launch(b, stream);
    if (e != hipSuccess) return fail_hip(e, "render launch");
    if (stats) return read_stats(b, stream, stats);
    return RISE_HIP_OK;
}

// ---------------------------------------------------------------------------
// Serving: continuous batching of render requests
// ---------------------------------------------------------------------------
//
// Requests are submitted at any time and admitted while their film
// accumulators fit the HBM budget and a request slot is free.  Each pass takes
// the next chunk of every admitted request (the chunks of rise_hip_batch), so
// a request that arrives while others render shares the next pass with them.
// Passes are launched ahead on the server's stream (a ring of kRing passes in
// flight), each with its chunk table staged in pinned memory; a request is
// done once the pass holding its last chunk, and its resolve, have run.  The
// chunking and the per-pixel summation order are those of rise_hip_render, so
// every film is bit-identical to rendering that request alone.

namespace {

constexpr int kRing = 4;          // passes in flight
constexpr int kServerDirLights = 1;  // directional lights a served scene may have (a sky's sun)
constexpr int kServerSlots = 256; // admitted requests at most
constexpr int kPassEntries = 1024; // chunks per server pass at most
constexpr int kAllocRetries = 3;    // failed film allocations before a request that runs alone fails

struct ServerReq {
    uint64_t ticket = 0;
    rise_hip_render_request q{};
    std::vector<PassEntry> chunks;
    size_t next = 0;
    int slot = -1;
    float* acc = nullptr;          // film accumulators, 12 floats per pixel
    size_t acc_bytes = 0;
    bool any_launched = false;     // some chunk has been put in a pass
    uint64_t last_pass = 0;        // sequence number of the last pass holding one of its chunks
    int status = RISE_HIP_REQUEST_DONE;
    int64_t deadline_ns = 0;       // 0: none
    int alloc_failures = 0;        // failed film allocations so far
    int64_t submit_ns = 0;
    int64_t px = 0;
    int hits = 1;                  // bounce iterations this request needs
    int depth = 1;                 // tree depth of its scene
    int64_t tris = 0;
    unsigned kinds = 0;
};

int64_t now_ns()
{
    return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

}  // namespace

struct rise_hip_server {
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t budget = 0, used = 0;
    int max_slots = kServerSlots;
    rise_hip_batch_tuning tuning{};
    int64_t cap = 0;
    void* work = nullptr;
    Paths S{};
    unsigned* d_counters = nullptr;  // kRing blocks of kCounters
    unsigned* h_counters = nullptr;  // pinned copies of the same
    ReqInfo* d_info = nullptr;       // kServerSlots
    ReqInfo* h_info = nullptr;       // pinned staging, kServerSlots
    PassEntry* d_ent = nullptr;      // kRing * kPassEntries
    PassEntry* h_ent = nullptr;      // pinned staging
    hipEvent_t ev[kRing] = {};
    bool busy[kRing] = {};
    int pass_bounces[kRing] = {};
    int64_t pass_paths[kRing] = {};
    uint64_t pass_seq[kRing] = {};
    uint64_t seq = 0;                // passes launched
    uint64_t retired_upto = 0;       // every pass with sequence < this has completed
    int ring_next = 0;
    uint64_t next_ticket = 1;
    std::deque<ServerReq> waiting;
    std::vector<ServerReq> active;   // admission order
    std::vector<int> free_slots;
    std::deque<rise_hip_server_result> done;  // finished, not yet reported
    int inject_alloc_failures = 0;   // test hook: film allocations to report as failed
    rise_hip_server_stats stats{};
};

namespace {

void server_free(rise_hip_server* sv)
{
    if (!sv) return;
    for (auto& r : sv->active)
        if (r.acc) (void)hipFreeAsync(r.acc, sv->stream);
    (void)hipStreamSynchronize(sv->stream);
    for (int k = 0; k < kRing; ++k)
        if (sv->ev[k]) (void)hipEventDestroy(sv->ev[k]);
    (void)hipFree(sv->work);
    (void)hipFree(sv->d_counters);
    (void)hipFree(sv->d_info);
    (void)hipFree(sv->d_ent);
    (void)hipHostFree(sv->h_counters);
    (void)hipHostFree(sv->h_info);
    (void)hipHostFree(sv->h_ent);
    delete sv;
}

// Marks completed passes and requests; counts their rays.
void server_retire(rise_hip_server* sv, bool block)
{
    for (int k = 0; k < kRing; ++k) {
        if (!sv->busy[k]) continue;
        const hipError_t e = block ? hipEventSynchronize(sv->ev[k]) : hipEventQuery(sv->ev[k]);
        if (e != hipSuccess) continue;
        sv->busy[k] = false;
        const unsigned* c = sv->h_counters + (size_t)k * kCounters;
        for (int d = 0; d < sv->pass_bounces[k]; ++d) {
            sv->stats.extension_rays += c[d];
            sv->stats.shadow_rays += c[kMaxPathDepth + 1 + d];
        }
        sv->stats.ior_overflows += c[kOverflowCounter];
        sv->stats.paths += sv->pass_paths[k];
    }
    // Passes complete in launch order on one stream: the oldest still busy bounds retired_upto.
    uint64_t lim = sv->seq;
    for (int k = 0; k < kRing; ++k)
        if (sv->busy[k]) lim = std::min(lim, sv->pass_seq[k]);
    sv->retired_upto = lim;
    const int64_t t = now_ns();
    for (size_t i = 0; i < sv->active.size();) {
        ServerReq& r = sv->active[i];
        if (r.next == r.chunks.size() && (!r.any_launched || r.last_pass < sv->retired_upto)) {
            sv->done.push_back(rise_hip_server_result{r.ticket, t - r.submit_ns, r.status, 0});
            sv->used -= (int64_t)r.acc_bytes;
            sv->free_slots.push_back(r.slot);
            sv->stats.completed += 1;
            if (r.status == RISE_HIP_REQUEST_CANCELLED) sv->stats.cancelled += 1;
            if (r.status == RISE_HIP_REQUEST_TIMED_OUT) sv->stats.timed_out += 1;
            sv->active.erase(sv->active.begin() + (long)i);
        } else {
            ++i;
        }
    }
}

// Ends a request early with `status` (cancelled or timed out).  A waiting
// request finishes at once; an admitted one gets no further chunks, its
// accumulators are freed in stream order after the passes already holding
// its chunks, its film is left as it is (not resolved), and it finishes once
// those passes have run.  A request whose last chunk is already in a pass
// completes normally.  Returns false for an unknown or finished ticket.
bool server_end(rise_hip_server* sv, uint64_t ticket, int status)
{
    for (size_t i = 0; i < sv->waiting.size(); ++i) {
        if (sv->waiting[i].ticket != ticket) continue;
        ServerReq& r = sv->waiting[i];
        sv->done.push_back(rise_hip_server_result{r.ticket, now_ns() - r.submit_ns, status, 0});
        sv->stats.completed += 1;
        if (status == RISE_HIP_REQUEST_CANCELLED) sv->stats.cancelled += 1;
        if (status == RISE_HIP_REQUEST_TIMED_OUT) sv->stats.timed_out += 1;
        sv->waiting.erase(sv->waiting.begin() + (long)i);
        return true;
    }
    for (ServerReq& r : sv->active) {
        if (r.ticket != ticket) continue;
        if (r.acc) {  // its last chunk is not in a pass yet
            (void)hipFreeAsync(r.acc, sv->stream);
            r.acc = nullptr;
            r.next = r.chunks.size();
            r.status = status;
        }
        return true;
    }
    return false;
}

// Ends every request whose deadline has passed.
void server_expire(rise_hip_server* sv)
{
    const int64_t t = now_ns();
    std::vector<uint64_t> late;
    for (const ServerReq& r : sv->waiting)
        if (r.deadline_ns > 0 && t > r.deadline_ns) late.push_back(r.ticket);
    for (const ServerReq& r : sv->active)
        if (r.deadline_ns > 0 && t > r.deadline_ns && r.acc) late.push_back(r.ticket);
    for (uint64_t k : late) (void)server_end(sv, k, RISE_HIP_REQUEST_TIMED_OUT);
}

// Admits waiting requests in arrival order while memory and slots allow.  A
// request whose film accumulators cannot be allocated stays at the head of
// the queue and is tried again at a later call (memory frees as admitted
// requests finish); one that fails kAllocRetries times while nothing else is
// admitted is reported failed.
rise_hip_status server_admit(rise_hip_server* sv)
{
    while (!sv->waiting.empty() && !sv->free_slots.empty()) {
        ServerReq& r = sv->waiting.front();
        if (sv->used + (int64_t)r.acc_bytes > sv->budget && !sv->active.empty()) break;
        const int slot = sv->free_slots.back();
        ReqInfo info;
        rise_hip_status st = make_info(r.q, info);
        if (st != RISE_HIP_OK) return st;
        hipError_t e;
        if (sv->inject_alloc_failures > 0) {
            --sv->inject_alloc_failures;
            e = hipErrorOutOfMemory;
        } else {
            e = hipMallocAsync((void**)&r.acc, r.acc_bytes, sv->stream);
        }
        if (e == hipErrorOutOfMemory) {
            (void)hipGetLastError();
            r.acc = nullptr;
            sv->stats.alloc_retries += 1;
            if (++r.alloc_failures >= kAllocRetries && sv->active.empty()) {
                sv->done.push_back(rise_hip_server_result{r.ticket, now_ns() - r.submit_ns, RISE_HIP_REQUEST_FAILED, 0});
                sv->stats.completed += 1;
                sv->stats.failed += 1;
                sv->waiting.pop_front();
                continue;
            }
            break;
        }
        if (e != hipSuccess) return fail_hip(e, "film accumulators");
        const size_t f = 3 * (size_t)r.px;
        info.fsum = r.acc;
        info.fcomp = r.acc + f;
        info.fsq = r.acc + 2 * f;
        info.fsqc = r.acc + 3 * f;
        sv->h_info[slot] = info;
        e = hipMemcpyAsync(sv->d_info + slot, sv->h_info + slot, sizeof(ReqInfo), hipMemcpyHostToDevice, sv->stream);
        if (e != hipSuccess) return fail_hip(e, "request table");
        hipLaunchKernelGGL(k_fill, dim3(grid_for((int64_t)(4 * f), 256, 4096)), dim3(256), 0, sv->stream, r.acc, 4 * f);
        sv->free_slots.pop_back();
        r.slot = slot;
        for (PassEntry& c : r.chunks) c.req = slot;
        sv->used += (int64_t)r.acc_bytes;
        sv->active.push_back(std::move(r));
        sv->waiting.pop_front();
    }
    return RISE_HIP_OK;
}

// Launches passes while the ring has room and admitted requests have chunks left.
rise_hip_status server_launch(rise_hip_server* sv)
{
    for (;;) {
        const int k = sv->ring_next;
        if (sv->busy[k]) return RISE_HIP_OK;
        PassEntry* ent = sv->h_ent + (size_t)k * kPassEntries;
        int64_t paths = 0, film = 0, tris = 0;
        int bounces = 1, depth = 1;
        unsigned kinds = 0;
        std::vector<ServerReq*> finishing;
        std::vector<PassSource> src;
        src.reserve(sv->active.size());
        for (ServerReq& r : sv->active) src.push_back(PassSource{&r.chunks, &r.next});
        std::vector<PassEntry> pass;
        std::vector<int> taken;
        fill_pass(src, sv->cap, kPassEntries, pass, paths, film, taken);
        const int n_ent = (int)pass.size();
        if (n_ent == 0) return RISE_HIP_OK;
        std::copy(pass.begin(), pass.end(), ent);
        for (size_t i = 0; i < sv->active.size(); ++i) {
            if (!taken[i]) continue;
            ServerReq& r = sv->active[i];
            r.any_launched = true;
            r.last_pass = sv->seq;
            bounces = std::max(bounces, r.hits);
            depth = std::max(depth, r.depth);
            tris = std::max(tris, r.tris);
            kinds |= r.kinds;
            if (r.next == r.chunks.size()) finishing.push_back(&r);
        }
        PassShape sh{};
        int stack = 0, block = 128;
        bool persist = true;
        choose_traversal(sv->tuning, depth, tris, sv->cap, stack, block, persist, sh.ext_fn, sh.con_fn, sh.ge, sh.gc);
        sh.block = (unsigned)block;
        sh.cap = sv->cap;
        sh.bounces = bounces;
        sh.mat_mask = kinds;
        PassEntry* d_ent = sv->d_ent + (size_t)k * kPassEntries;
        hipError_t e = hipMemcpyAsync(d_ent, ent, (size_t)n_ent * sizeof(PassEntry), hipMemcpyHostToDevice, sv->stream);
        if (e != hipSuccess) return fail_hip(e, "pass table");
        unsigned* cnt = sv->d_counters + (size_t)k * kCounters;
        launch_pass(sv->d_info, pass_counters(sv->S, cnt), d_ent, n_ent, (int)paths, (int)film, sh, sv->stream);
        for (ServerReq* r : finishing) {
            hipLaunchKernelGGL(k_resolve, dim3(grid_for(3 * r->px, 256, 2048)), dim3(256), 0, sv->stream, sv->d_info,
                               r->slot);
            e = hipFreeAsync(r->acc, sv->stream);
            if (e != hipSuccess) return fail_hip(e, "film accumulators");
            r->acc = nullptr;
        }
        e = hipMemcpyAsync(sv->h_counters + (size_t)k * kCounters, cnt, kCounters * sizeof(unsigned),
                           hipMemcpyDeviceToHost, sv->stream);
        if (e == hipSuccess) e = hipEventRecord(sv->ev[k], sv->stream);
        if (e == hipSuccess) e = hipGetLastError();
        if (e != hipSuccess) return fail_hip(e, "pass launch");
        sv->busy[k] = true;
        sv->pass_bounces[k] = bounces;
        sv->pass_paths[k] = paths;
        sv->pass_seq[k] = sv->seq++;
        sv->stats.passes += 1;
        sv->ring_next = (k + 1) % kRing;
    }
}

rise_hip_status server_report(rise_hip_server* sv, uint64_t* tickets, int64_t* latency_ns,
                              rise_hip_server_result* results, size_t cap, size_t* n_done)
{
    size_t n = 0;
    while (n < cap && !sv->done.empty()) {
        const rise_hip_server_result& d = sv->done.front();
        if (tickets) tickets[n] = d.ticket;
        if (latency_ns) latency_ns[n] = d.latency_ns;
        if (results) results[n] = d;
        sv->done.pop_front();
        ++n;
    }
    if (n_done) *n_done = n;
    return RISE_HIP_OK;
}

rise_hip_status server_poll(rise_hip_server* server, uint64_t* tickets, int64_t* latency_ns,
                            rise_hip_server_result* results, size_t capacity, size_t* n_done)
{
    clear_error();
    if (!server) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "server is NULL");
    server_retire(server, false);
    server_expire(server);
    rise_hip_status st = server_admit(server);
    if (st == RISE_HIP_OK) st = server_launch(server);
    if (st != RISE_HIP_OK) return st;
    return server_report(server, tickets, latency_ns, results, capacity, n_done);
}

rise_hip_status server_drain(rise_hip_server* server, uint64_t* tickets, int64_t* latency_ns,
                             rise_hip_server_result* results, size_t capacity, size_t* n_done)
{
    clear_error();
    if (!server) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "server is NULL");
    while (!server->waiting.empty() || !server->active.empty()) {
        server_retire(server, false);
        server_expire(server);
        rise_hip_status st = server_admit(server);
        if (st == RISE_HIP_OK) st = server_launch(server);
        if (st != RISE_HIP_OK) return st;
        if (!server->active.empty() || !server->waiting.empty()) {
            // wait for the oldest pass in flight before going round again
            int oldest = -1;
            for (int k = 0; k < kRing; ++k)
                if (server->busy[k] && (oldest < 0 || server->pass_seq[k] < server->pass_seq[oldest])) oldest = k;
            if (oldest >= 0) (void)hipEventSynchronize(server->ev[oldest]);
            else if (server->waiting.empty() && server->active.empty()) break;
        }
    }
    server_retire(server, true);
    return server_report(server, tickets, latency_ns, results, capacity, n_done);
}

}  // namespace

extern "C" rise_hip_status rise_hip_server_create(const rise_hip_server_config* config, hipStream_t stream,
                                                  rise_hip_server** out_server)
{
    clear_error();
    if (!out_server) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "out_server is NULL");
    *out_server = nullptr;
    rise_hip_server_config cfg{};
    if (config) cfg = *config;
    if (cfg.hbm_budget_bytes < 0 || cfg.max_requests < 0 || cfg.max_requests > kServerSlots)
        return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "bad server configuration");
    if (cfg.tuning.max_paths < 0 || cfg.tuning.persistent < 0 || cfg.tuning.persistent > 2 ||
        (cfg.tuning.trace_block != 0 && cfg.tuning.trace_block != 64 && cfg.tuning.trace_block != 128 &&
         cfg.tuning.trace_block != 256) ||
        (cfg.tuning.stack != 0 && cfg.tuning.stack != 8 && cfg.tuning.stack != 16 && cfg.tuning.stack != -1))
        return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "bad tuning values");
    rise_hip_server* sv = new (std::nothrow) rise_hip_server;
    if (!sv) return fail(RISE_HIP_ERR_OUT_OF_MEMORY, "host allocation failed");
    sv->stream = stream;
    sv->budget = cfg.hbm_budget_bytes > 0 ? cfg.hbm_budget_bytes : (int64_t(8) << 30);
    sv->max_slots = cfg.max_requests > 0 ? cfg.max_requests : kServerSlots;
    sv->tuning = cfg.tuning;
    sv->cap = cfg.tuning.max_paths > 0 ? std::max<int64_t>(kChunkPaths, std::min<int64_t>(cfg.tuning.max_paths, kMaxPaths))
                                       : kMaxPaths;
    hipError_t e = hipGetDevice(&sv->device);
    const size_t need = carve(nullptr, sv->cap, 1, true, true, kServerDirLights, nullptr, nullptr);
    if (e == hipSuccess) e = hipMalloc(&sv->work, need);
    if (e == hipSuccess) e = hipMalloc((void**)&sv->d_counters, sizeof(unsigned) * kCounters * kRing);
    if (e == hipSuccess) e = hipMalloc((void**)&sv->d_info, sizeof(ReqInfo) * kServerSlots);
    if (e == hipSuccess) e = hipMalloc((void**)&sv->d_ent, sizeof(PassEntry) * kPassEntries * kRing);
    if (e == hipSuccess) e = hipHostMalloc((void**)&sv->h_counters, sizeof(unsigned) * kCounters * kRing, 0);
    if (e == hipSuccess) e = hipHostMalloc((void**)&sv->h_info, sizeof(ReqInfo) * kServerSlots, 0);
    if (e == hipSuccess) e = hipHostMalloc((void**)&sv->h_ent, sizeof(PassEntry) * kPassEntries * kRing, 0);
    for (int k = 0; k < kRing && e == hipSuccess; ++k) e = hipEventCreateWithFlags(&sv->ev[k], hipEventDisableTiming);
    if (e != hipSuccess) {
        server_free(sv);
        return fail_hip(e, "server allocation");
    }
    carve((char*)sv->work, sv->cap, 1, true, true, kServerDirLights, &sv->S, nullptr);
    for (int k = sv->max_slots - 1; k >= 0; --k) sv->free_slots.push_back(k);
    *out_server = sv;
    return RISE_HIP_OK;
}

extern "C" rise_hip_status rise_hip_server_submit(rise_hip_server* server, const rise_hip_render_request* request,
                                                  uint64_t* out_ticket)
{
    clear_error();
    if (!server || !request) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    rise_hip_status st = validate(*request, 0);
    if (st != RISE_HIP_OK) return st;
    if (request->scene->view.n_dir > kServerDirLights)
        return fail(RISE_HIP_ERR_NOT_SUPPORTED, "the server renders scenes with at most " +
                                                    std::to_string(kServerDirLights) + " directional light");
    {
        ReqInfo probe;
        st = make_info(*request, probe);  // camera checks, so admission cannot fail later
        if (st != RISE_HIP_OK) return st;
    }
    ServerReq r;
    r.q = *request;
    r.ticket = server->next_ticket++;
    r.submit_ns = now_ns();
    const rise_hip_render_desc& D = request->desc;
    r.px = (int64_t)D.width * D.height;
    r.acc_bytes = 12 * sizeof(float) * (size_t)r.px;
    chunks_of(D, 0, r.chunks);
    const rise_hip_pt_scene* sc = request->scene;
    r.hits = std::min(kMaxPathDepth, 1 + D.max_diffuse_bounce + (sc->has_glossy ? D.max_glossy_bounce : 0) +
                                         (sc->has_transmission ? D.max_transmission_bounce : 0) +
                                         (sc->has_media ? kMaxVolumeBounce : 0));
    r.depth = sc->view.depth;
    r.tris = sc->n_tris;
    r.kinds = sc->material_kinds | (sc->has_media ? 1u << kMatVolume : 0u);
    for (const PassEntry& c : r.chunks)
        if ((int64_t)c.np * c.ns > server->cap) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "chunk exceeds path pool");
    if (out_ticket) *out_ticket = r.ticket;
    server->waiting.push_back(std::move(r));
    server->stats.submitted += 1;
    return RISE_HIP_OK;
}

extern "C" rise_hip_status rise_hip_server_poll(rise_hip_server* server, uint64_t* tickets, int64_t* latency_ns,
                                                size_t capacity, size_t* n_done)
{
    return server_poll(server, tickets, latency_ns, nullptr, capacity, n_done);
}

extern "C" rise_hip_status rise_hip_server_drain(rise_hip_server* server, uint64_t* tickets, int64_t* latency_ns,
                                                 size_t capacity, size_t* n_done)
{
    return server_drain(server, tickets, latency_ns, nullptr, capacity, n_done);
}

extern "C" rise_hip_status rise_hip_server_poll_results(rise_hip_server* server, rise_hip_server_result* results,
                                                        size_t capacity, size_t* n_done)
{
    return server_poll(server, nullptr, nullptr, results, capacity, n_done);
}

extern "C" rise_hip_status rise_hip_server_drain_results(rise_hip_server* server, rise_hip_server_result* results,
                                                         size_t capacity, size_t* n_done)
{
    return server_drain(server, nullptr, nullptr, results, capacity, n_done);
}

extern "C" rise_hip_status rise_hip_server_cancel(rise_hip_server* server, uint64_t ticket)
{
    clear_error();
    if (!server) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "server is NULL");
    if (!server_end(server, ticket, RISE_HIP_REQUEST_CANCELLED))
        return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "unknown or finished ticket");
    return RISE_HIP_OK;
}

extern "C" rise_hip_status rise_hip_server_set_timeout(rise_hip_server* server, uint64_t ticket, int64_t timeout_ns)
{
    clear_error();
    if (!server) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "server is NULL");
    if (timeout_ns < 0) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "negative timeout");
    for (ServerReq& r : server->waiting)
        if (r.ticket == ticket) {
            r.deadline_ns = timeout_ns > 0 ? r.submit_ns + timeout_ns : 0;
            return RISE_HIP_OK;
        }
    for (ServerReq& r : server->active)
        if (r.ticket == ticket) {
            r.deadline_ns = timeout_ns > 0 ? r.submit_ns + timeout_ns : 0;
            return RISE_HIP_OK;
        }
    return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "unknown or finished ticket");
}

extern "C" rise_hip_status rise_hip_server_fail_next_allocations(rise_hip_server* server, int32_t n)
{
    clear_error();
    if (!server || n < 0) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "bad argument");
    server->inject_alloc_failures = n;
    return RISE_HIP_OK;
}

extern "C" rise_hip_status rise_hip_server_get_stats(const rise_hip_server* server, rise_hip_server_stats* out)
{
    clear_error();
    if (!server || !out) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = server->stats;
    out->admitted = (int64_t)server->active.size();
    out->waiting = (int64_t)server->waiting.size();
    out->hbm_in_use = server->used;
    return RISE_HIP_OK;
}

extern "C" rise_hip_status rise_hip_server_release(rise_hip_server* server)
{
    clear_error();
    server_free(server);
    return RISE_HIP_OK;
}
