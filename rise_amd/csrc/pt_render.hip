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
// Queues are compacted with a wave-wide ballot and one atomic per wave; every
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
#include "pt_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

using namespace rise_hip;

namespace {

__constant__ float c_cie_x[81] = {RISE_CIE_X_VALUES};
__constant__ float c_cie_y[81] = {RISE_CIE_Y_VALUES};
__constant__ float c_cie_z[81] = {RISE_CIE_Z_VALUES};
const double h_cie_y[81] = {RISE_CIE_Y_VALUES};

constexpr int kTraceBlock = 128;
constexpr int kShadeBlock = 256;
constexpr int64_t kChunkPaths = int64_t(1) << 20;   // paths per chunk (at most)
constexpr int64_t kMaxPaths = int64_t(1) << 22;     // paths in flight per pass (at most)
constexpr int kCounters = 2 * (kMaxPathDepth + 1);  // extension counts, then shadow counts, per pass

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

// One request: scene tables, render parameters, film accumulators, output.
struct ReqInfo {
    PtSceneView sc;
    int width, height, spp, spp_total;
    int max_diffuse, max_glossy, max_transmission, rr_min_depth;
    float rr_threshold, lambda_begin, lambda_end;
    float xyz_scale;  // mYNormalization / N (PathTracingSpectralRasterizer.cpp:168-173)
    unsigned long long seed;
    float cam_pos[3], cam_f[3], cam_r[3], cam_u[3];
    float cam_h;      // 2 tan(fov / 2)
    float *fsum, *fcomp, *fsq, *fsqc;  // 3 * width * height each
    float *out, *out_sq;
};

// One chunk of a pass: pixels [p0, p0 + np) x samples [s0, s0 + ns) of
// request `req`, in pool slots [slot0, slot0 + np * ns) (pixel-major), and
// film threads [film0, film0 + np).
struct PassEntry {
    int req, p0, np, s0, ns, slot0, film0, pad;
};

// Path state, structure of arrays, `cap` entries each.
struct Paths {
    int cap;
    float *ox, *oy, *oz, *dx, *dy, *dz;
    float* thr;      // 4 * cap, wavelength-major
    float* res;      // 4 * cap
    float* lambda0;
    float* pdf;      // bsdfPdf of the ray being traced (0: camera or delta)
    unsigned* path_id;
    unsigned* req;
    unsigned* bounces;   // diffuse count (PropagateBounceLimits)
    unsigned* hit_leaf;
    float* hit_t;
    unsigned* queue[2];
    // shadow records, 4 * cap
    float *sox, *soy, *soz, *sdx, *sdy, *sdz, *stmax, *sval;
    unsigned* sslot;     // slot * 4 + wavelength
    unsigned* cnt_ext;   // [kMaxPathDepth + 1] of the current pass
    unsigned* cnt_sh;    // [kMaxPathDepth + 1]
};

// Wave-wide compaction: one atomic per wave, ranks from the ballot.
__device__ inline unsigned wave_push(unsigned* counter, bool pred)
{
    const unsigned long long m = __ballot(pred);
    if (m == 0ull) return 0xffffffffu;
    const unsigned rank =
        __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
    const int leader = __ffsll((unsigned long long)m) - 1;
    unsigned base = 0;
    if ((int)__lane_id() == leader) base = atomicAdd(counter, (unsigned)__popcll(m));
    base = __shfl(base, leader);
    return pred ? base + rank : 0xffffffffu;
}

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
        const float jx = rng_u(R.seed, path, 0), jy = rng_u(R.seed, path, 1), ul = rng_u(R.seed, path, 2);
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
        }
        S.lambda0[i] = R.lambda_begin + ul * (R.lambda_end - R.lambda_begin);
        S.pdf[i] = 0.0f;
        S.path_id[i] = path;
        S.req[i] = (unsigned)E.req;
        S.bounces[i] = 0u;
        S.queue[0][i] = (unsigned)i;
    }
}

// Closest hit for the active queue (ObjectManager::IntersectRay).
template <int STACK>
__global__ __launch_bounds__(kTraceBlock) void k_extend(const ReqInfo* __restrict__ reqs, Paths S, int depth)
{
    __shared__ uint2 lds[STACK > 0 ? STACK * kTraceBlock : 1];
    const unsigned n = S.cnt_ext[depth];
    const unsigned* q = S.queue[depth & 1];
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const unsigned slot = q[i];
        const unsigned rq = S.req[slot];
        const NodeSlot* nodes = reqs[rq].sc.nodes;
        const Tri48* tris = reqs[rq].sc.tris;
        RayState r;
        init_ray(r, S.ox[slot], S.oy[slot], S.oz[slot], S.dx[slot], S.dy[slot], S.dz[slot]);
        double best;
        unsigned leaf, nn = 0, nt = 0;
        int id;
        if (STACK > 0) {
            LdsStack<STACK> st{lds + threadIdx.x, kTraceBlock, 0};
            id = trace_closest<false>(r, nodes, tris, st, best, leaf, nn, nt);
        } else {
            PrivStack st;
            st.sp = 0;
            id = trace_closest<false>(r, nodes, tris, st, best, leaf, nn, nt);
        }
        S.hit_leaf[slot] = id >= 0 ? leaf : 0xffffffffu;
        S.hit_t[slot] = id >= 0 ? (float)best : 0.0f;
    }
}

// One bounce of IntegrateFromHitHWSS for the active queue.  Every lane of a
// wave runs the same loop trips (the queue bound is uniform per block) so the
// compaction ballots see all lanes; each shadow record is written as soon as
// it is made, which keeps one record live at a time.
__global__ __launch_bounds__(kShadeBlock) void k_shade(const ReqInfo* __restrict__ reqs, Paths S, int depth)
{
    const unsigned n = S.cnt_ext[depth];
    const unsigned* qin = S.queue[depth & 1];
    unsigned* qout = S.queue[(depth + 1) & 1];
    unsigned* cnt_next = S.cnt_ext + depth + 1;
    unsigned* cnt_sh = S.cnt_sh + depth;
    const unsigned base_dim = 32u * (unsigned)(16 + depth);
    const unsigned stride = gridDim.x * blockDim.x;
    for (unsigned b = blockIdx.x * blockDim.x; b < n; b += stride) {
        const unsigned i = b + threadIdx.x;
        unsigned slot = 0, leaf = 0xffffffffu;
        if (i < n) {
            slot = qin[i];
            leaf = S.hit_leaf[slot];
        }
        const bool valid = leaf != 0xffffffffu;  // a miss ends the path (no environment)
        const ReqInfo* R = reqs + (valid ? S.req[slot] : 0u);
        const PtSceneView& sc = R->sc;

        // Hit record (TriangleMeshGeometryIndexed + Object layers).
        Surf h;
        DevObject ob{};
        DevMaterial mat{};
        unsigned path = 0;
        f3 o = mk3(0, 0, 0);
        float thr[kHwss] = {0, 0, 0, 0}, lam[kHwss] = {0, 0, 0, 0};
        if (valid) {
            path = S.path_id[slot];
            o = mk3(S.ox[slot], S.oy[slot], S.oz[slot]);
            const f3 d = mk3(S.dx[slot], S.dy[slot], S.dz[slot]);
            for (int w = 0; w < kHwss; ++w) thr[w] = S.thr[w * S.cap + slot];
            wavelengths(S.lambda0[slot], R->lambda_begin, R->lambda_end, lam);
            const float4* tp = reinterpret_cast<const float4*>(sc.tris + leaf);
            const float4 ta = tp[0], tb = tp[1], tc = tp[2];
            const f3 v0 = mk3(ta.x, ta.y, ta.z), v1 = mk3(tb.x, tb.y, tb.z), v2 = mk3(tc.x, tc.y, tc.z);
            h.object = float_to_bits(tb.w);
            ob = sc.objects[h.object];
            h.material = ob.material;
            mat = sc.materials[ob.material];
            h.p = o + d * S.hit_t[slot];
            h.dir = d;
            const f3 fn = norm3(cross3(v1 - v0, v2 - v0));
            h.n = fn;
            h.ng = fn;
            if (ob.double_sided) {  // TriangleMeshGeometryIndexed.cpp:171-188
                if (dot3(h.n, d) > 0.0f) h.n = -h.n;
                if (dot3(h.ng, d) > 0.0f) h.ng = -h.ng;
            }
            h.frame = onb_from_w(h.n);

            // Emission with MIS against NEE (PathTracingIntegrator.cpp:4422-4515).
            if (mat.emit_spd >= 0) {
                const float bsdf_pdf = S.pdf[slot];
                float mis = 1.0f;
                if (bsdf_pdf > 0.0f && ob.luminary >= 0) {
                    const DevLuminary L = sc.lums[ob.luminary];
                    const float cos_l = fabsf(dot3(d, h.ng));
                    if (L.area > 0.0f && cos_l > 0.0f) {
                        const float dist = len3(h.p - o);
                        const float ps = L.pdf_select > 0.0f ? L.pdf_select : 1.0f;
                        const float p_nee = ps * (dist * dist) / (L.area * cos_l);
                        mis = power_heuristic(bsdf_pdf, p_nee);
                    }
                }
                for (int w = 0; w < kHwss; ++w)
                    S.res[w * S.cap + slot] += thr[w] * (emitted(sc, mat, -d, h.ng, lam[w]) * mis);
            }
        }

        // Next-event estimation, one light sample per wavelength
        // (PathTracingIntegrator.cpp:4521-4552 -> LightSampler::EvaluateDirectLightingNM,
        // Lights/LightSampler.cpp:1987-2269).
        for (int w = 0; w < kHwss; ++w) {
            bool has = false;
            float v = 0.0f, stm = 0.0f;
            f3 sorg = mk3(0, 0, 0), to = mk3(0, 0, 0);
            if (valid && sc.n_lums > 0) {
                const unsigned dim = base_dim + 4u * (unsigned)w;
                const float usel = rng_u(R->seed, path, dim);
                int li = 0;
                while (li < sc.n_lums - 1 && !(usel < sc.lums[li].cdf_select)) ++li;
                if (li != ob.luminary) {  // never the shading object itself (:2061)
                    const DevLuminary L = sc.lums[li];
                    const float ux = rng_u(R->seed, path, dim + 1), uy = rng_u(R->seed, path, dim + 2),
                                uz = rng_u(R->seed, path, dim + 3);
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
                        float contrib = le * cos_s * geom * lambert_value(sc, h, mat.refl_spd, to, lam[w]);
                        const float p_light = L.pdf_select * (dist * dist) / (L.area * cos_l);
                        const float p_bsdf = lambert_pdf(h, to);
                        if (p_bsdf > 0.0f) contrib *= power_heuristic(p_light, p_bsdf);
                        contrib /= L.pdf_select;
                        v = thr[w] * contrib;
                        if (v > 0.0f) {
                            has = true;
                            sorg = offset_ray(h.p, dot3(to, h.ng) > 0.0f ? h.ng : -h.ng);
                            stm = dist - 0.001f;  // LightSampler.cpp:2206
                        }
                    }
                }
            }
            const unsigned si = wave_push(cnt_sh, has);
            if (has) {
                S.sox[si] = sorg.x;
                S.soy[si] = sorg.y;
                S.soz[si] = sorg.z;
                S.sdx[si] = to.x;
                S.sdy[si] = to.y;
                S.sdz[si] = to.z;
                S.stmax[si] = stm;
                S.sval[si] = v;
                S.sslot[si] = slot * 4u + (unsigned)w;
            }
        }

        // LambertianSPF::ScatterNM at the hero wavelength (Materials/LambertianSPF.cpp:90-131),
        // companions, Russian roulette, bounce caps.
        bool cont = false;
        f3 new_o = mk3(0, 0, 0), nd = mk3(0, 0, 0);
        if (valid) {
            const float ua = rng_u(R->seed, path, base_dim + 16), ub = rng_u(R->seed, path, base_dim + 17);
            Onb fr = h.frame;
            const bool back = dot3(h.dir, h.frame.w) > kNearZeroF;
            if (back) {
                fr.w = -fr.w;
                fr.u = -fr.u;
            }
            const float phi = kTwoPi * ua, st = sqrtf(ub), ct = sqrtf(1.0f - ub);
            nd = onb_xform(fr, mk3(cosf(phi) * st, sinf(phi) * st, ct));
            const float pdf = fabsf(dot3(nd, h.frame.w)) * kInvPi;
            const f3 neff = back ? -h.frame.w : h.frame.w;
            const f3 g = geom_against_ray(h, neff);
            if (dot3(nd, g) > 0.0f) {
                float sc4[kHwss];
                sc4[0] = spd_value(sc, mat.refl_spd, lam[0]);
                for (int w = 1; w < kHwss; ++w) {
                    // Companion weight: BSDF value * |cos| / pdf (:4693-4716).
                    float cw = lambert_value(sc, h, mat.refl_spd, nd, lam[w]) * fabsf(dot3(nd, h.n));
                    if (pdf > 0.0f) cw /= pdf;
                    sc4[w] = cw > 0.0f ? cw : 0.0f;
                }
                // Russian roulette over the max across wavelengths (:4721-4768,
                // PathTransportUtilities.h:86-110).
                float mprev = fabsf(thr[0]), mcurr = fabsf(thr[0] * sc4[0]);
                for (int w = 1; w < kHwss; ++w) {
                    mprev = fmaxf(mprev, fabsf(thr[w]));
                    mcurr = fmaxf(mcurr, fabsf(thr[w] * sc4[w]));
                }
                bool go = mcurr > kNearZeroF;
                if (go && depth >= R->rr_min_depth) {
                    const float p = fminf(1.0f, mcurr / fmaxf(mprev, R->rr_threshold));
                    if (rng_u(R->seed, path, base_dim + 19) >= p) {
                        go = false;
                    } else if (p > 0.0f && p < 1.0f) {
                        const float inv = 1.0f / p;
                        for (int w = 0; w < kHwss; ++w) sc4[w] *= inv;
                    }
                }
                // Bounce caps (PathTransportUtilities.h:170-210).
                const unsigned nb = S.bounces[slot] + 1u;
                if (go && (int)nb > R->max_diffuse) go = false;
                if (go) {
                    // Runaway guard for the next bounce (:4096-4110).
                    float mx = 0.0f;
                    bool bad = false;
                    for (int w = 0; w < kHwss; ++w) {
                        const float t = thr[w] * sc4[w];
                        if (!(fabsf(t) <= 3.0e38f)) bad = true;
                        mx = fmaxf(mx, fabsf(t));
                        S.thr[w * S.cap + slot] = t;
                    }
                    if (!bad && mx <= 1e6f) {
                        S.pdf[slot] = pdf;
                        S.bounces[slot] = nb;
                        new_o = offset_ray(h.p, dot3(nd, h.ng) > 0.0f ? h.ng : -h.ng);
                        cont = true;
                    }
                }
            }
        }
        const unsigned qi = wave_push(cnt_next, cont);
        if (cont) {
            qout[qi] = slot;
            S.ox[slot] = new_o.x;
            S.oy[slot] = new_o.y;
            S.oz[slot] = new_o.z;
            S.dx[slot] = nd.x;
            S.dy[slot] = nd.y;
            S.dz[slot] = nd.z;
        }
    }
}

// Shadow rays (LightSampler::ShadowOccludedNM -> ObjectManager::IntersectShadowRay);
// adds each unoccluded contribution to its path and wavelength.  There is one
// record per (path, wavelength) and bounce, so the additions never collide.
template <int STACK>
__global__ __launch_bounds__(kTraceBlock) void k_connect(const ReqInfo* __restrict__ reqs, Paths S, int depth)
{
    __shared__ uint2 lds[STACK > 0 ? STACK * kTraceBlock : 1];
    const unsigned n = S.cnt_sh[depth];
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const unsigned code = S.sslot[i];
        const unsigned slot = code >> 2, w = code & 3u;
        const unsigned rq = S.req[slot];
        const NodeSlot* nodes = reqs[rq].sc.nodes;
        const Tri48* tris = reqs[rq].sc.tris;
        RayState r;
        init_ray(r, S.sox[i], S.soy[i], S.soz[i], S.sdx[i], S.sdy[i], S.sdz[i]);
        const double how_far2 = scale_how_far(r, (double)S.stmax[i]);
        unsigned nn = 0, nt = 0;
        bool occ;
        if (STACK > 0) {
            LdsStack<STACK> st{lds + threadIdx.x, kTraceBlock, 0};
            occ = trace_any<false>(r, nodes, tris, st, how_far2, nn, nt);
        } else {
            PrivStack st;
            st.sp = 0;
            occ = trace_any<false>(r, nodes, tris, st, how_far2, nn, nt);
        }
        if (!occ) S.res[w * S.cap + slot] += S.sval[i];
    }
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
        const PassEntry E = ent[find_entry<&PassEntry::film0>(ent, n_ent, i)];
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
        for (int j = 0; j < E.ns; ++j) {
            const int slot = E.slot0 + local * E.ns + j;
            float lam[kHwss];
            wavelengths(S.lambda0[slot], R.lambda_begin, R.lambda_end, lam);
            float X = 0.0f, Y = 0.0f, Z = 0.0f;
            for (int w = 0; w < kHwss; ++w) {
                float x, y, z;
                if (xyz_from_nm(lam[w], x, y, z)) {
                    const float r = S.res[w * S.cap + slot];
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
// `base` (NULL: sizes only); returns the size.
size_t carve(char* base, int64_t cap, int passes, Paths* P, unsigned** counters)
{
    size_t off = 0;
    auto take = [&](size_t bytes) -> char* {
        char* p = base ? base + off : nullptr;
        off += align_up(bytes);
        return p;
    };
    const size_t c = (size_t)cap, c4 = 4 * c;
    Paths S{};
    S.cap = (int)cap;
    S.ox = (float*)take(c * 4);
    S.oy = (float*)take(c * 4);
    S.oz = (float*)take(c * 4);
    S.dx = (float*)take(c * 4);
    S.dy = (float*)take(c * 4);
    S.dz = (float*)take(c * 4);
    S.thr = (float*)take(c4 * 4);
    S.res = (float*)take(c4 * 4);
    S.lambda0 = (float*)take(c * 4);
    S.pdf = (float*)take(c * 4);
    S.path_id = (unsigned*)take(c * 4);
    S.req = (unsigned*)take(c * 4);
    S.bounces = (unsigned*)take(c * 4);
    S.hit_leaf = (unsigned*)take(c * 4);
    S.hit_t = (float*)take(c * 4);
    S.queue[0] = (unsigned*)take(c * 4);
    S.queue[1] = (unsigned*)take(c * 4);
    S.sox = (float*)take(c4 * 4);
    S.soy = (float*)take(c4 * 4);
    S.soz = (float*)take(c4 * 4);
    S.sdx = (float*)take(c4 * 4);
    S.sdy = (float*)take(c4 * 4);
    S.sdz = (float*)take(c4 * 4);
    S.stmax = (float*)take(c4 * 4);
    S.sval = (float*)take(c4 * 4);
    S.sslot = (unsigned*)take(c4 * 4);
    unsigned* cnt = (unsigned*)take(sizeof(unsigned) * (size_t)kCounters * (size_t)std::max(1, passes));
    if (P) *P = S;
    if (counters) *counters = cnt;
    return off;
}

unsigned grid_for(int64_t n, int block, unsigned max_blocks)
{
    const int64_t g = (n + block - 1) / block;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, max_blocks));
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

}  // namespace

// ---------------------------------------------------------------------------
// Batches
// ---------------------------------------------------------------------------

struct rise_hip_batch {
    int device = 0;
    int n_req = 0;
    int bounces = 1;                   // bounce iterations per pass
    int stack = 0;                     // LDS stack entries per lane (0: private array)
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
rise_hip_status plan(rise_hip_batch* b, const rise_hip_render_request* reqs, size_t n)
{
    if (n == 0 || !reqs) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "empty batch");
    if (n > (size_t)0x7fffffff) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "too many requests");
    b->requests.clear();
    std::vector<ReqInfo> info(n);
    std::vector<std::vector<PassEntry>> chunks(n);
    int64_t total_paths = 0, max_chunk = 0;
    int bounces = 1, depth = 1;
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
        const int hits = 1 + D.max_diffuse_bounce + (sc->has_glossy ? D.max_glossy_bounce : 0) +
                         (sc->has_transmission ? D.max_transmission_bounce : 0);
        bounces = std::max(bounces, std::min(kMaxPathDepth, hits));
        depth = std::max(depth, reqs[k].scene->view.depth);
        film_floats += 12 * (size_t)req_px[k];
    }
    const int64_t cap = std::max<int64_t>(max_chunk, std::min<int64_t>(kMaxPaths, total_paths));

    // Passes: each takes the next chunk of every request, in request order, while they fit.
    std::vector<PassEntry> entries;
    std::vector<int> pass_first(1, 0), pass_paths, pass_film;
    std::vector<size_t> next(n, 0);
    size_t left = 0;
    for (const auto& c : chunks) left += c.size();
    while (left > 0) {
        int64_t paths = 0;
        int64_t film = 0;
        const size_t first = entries.size();
        for (size_t k = 0; k < n; ++k) {
            if (next[k] >= chunks[k].size()) continue;
            PassEntry e = chunks[k][next[k]];
            const int64_t pp = (int64_t)e.np * e.ns;
            if (paths + pp > cap) continue;
            e.slot0 = (int)paths;
            e.film0 = (int)film;
            paths += pp;
            film += e.np;
            entries.push_back(e);
            ++next[k];
            --left;
        }
        if (entries.size() == first) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "internal error: empty pass");
        pass_first.push_back((int)entries.size());
        pass_paths.push_back((int)paths);
        pass_film.push_back((int)film);
    }
    const int passes = (int)pass_paths.size();

    // Device buffers.
    hipError_t e = hipGetDevice(&b->device);
    if (e != hipSuccess) return fail_hip(e, "hipGetDevice");
    const size_t need = carve(nullptr, cap, passes, nullptr, nullptr);
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
    carve((char*)b->work, cap, passes, &b->S, &b->counters);
    b->film_floats = film_floats;
    b->entries.swap(entries);
    b->pass_first.swap(pass_first);
    b->pass_paths.swap(pass_paths);
    b->pass_film.swap(pass_film);
    b->req_px.swap(req_px);
    const int need_stack = depth - 1;
    b->stack = need_stack <= 8 ? 8 : (need_stack <= 16 ? 16 : 0);
    b->bounces = bounces;
    b->n_req = (int)n;
    b->requests.assign(reqs, reqs + n);
    b->launched = false;
    return RISE_HIP_OK;
}

// Kernel launches of a planned batch (no host synchronization, no copies).
hipError_t launch(rise_hip_batch* b, hipStream_t stream)
{
    const unsigned gt = grid_for(b->cap, kTraceBlock, 4096);
    const unsigned gs = grid_for(b->cap, kShadeBlock, 2048);
    hipLaunchKernelGGL(k_fill, dim3(grid_for((int64_t)b->film_floats, 256, 4096)), dim3(256), 0, stream, b->d_film,
                       b->film_floats);
    const int passes = (int)b->pass_paths.size();
    for (int p = 0; p < passes; ++p) {
        Paths S = b->S;
        S.cnt_ext = b->counters + (size_t)p * kCounters;
        S.cnt_sh = S.cnt_ext + (kMaxPathDepth + 1);
        const PassEntry* ent = b->d_entries + b->pass_first[p];
        const int n_ent = b->pass_first[p + 1] - b->pass_first[p];
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(256), 0, stream, S.cnt_ext, kCounters);
        hipLaunchKernelGGL(k_generate, dim3(grid_for(b->pass_paths[p], 256, 4096)), dim3(256), 0, stream, b->d_info,
                           ent, n_ent, b->pass_paths[p], S);
        for (int d = 0; d < b->bounces; ++d) {
            switch (b->stack) {
            case 8: hipLaunchKernelGGL(k_extend<8>, dim3(gt), dim3(kTraceBlock), 0, stream, b->d_info, S, d); break;
            case 16: hipLaunchKernelGGL(k_extend<16>, dim3(gt), dim3(kTraceBlock), 0, stream, b->d_info, S, d); break;
            default: hipLaunchKernelGGL(k_extend<0>, dim3(gt), dim3(kTraceBlock), 0, stream, b->d_info, S, d); break;
            }
            hipLaunchKernelGGL(k_shade, dim3(gs), dim3(kShadeBlock), 0, stream, b->d_info, S, d);
            switch (b->stack) {
            case 8: hipLaunchKernelGGL(k_connect<8>, dim3(gt), dim3(kTraceBlock), 0, stream, b->d_info, S, d); break;
            case 16: hipLaunchKernelGGL(k_connect<16>, dim3(gt), dim3(kTraceBlock), 0, stream, b->d_info, S, d); break;
            default: hipLaunchKernelGGL(k_connect<0>, dim3(gt), dim3(kTraceBlock), 0, stream, b->d_info, S, d); break;
            }
        }
        hipLaunchKernelGGL(k_film, dim3(grid_for(b->pass_film[p], 256, 4096)), dim3(256), 0, stream, b->d_info, ent,
                           n_ent, b->pass_film[p], S);
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

extern "C" rise_hip_status rise_hip_batch_create(const rise_hip_render_request* requests, size_t n_requests,
                                                 rise_hip_batch** out_batch)
{
    clear_error();
    if (!out_batch) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "out_batch is NULL");
    *out_batch = nullptr;
    rise_hip_batch* b = new (std::nothrow) rise_hip_batch;
    if (!b) return fail(RISE_HIP_ERR_OUT_OF_MEMORY, "host allocation failed");
    const rise_hip_status st = plan(b, requests, n_requests);
    if (st != RISE_HIP_OK) {
        free_batch(b);
        return st;
    }
    *out_batch = b;
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
        const rise_hip_status st = plan(b, &q, 1);
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
