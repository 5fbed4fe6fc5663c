// pt_render.hip -- wavefront spectral path tracer and its C ABI
// (include/rise_hip_render.h).
//
// Per pass (a block of pixels x a block of samples, sized to the path
// capacity), kernels run in this order on the caller's stream:
//   generate            camera rays, hero wavelength, path state
//   for each bounce d:  extend(d)   closest hit for the active queue
//                       shade(d)    emission + MIS, one NEE sample per
//                                   wavelength (shadow queue), hero BSDF
//                                   sample, companions, RR, bounce caps
//                                   (continuation queue)
//                       connect(d)  any-hit for the shadow queue; adds the
//                                   unoccluded contributions
//   film                per-pixel FP32 Kahan sums over the pass's samples
// and one resolve at the end.  Queues are compacted with a wave-wide ballot
// and one atomic per wave.  Every count lives in device memory and every
// kernel is a grid-stride loop over it, so a render is a fixed launch
// sequence with no host synchronization.
//
// The bounce loop follows PathTracingIntegrator::IntegrateFromHitHWSS
// (Shaders/PathTracingIntegrator.cpp:3939-4822) for Lambertian materials and
// Lambertian area luminaries; each step cites the lines it follows.
#include "rise_hip_render.h"

#include "api_common.h"
#include "bvh_build.h"
#include "cie_table.h"
#include "pt_device.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
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
constexpr int kLdsStack = 16;
constexpr int64_t kMaxPaths = int64_t(1) << 21;  // paths in flight per pass

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

struct PtParams {
    int width, height;
    int spp_total;
    int p0, np;        // pixel range of this pass
    int s0, ns;        // sample range of this pass (absolute sample index s0)
    int n_paths;       // np * ns
    int cap;           // path capacity (stride of the SoA arrays)
    float lambda_begin, lambda_end;
    float xyz_scale;   // mYNormalization / N (PathTracingSpectralRasterizer.cpp:168-173)
    int max_diffuse, max_glossy, max_transmission;
    int rr_min_depth;
    float rr_threshold;
    unsigned long long seed;
    float cam_pos[3], cam_f[3], cam_r[3], cam_u[3];
    float cam_h;       // 2 tan(fov / 2)
};

// Path state, structure of arrays, `cap` entries each.
struct Paths {
    float *ox, *oy, *oz, *dx, *dy, *dz;
    float* thr;      // 4 * cap, wavelength-major
    float* res;      // 4 * cap
    float* lambda0;
    float* pdf;      // bsdfPdf of the ray being traced (0: camera or delta)
    unsigned* path_id;
    unsigned* bounces;   // diffuse count (PropagateBounceLimits)
    unsigned* hit_leaf;
    float* hit_t;
    unsigned* queue[2];
    // shadow records, 4 * cap
    float *sox, *soy, *soz, *sdx, *sdy, *sdz, *stmax, *sval;
    unsigned* sslot;     // slot * 4 + wavelength
    unsigned* cnt_ext;   // [kMaxPathDepth + 1]
    unsigned* cnt_sh;    // [kMaxPathDepth + 1]
    // film, 3 * width * height each
    float *fsum, *fcomp, *fsq, *fsqc;
};

__device__ inline unsigned lane_id() { return __lane_id(); }

// Wave-wide compaction: one atomic per wave, ranks from the ballot.
__device__ inline unsigned wave_push(unsigned* counter, bool pred)
{
    const unsigned long long m = __ballot(pred);
    if (m == 0ull) return 0xffffffffu;
    const unsigned rank =
        __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
    const int leader = __ffsll((unsigned long long)m) - 1;
    unsigned base = 0;
    if ((int)lane_id() == leader) base = atomicAdd(counter, (unsigned)__popcll(m));
    base = __shfl(base, leader);
    return pred ? base + rank : 0xffffffffu;
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

// Camera sample: PathTracingSpectralRasterizer::IntegratePixel (box filter,
// :310-338) and IntegratePixelSpectral (:121-150).
__global__ void k_generate(PtParams P, Paths S)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) S.cnt_ext[0] = (unsigned)P.n_paths;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < P.n_paths; i += gridDim.x * blockDim.x) {
        const int pix = P.p0 + i / P.ns;
        const int s = P.s0 + i % P.ns;
        const int px = pix % P.width, py = pix / P.width;
        const unsigned path = (unsigned)((unsigned long long)pix * (unsigned long long)P.spp_total + (unsigned)s);
        const float jx = rng_u(P.seed, path, 0), jy = rng_u(P.seed, path, 1), ul = rng_u(P.seed, path, 2);
        const float sx = (float)px + jx, sy = (float)py + jy;
        const float cx = (sx - 0.5f * (float)P.width) * P.cam_h / (float)P.height;
        const float cy = -(sy - 0.5f * (float)P.height) * P.cam_h / (float)P.height;
        const f3 f = mk3(P.cam_f[0], P.cam_f[1], P.cam_f[2]);
        const f3 r = mk3(P.cam_r[0], P.cam_r[1], P.cam_r[2]);
        const f3 u = mk3(P.cam_u[0], P.cam_u[1], P.cam_u[2]);
        const f3 d = norm3(f + r * cx + u * cy);
        S.ox[i] = P.cam_pos[0];
        S.oy[i] = P.cam_pos[1];
        S.oz[i] = P.cam_pos[2];
        S.dx[i] = d.x;
        S.dy[i] = d.y;
        S.dz[i] = d.z;
        for (int w = 0; w < kHwss; ++w) {
            S.thr[w * P.cap + i] = 1.0f;
            S.res[w * P.cap + i] = 0.0f;
        }
        S.lambda0[i] = P.lambda_begin + ul * (P.lambda_end - P.lambda_begin);
        S.pdf[i] = 0.0f;
        S.path_id[i] = path;
        S.bounces[i] = 0u;
        S.queue[0][i] = (unsigned)i;
    }
}

// Closest hit for the active queue (ObjectManager::IntersectRay).
template <bool LDS>
__global__ __launch_bounds__(kTraceBlock) void k_extend(PtSceneView sc, PtParams P, Paths S, int depth)
{
    __shared__ uint2 lds[LDS ? kLdsStack * kTraceBlock : 1];
    const unsigned n = S.cnt_ext[depth];
    const unsigned* q = S.queue[depth & 1];
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const unsigned slot = q[i];
        RayState r;
        init_ray(r, S.ox[slot], S.oy[slot], S.oz[slot], S.dx[slot], S.dy[slot], S.dz[slot]);
        double best;
        unsigned leaf, nn = 0, nt = 0;
        int id;
        if (LDS) {
            LdsStack<kLdsStack> st{lds + threadIdx.x, kTraceBlock, 0};
            id = trace_closest<false>(r, sc.nodes, sc.tris, st, best, leaf, nn, nt);
        } else {
            PrivStack st;
            st.sp = 0;
            id = trace_closest<false>(r, sc.nodes, sc.tris, st, best, leaf, nn, nt);
        }
        S.hit_leaf[slot] = id >= 0 ? leaf : 0xffffffffu;
        S.hit_t[slot] = id >= 0 ? (float)best : 0.0f;
    }
}

// One bounce of IntegrateFromHitHWSS for the active queue.  Every lane of a
// wave runs the same loop trips (the queue bound is uniform per block) so the
// compaction ballots see all lanes; each shadow record is written as soon as
// it is made, which keeps one record live at a time.
__global__ __launch_bounds__(kShadeBlock) void k_shade(PtSceneView sc, PtParams P, Paths S, int depth)
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
            for (int w = 0; w < kHwss; ++w) thr[w] = S.thr[w * P.cap + slot];
            wavelengths(S.lambda0[slot], P.lambda_begin, P.lambda_end, lam);
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
                    S.res[w * P.cap + slot] += thr[w] * (emitted(sc, mat, -d, h.ng, lam[w]) * mis);
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
                const float usel = rng_u(P.seed, path, dim);
                int li = 0;
                while (li < sc.n_lums - 1 && !(usel < sc.lums[li].cdf_select)) ++li;
                if (li != ob.luminary) {  // never the shading object itself (:2061)
                    const DevLuminary L = sc.lums[li];
                    const float ux = rng_u(P.seed, path, dim + 1), uy = rng_u(P.seed, path, dim + 2),
                                uz = rng_u(P.seed, path, dim + 3);
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
            const float ua = rng_u(P.seed, path, base_dim + 16), ub = rng_u(P.seed, path, base_dim + 17);
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
                if (go && depth >= P.rr_min_depth) {
                    const float p = fminf(1.0f, mcurr / fmaxf(mprev, P.rr_threshold));
                    if (rng_u(P.seed, path, base_dim + 19) >= p) {
                        go = false;
                    } else if (p > 0.0f && p < 1.0f) {
                        const float inv = 1.0f / p;
                        for (int w = 0; w < kHwss; ++w) sc4[w] *= inv;
                    }
                }
                // Bounce caps (PathTransportUtilities.h:170-210).
                const unsigned nb = S.bounces[slot] + 1u;
                if (go && (int)nb > P.max_diffuse) go = false;
                if (go) {
                    // Runaway guard for the next bounce (:4096-4110).
                    float mx = 0.0f;
                    bool bad = false;
                    for (int w = 0; w < kHwss; ++w) {
                        const float t = thr[w] * sc4[w];
                        if (!(fabsf(t) <= 3.0e38f)) bad = true;
                        mx = fmaxf(mx, fabsf(t));
                        S.thr[w * P.cap + slot] = t;
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
// adds each unoccluded contribution to its path and wavelength.
template <bool LDS>
__global__ __launch_bounds__(kTraceBlock) void k_connect(PtSceneView sc, PtParams P, Paths S, int depth)
{
    __shared__ uint2 lds[LDS ? kLdsStack * kTraceBlock : 1];
    const unsigned n = S.cnt_sh[depth];
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        RayState r;
        init_ray(r, S.sox[i], S.soy[i], S.soz[i], S.sdx[i], S.sdy[i], S.sdz[i]);
        const double how_far2 = scale_how_far(r, (double)S.stmax[i]);
        unsigned nn = 0, nt = 0;
        bool occ;
        if (LDS) {
            LdsStack<kLdsStack> st{lds + threadIdx.x, kTraceBlock, 0};
            occ = trace_any<false>(r, sc.nodes, sc.tris, st, how_far2, nn, nt);
        } else {
            PrivStack st;
            st.sp = 0;
            occ = trace_any<false>(r, sc.nodes, sc.tris, st, how_far2, nn, nt);
        }
        if (!occ) {
            const unsigned code = S.sslot[i];
            const unsigned slot = code >> 2, w = code & 3u;
            S.res[w * P.cap + slot] += S.sval[i];
        }
    }
}

// Per-pixel FP32 sums with a Kahan compensation term, samples in index
// order: deterministic for a given pass layout.
__device__ inline void kahan_add(float& sum, float& comp, float x)
{
    const float y = x - comp;
    const float t = sum + y;
    comp = (t - sum) - y;
    sum = t;
}

__global__ void k_film(PtParams P, Paths S)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < P.np; i += gridDim.x * blockDim.x) {
        const int pix = P.p0 + i;
        float s[3], c[3], q[3], qc[3];
        for (int k = 0; k < 3; ++k) {
            s[k] = S.fsum[3 * pix + k];
            c[k] = S.fcomp[3 * pix + k];
            q[k] = S.fsq[3 * pix + k];
            qc[k] = S.fsqc[3 * pix + k];
        }
        for (int j = 0; j < P.ns; ++j) {
            const int slot = i * P.ns + j;
            float lam[kHwss];
            wavelengths(S.lambda0[slot], P.lambda_begin, P.lambda_end, lam);
            float X = 0.0f, Y = 0.0f, Z = 0.0f;
            for (int w = 0; w < kHwss; ++w) {
                float x, y, z;
                if (xyz_from_nm(lam[w], x, y, z)) {
                    const float r = S.res[w * P.cap + slot];
                    X += x * r;
                    Y += y * r;
                    Z += z * r;
                }
            }
            const float v[3] = {X * P.xyz_scale, Y * P.xyz_scale, Z * P.xyz_scale};
            for (int k = 0; k < 3; ++k) {
                kahan_add(s[k], c[k], v[k]);
                kahan_add(q[k], qc[k], v[k] * v[k]);
            }
        }
        for (int k = 0; k < 3; ++k) {
            S.fsum[3 * pix + k] = s[k];
            S.fcomp[3 * pix + k] = c[k];
            S.fsq[3 * pix + k] = q[k];
            S.fsqc[3 * pix + k] = qc[k];
        }
    }
}

__global__ void k_resolve(int npx, float inv_n, const float* fsum, const float* fsq, float* out, float* out_sq)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < 3 * npx; i += gridDim.x * blockDim.x) {
        out[i] = fsum[i] * inv_n;
        if (out_sq) out_sq[i] = fsq[i] * inv_n;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Scene
// ---------------------------------------------------------------------------

struct rise_hip_pt_scene {
    int device = 0;
    std::vector<void*> allocs;
    PtSceneView view{};
    int64_t n_tris = 0;
    // work buffers, reused between renders
    void* work = nullptr;
    size_t work_bytes = 0;
    int64_t cap = 0;
    int64_t film_px = 0;
};

namespace {

template <class T>
rise_hip_status upload(rise_hip_pt_scene* s, const std::vector<T>& v, const T** out)
{
    const size_t bytes = std::max<size_t>(1, v.size()) * sizeof(T);
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return fail_hip(e, "hipMalloc (scene)");
    s->allocs.push_back(p);
    if (!v.empty()) {
        e = hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
        if (e != hipSuccess) return fail_hip(e, "hipMemcpy (scene)");
    }
    *out = static_cast<const T*>(p);
    return RISE_HIP_OK;
}

void free_scene(rise_hip_pt_scene* s)
{
    for (void* p : s->allocs) (void)hipFree(p);
    if (s->work) (void)hipFree(s->work);
    delete s;
}

}  // namespace

extern "C" rise_hip_status rise_hip_pt_scene_create(const rise_hip_object_desc* objects, size_t n_objects,
                                                    const rise_hip_material_desc* materials, size_t n_materials,
                                                    const rise_hip_spd_desc* spds, size_t n_spds,
                                                    rise_hip_pt_scene** out_scene)
{
    clear_error();
    if (!out_scene) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "out_scene is NULL");
    *out_scene = nullptr;
    if ((n_objects && !objects) || (n_materials && !materials) || (n_spds && !spds))
        return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL array with a nonzero count");
    if (n_objects > 0x7fffffff || n_materials > 0x7fffffff || n_spds > 0x7fffffff)
        return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "too many objects, materials or spectra");
    // Spectra.
    std::vector<DevSpd> dspd(n_spds);
    std::vector<float> amps;
    for (size_t i = 0; i < n_spds; ++i) {
        const rise_hip_spd_desc& d = spds[i];
        if (d.n <= 0 || !d.amplitudes || !(d.lambda_end > d.lambda_begin))
            return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "spectrum " + std::to_string(i) + " is empty or inverted");
        dspd[i].begin = d.lambda_begin;
        dspd[i].end = d.lambda_end;
        dspd[i].n = d.n;
        dspd[i].inv_delta = (float)((double)d.n / ((double)d.lambda_end - (double)d.lambda_begin));
        dspd[i].offset = (int)amps.size();
        amps.insert(amps.end(), d.amplitudes, d.amplitudes + d.n);
    }
    // Materials.
    std::vector<DevMaterial> dmat(n_materials);
    for (size_t i = 0; i < n_materials; ++i) {
        const rise_hip_material_desc& m = materials[i];
        if (m.type != RISE_HIP_MATERIAL_LAMBERTIAN)
            return fail(RISE_HIP_ERR_NOT_SUPPORTED, "material " + std::to_string(i) + ": unsupported type");
        if (m.reflectance_spd < 0 || (size_t)m.reflectance_spd >= n_spds || m.emission_spd >= (int)n_spds ||
            m.emission_spd < -1)
            return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "material " + std::to_string(i) + ": spectrum index out of range");
        dmat[i] = DevMaterial{m.type, m.reflectance_spd, m.emission_spd, m.emission_scale};
    }
    // Objects -> one triangle list.
    int64_t total = 0;
    for (size_t o = 0; o < n_objects; ++o) {
        if (objects[o].n_tris < 0 || (objects[o].n_tris > 0 && !objects[o].vertices))
            return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "object " + std::to_string(o) + ": bad triangle buffer");
        if (objects[o].material < 0 || (size_t)objects[o].material >= n_materials)
            return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "object " + std::to_string(o) + ": material out of range");
        total += objects[o].n_tris;
    }
    if (total > 0x7fffffff) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "more than 2^31-1 triangles");
    std::vector<float> verts((size_t)total * 9);
    std::vector<int32_t> tri_obj((size_t)total);
    {
        int64_t off = 0;
        for (size_t o = 0; o < n_objects; ++o) {
            const int64_t k = objects[o].n_tris;
            if (k) std::memcpy(verts.data() + 9 * off, objects[o].vertices, (size_t)k * 9 * sizeof(float));
            for (int64_t i = 0; i < k; ++i) tri_obj[off + i] = (int32_t)o;
            off += k;
        }
    }
    WideBVH bvh;
    std::string err;
    try {
        if (!build_wide_bvh(verts.data(), total, nullptr, BuildParams(), bvh, err))
            return fail(RISE_HIP_ERR_BUILD, "BVH build failed: " + err);
    } catch (const std::bad_alloc&) {
        return fail(RISE_HIP_ERR_OUT_OF_MEMORY, "host allocation failed during BVH build");
    }
    // Object id in v1.w; inverse map global id -> leaf index.
    std::vector<uint32_t> leaf_of((size_t)total);
    for (size_t l = 0; l < bvh.tris.size(); ++l) {
        int32_t id;
        std::memcpy(&id, &bvh.tris[l].v0[3], 4);
        const int32_t ob = tri_obj[(size_t)id];
        std::memcpy(&bvh.tris[l].v1[3], &ob, 4);
        leaf_of[(size_t)id] = (uint32_t)l;
    }
    // Luminaries (LuminaryManager + LightSampler::Prepare): emitting objects,
    // area CDF over their triangles, selection weight = mean exitance * scale * area.
    std::vector<DevObject> dobj(n_objects);
    std::vector<DevLuminary> lums;
    std::vector<uint32_t> lum_tris;
    std::vector<float> lum_cdf;
    std::vector<double> weights;
    {
        int64_t off = 0;
        for (size_t o = 0; o < n_objects; ++o) {
            const int64_t k = objects[o].n_tris;
            dobj[o] = DevObject{objects[o].material, objects[o].double_sided ? 1 : 0, -1, 0};
            const DevMaterial& m = dmat[objects[o].material];
            if (m.emit_spd >= 0 && k > 0) {
                std::vector<double> areas((size_t)k);
                double tot = 0;
                for (int64_t i = 0; i < k; ++i) {
                    const float* f = verts.data() + 9 * (off + i);
                    const double ax = (double)f[3] - f[0], ay = (double)f[4] - f[1], az = (double)f[5] - f[2];
                    const double bx = (double)f[6] - f[0], by = (double)f[7] - f[1], bz = (double)f[8] - f[2];
                    const double cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
                    areas[(size_t)i] = std::sqrt(cx * cx + cy * cy + cz * cz) * 0.5;
                    tot += areas[(size_t)i];
                }
                const DevSpd& e = dspd[(size_t)m.emit_spd];
                double mean = 0;
                for (int j = 0; j < e.n; ++j) mean += amps[(size_t)(e.offset + j)];
                mean /= e.n;
                const double wgt = mean * m.emit_scale * tot;
                if (tot > 0 && wgt > 0) {
                    DevLuminary L{};
                    L.object = (int)o;
                    L.tri_first = (int)lum_tris.size();
                    L.tri_count = (int)k;
                    L.area = (float)tot;
                    L.material = objects[o].material;
                    double acc = 0;
                    for (int64_t i = 0; i < k; ++i) {
                        acc += areas[(size_t)i] / tot;
                        lum_tris.push_back(leaf_of[(size_t)(off + i)]);
                        lum_cdf.push_back((float)acc);
                    }
                    dobj[o].luminary = (int)lums.size();
                    lums.push_back(L);
                    weights.push_back(wgt);
                }
            }
            off += k;
        }
        double wt = 0;
        for (double w : weights) wt += w;
        double acc = 0;
        for (size_t i = 0; i < lums.size(); ++i) {
            acc += weights[i];
            lums[i].pdf_select = (float)(weights[i] / wt);
            lums[i].cdf_select = (float)(acc / wt);
        }
        if (!lums.empty()) lums.back().cdf_select = 1.0f;
    }

    rise_hip_pt_scene* s = new (std::nothrow) rise_hip_pt_scene;
    if (!s) return fail(RISE_HIP_ERR_OUT_OF_MEMORY, "host allocation failed");
    hipError_t he = hipGetDevice(&s->device);
    if (he != hipSuccess) {
        delete s;
        return fail_hip(he, "hipGetDevice");
    }
    rise_hip_status st = RISE_HIP_OK;
    PtSceneView& v = s->view;
    if (st == RISE_HIP_OK) st = upload(s, bvh.nodes, &v.nodes);
    if (st == RISE_HIP_OK) st = upload(s, bvh.tris, &v.tris);
    if (st == RISE_HIP_OK) st = upload(s, dobj, &v.objects);
    if (st == RISE_HIP_OK) st = upload(s, dmat, &v.materials);
    if (st == RISE_HIP_OK) st = upload(s, dspd, &v.spds);
    if (st == RISE_HIP_OK) st = upload(s, amps, &v.spd_amp);
    if (st == RISE_HIP_OK) st = upload(s, lums, &v.lums);
    if (st == RISE_HIP_OK) st = upload(s, lum_tris, &v.lum_tris);
    if (st == RISE_HIP_OK) st = upload(s, lum_cdf, &v.lum_cdf);
    if (st != RISE_HIP_OK) {
        free_scene(s);
        return st;
    }
    v.depth = bvh.stats.depth;
    v.n_lums = (int)lums.size();
    s->n_tris = total;
    *out_scene = s;
    return RISE_HIP_OK;
}

extern "C" rise_hip_status rise_hip_pt_scene_release(rise_hip_pt_scene* scene)
{
    clear_error();
    if (scene) free_scene(scene);
    return RISE_HIP_OK;
}

// ---------------------------------------------------------------------------
// Render
// ---------------------------------------------------------------------------

namespace {

size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

// Carves the work buffer for `cap` paths and `npx` film pixels; returns the size.
size_t carve(char* base, int64_t cap, int64_t npx, Paths* P)
{
    size_t off = 0;
    auto take = [&](size_t bytes) -> char* {
        char* p = base ? base + off : nullptr;
        off += align_up(bytes);
        return p;
    };
    const size_t c = (size_t)cap, c4 = 4 * c, f = 3 * (size_t)npx;
    Paths S{};
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
    S.cnt_ext = (unsigned*)take(8 * (kMaxPathDepth + 1));  // extension counts, then shadow counts
    S.cnt_sh = S.cnt_ext ? S.cnt_ext + (kMaxPathDepth + 1) : nullptr;
    S.fsum = (float*)take(f * 4);
    S.fcomp = (float*)take(f * 4);
    S.fsq = (float*)take(f * 4);
    S.fsqc = (float*)take(f * 4);
    if (P) *P = S;
    return off;
}

unsigned grid_for(int64_t n, int block, unsigned max_blocks)
{
    const int64_t g = (n + block - 1) / block;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, max_blocks));
}

}  // namespace

extern "C" rise_hip_status rise_hip_render(rise_hip_pt_scene* scene, const rise_hip_render_desc* desc, float* film_xyz,
                                           float* film_xyz_sq, rise_hip_render_stats* stats, hipStream_t stream)
{
    clear_error();
    if (!scene || !desc || !film_xyz) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    const rise_hip_render_desc& D = *desc;
    if (D.width <= 0 || D.height <= 0 || D.spp <= 0 || D.spp_offset < 0 || D.spp_total < D.spp_offset + D.spp)
        return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "bad image size or sample range");
    const int64_t npx = (int64_t)D.width * D.height;
    if ((uint64_t)npx * (uint64_t)D.spp_total > 0xffffffffull)
        return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "width * height * spp_total exceeds 2^32 (path id space)");
    if (!(D.lambda_end > D.lambda_begin) || D.lambda_begin < 0)
        return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "bad wavelength range");
    if (D.max_diffuse_bounce < 0 || D.max_glossy_bounce < 0 || D.max_transmission_bounce < 0 || D.rr_min_depth < 0)
        return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "negative bounce cap or RR depth");
    if (!(D.camera.fov_deg > 0 && D.camera.fov_deg < 180)) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "bad field of view");

    // Pass layout: all samples of a block of pixels, or a block of samples of all pixels.
    const int64_t cap = std::min<int64_t>(kMaxPaths, npx * (int64_t)D.spp);
    int ns = (int)std::min<int64_t>(D.spp, std::max<int64_t>(1, cap / npx));
    int np = (int)std::min<int64_t>(npx, cap / ns);

    // Work buffers.
    Paths S{};
    const size_t need = carve(nullptr, cap, npx, nullptr);
    if (need > scene->work_bytes) {
        if (scene->work) (void)hipFree(scene->work);
        scene->work = nullptr;
        scene->work_bytes = 0;
        const hipError_t e = hipMalloc(&scene->work, need);
        if (e != hipSuccess) return fail_hip(e, "hipMalloc (render work buffers)");
        scene->work_bytes = need;
    }
    carve((char*)scene->work, cap, npx, &S);

    // Parameters.
    PtParams P{};
    P.width = D.width;
    P.height = D.height;
    P.spp_total = D.spp_total;
    P.cap = (int)cap;
    P.lambda_begin = D.lambda_begin;
    P.lambda_end = D.lambda_end;
    const double ky = cie_y_integral(D.lambda_begin, D.lambda_end);
    const double ynorm = ky > 1e-12 ? ((double)D.lambda_end - D.lambda_begin) / ky : 1.0;
    P.xyz_scale = (float)(ynorm / kHwss);
    P.max_diffuse = D.max_diffuse_bounce;
    P.max_glossy = D.max_glossy_bounce;
    P.max_transmission = D.max_transmission_bounce;
    P.rr_min_depth = D.rr_min_depth;
    P.rr_threshold = D.rr_threshold;
    P.seed = D.seed;
    {
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
            P.cam_pos[k] = D.camera.position[k];
            P.cam_f[k] = (float)f[k];
            P.cam_r[k] = (float)r[k];
            P.cam_u[k] = (float)u[k];
        }
        P.cam_h = (float)(2.0 * std::tan(D.camera.fov_deg * (M_PI / 180.0) * 0.5));
    }
    const int bounces = std::min(kMaxPathDepth, 1 + D.max_diffuse_bounce + D.max_glossy_bounce +
                                                    D.max_transmission_bounce);
    const bool lds = scene->view.depth - 1 <= kLdsStack;
    const unsigned gt = grid_for(cap, kTraceBlock, 4096);
    const unsigned gs = grid_for(cap, kShadeBlock, 2048);

    hipLaunchKernelGGL(k_fill, dim3(grid_for(12 * npx, 256, 4096)), dim3(256), 0, stream, S.fsum,
                       (size_t)(S.fsqc + 3 * npx - S.fsum));
    int64_t total_ext = 0, total_sh = 0, total_paths = 0, total_bounces = 0;
    std::vector<unsigned> hc(2 * (kMaxPathDepth + 1));
    for (int s0 = D.spp_offset; s0 < D.spp_offset + D.spp; s0 += ns) {
        const int ns_pass = std::min(ns, D.spp_offset + D.spp - s0);
        for (int p0 = 0; p0 < npx; p0 += np) {
            const int np_pass = (int)std::min<int64_t>(np, npx - p0);
            P.p0 = p0;
            P.np = np_pass;
            P.s0 = s0;
            P.ns = ns_pass;
            P.n_paths = np_pass * ns_pass;
            hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(256), 0, stream, S.cnt_ext, 2 * (kMaxPathDepth + 1));
            hipLaunchKernelGGL(k_generate, dim3(grid_for(P.n_paths, 256, 4096)), dim3(256), 0, stream, P, S);
            for (int d = 0; d < bounces; ++d) {
                if (lds)
                    hipLaunchKernelGGL(k_extend<true>, dim3(gt), dim3(kTraceBlock), 0, stream, scene->view, P, S, d);
                else
                    hipLaunchKernelGGL(k_extend<false>, dim3(gt), dim3(kTraceBlock), 0, stream, scene->view, P, S, d);
                hipLaunchKernelGGL(k_shade, dim3(gs), dim3(kShadeBlock), 0, stream, scene->view, P, S, d);
                if (lds)
                    hipLaunchKernelGGL(k_connect<true>, dim3(gt), dim3(kTraceBlock), 0, stream, scene->view, P, S, d);
                else
                    hipLaunchKernelGGL(k_connect<false>, dim3(gt), dim3(kTraceBlock), 0, stream, scene->view, P, S,
                                       d);
            }
            hipLaunchKernelGGL(k_film, dim3(grid_for(np_pass, 256, 4096)), dim3(256), 0, stream, P, S);
            if (stats) {
                hipError_t e = hipMemcpyAsync(hc.data(), S.cnt_ext, hc.size() * 4, hipMemcpyDeviceToHost, stream);
                if (e == hipSuccess) e = hipStreamSynchronize(stream);
                if (e != hipSuccess) return fail_hip(e, "render statistics");
                for (int d = 0; d < bounces; ++d) {
                    total_ext += hc[d];
                    total_sh += hc[kMaxPathDepth + 1 + d];
                    if (hc[d]) ++total_bounces;
                }
            }
            total_paths += P.n_paths;
        }
    }
    hipLaunchKernelGGL(k_resolve, dim3(grid_for(3 * npx, 256, 4096)), dim3(256), 0, stream, (int)npx,
                       1.0f / (float)D.spp, S.fsum, S.fsq, film_xyz, film_xyz_sq);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail_hip(e, "render launch");
    if (stats) {
        const hipError_t e2 = hipStreamSynchronize(stream);
        if (e2 != hipSuccess) return fail_hip(e2, "render");
        stats->extension_rays = total_ext;
        stats->shadow_rays = total_sh;
        stats->paths = total_paths;
        stats->bounces = total_bounces;
    }
    return RISE_HIP_OK;
}
