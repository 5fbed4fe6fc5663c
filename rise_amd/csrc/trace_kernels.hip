// trace_kernels.hip -- batch closest-hit and any-hit kernels over a ray SoA.
//
// These replace the per-ray IObjectManager::IntersectRay / IntersectShadowRay
// calls (Interfaces/IObjectManager.h:38-68, Managers/ObjectManager.cpp:270-373)
// with one launch over a whole batch.  One lane per ray; the node-group stack
// lives in LDS (16 entries per lane) when the tree is shallow enough, else in
// a private array.
#include "rise_hip.h"
#include "trace_core.h"
#include "trace_kernels.h"
#include "wave_fetch.h"

#include <algorithm>
#include <mutex>

namespace rise_hip {

static_assert(kArithIeee == RISE_HIP_ARITH_IEEE && kArithX86Fma == RISE_HIP_ARITH_X86_FMA,
              "arithmetic models share the ABI's values");

namespace {

#ifndef RISE_PERSISTENT
#define RISE_PERSISTENT 1
#endif
#ifndef RISE_TRACE_BLOCK
#define RISE_TRACE_BLOCK 128
#endif
constexpr int kBlock = RISE_TRACE_BLOCK;

__device__ inline void load_ray(const float* __restrict__ rays, int64_t n, int64_t i, float& ox, float& oy, float& oz,
                                float& dx, float& dy, float& dz)
{
    ox = rays[i];
    oy = rays[n + i];
    oz = rays[2 * n + i];
    dx = rays[3 * n + i];
    dy = rays[4 * n + i];
    dz = rays[5 * n + i];
}

template <bool COUNT>
__device__ inline void flush_counts(unsigned long long* counters, unsigned nn, unsigned nt)
{
    if (COUNT) {
        atomicAdd(counters, (unsigned long long)nn);
        atomicAdd(counters + 1, (unsigned long long)nt);
    }
}

// STACK: node-group entries per lane held in LDS; 0 selects a private array.
template <bool COUNT, int STACK, int A>
__global__ __launch_bounds__(kBlock) void closest_kernel(const NodeSlot* __restrict__ nodes,
                                                         const Tri48* __restrict__ tris,
                                                         const float* __restrict__ rays, int64_t n,
                                                         int32_t* __restrict__ prim_id, float* __restrict__ t_out,
                                                         unsigned long long* counters)
{
    __shared__ uint2 lds[STACK > 0 ? STACK * kBlock : 1];
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    float ox, oy, oz, dx, dy, dz;
    load_ray(rays, n, i, ox, oy, oz, dx, dy, dz);
    RayStateT<A> r;
    init_ray(r, ox, oy, oz, dx, dy, dz);
    unsigned nn = 0, nt = 0;
    double best;
    unsigned leaf;
    int id;
    if (STACK > 0) {
        LdsStack<STACK> st{lds + threadIdx.x, kBlock, 0};
        id = trace_closest<COUNT>(r, nodes, tris, st, best, leaf, nn, nt);
    } else {
        PrivStack st;
        st.sp = 0;
        id = trace_closest<COUNT>(r, nodes, tris, st, best, leaf, nn, nt);
    }
    prim_id[i] = id;
    t_out[i] = id >= 0 ? (float)world_range(r, best) : INFINITY;
    flush_counts<COUNT>(counters, nn, nt);
}

template <bool COUNT, int STACK, int A>
__global__ __launch_bounds__(kBlock) void any_kernel(const NodeSlot* __restrict__ nodes, const Tri48* __restrict__ tris,
                                                     const float* __restrict__ rays,
                                                     const float* __restrict__ tmax, int64_t n,
                                                     uint8_t* __restrict__ occluded, unsigned long long* counters)
{
    __shared__ uint2 lds[STACK > 0 ? STACK * kBlock : 1];
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    float ox, oy, oz, dx, dy, dz;
    load_ray(rays, n, i, ox, oy, oz, dx, dy, dz);
    RayStateT<A> r;
    init_ray(r, ox, oy, oz, dx, dy, dz);
    const double how_far = tmax ? (double)tmax[i] : kRiseInfinity;
    const double how_far2 = scale_how_far(r, how_far);
    unsigned nn = 0, nt = 0;
    bool occ;
    if (STACK > 0) {
        LdsStack<STACK> st{lds + threadIdx.x, kBlock, 0};
        occ = trace_any<COUNT>(r, nodes, tris, st, how_far2, nn, nt);
    } else {
        PrivStack st;
        st.sp = 0;
        occ = trace_any<COUNT>(r, nodes, tris, st, how_far2, nn, nt);
    }
    occluded[i] = occ ? 1 : 0;
    flush_counts<COUNT>(counters, nn, nt);
}

inline unsigned grid_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// F64: certify on the leaf records' double vertices (tri64).
template <bool COUNT, int STACK, int A, bool F64 = false>
__global__ __launch_bounds__(kBlock) void closest_persistent(const NodeSlot* __restrict__ nodes,
                                                             const Tri48* __restrict__ tris,
                                                             const float* __restrict__ rays, int64_t n,
                                                             int32_t* __restrict__ prim_id,
                                                             float* __restrict__ t_out, unsigned long long* counters,
                                                             unsigned long long* next,
                                                             const double* __restrict__ tri64)
{
    __shared__ uint2 lds[STACK > 0 ? STACK * kBlock : 1];
    LdsStack<STACK> lst{lds + threadIdx.x, kBlock, 0};
    PrivStack pst;
    pst.sp = 0;
    RayStateT<A> r;
    ClosestState s;
    long long ray = -1;  // -1 idle, -2 exhausted
    unsigned nn = 0, nt = 0;
    const int64_t grid_threads = (int64_t)gridDim.x * blockDim.x;
    for (bool first_pass = true;; first_pass = false) {
        // The first item of every thread is assigned statically; later ones come from the counter.
        const long long got = first_pass ? first_item(n) : refill(ray == -1, ray >= 0, next, n, grid_threads);
        if (ray == -1 && got != -1) {
            ray = got;
            if (ray >= 0) {
                float ox, oy, oz, dx, dy, dz;
                load_ray(rays, n, ray, ox, oy, oz, dx, dy, dz);
                init_ray(r, ox, oy, oz, dx, dy, dz);
                closest_begin(r, s);
                lst.sp = 0;
                pst.sp = 0;
            }
        }
        if (__ballot(ray >= 0) == 0ull && __ballot(ray == -1) == 0ull) break;
        if (ray >= 0) {
            const bool more = STACK > 0 ? closest_step<COUNT, false, F64>(r, nodes, tris, lst, s, nn, nt, tri64)
                                        : closest_step<COUNT, false, F64>(r, nodes, tris, pst, s, nn, nt, tri64);
            if (!more) {
                prim_id[ray] = s.best_id;
                t_out[ray] = s.best_id >= 0 ? (float)world_range(r, s.best_t) : INFINITY;
                ray = -1;
            }
        }
    }
    flush_counts<COUNT>(counters, nn, nt);
}

template <bool COUNT, int STACK, int A, bool F64 = false>
__global__ __launch_bounds__(kBlock) void any_persistent(const NodeSlot* __restrict__ nodes,
                                                         const Tri48* __restrict__ tris,
                                                         const float* __restrict__ rays,
                                                         const float* __restrict__ tmax, int64_t n,
                                                         uint8_t* __restrict__ occluded, unsigned long long* counters,
                                                         unsigned long long* next,
                                                         const double* __restrict__ tri64)
{
    __shared__ uint2 lds[STACK > 0 ? STACK * kBlock : 1];
    LdsStack<STACK> lst{lds + threadIdx.x, kBlock, 0};
    PrivStack pst;
    pst.sp = 0;
    RayStateT<A> r;
    AnyState s;
    long long ray = -1;
    unsigned nn = 0, nt = 0;
    const int64_t grid_threads = (int64_t)gridDim.x * blockDim.x;
    for (bool first_pass = true;; first_pass = false) {
        // The first item of every thread is assigned statically; later ones come from the counter.
        const long long got = first_pass ? first_item(n) : refill(ray == -1, ray >= 0, next, n, grid_threads);
        if (ray == -1 && got != -1) {
            ray = got;
            if (ray >= 0) {
                float ox, oy, oz, dx, dy, dz;
                load_ray(rays, n, ray, ox, oy, oz, dx, dy, dz);
                init_ray(r, ox, oy, oz, dx, dy, dz);
                const double how_far = tmax ? (double)tmax[ray] : kRiseInfinity;
                any_begin(r, scale_how_far(r, how_far), s);
                lst.sp = 0;
                pst.sp = 0;
            }
        }
        if (__ballot(ray >= 0) == 0ull && __ballot(ray == -1) == 0ull) break;
        if (ray >= 0) {
            const bool more = STACK > 0 ? any_step<COUNT, false, F64>(r, nodes, tris, lst, s, nn, nt, tri64)
                                        : any_step<COUNT, false, F64>(r, nodes, tris, pst, s, nn, nt, tri64);
            if (!more) {
                occluded[ray] = s.hit ? 1 : 0;
                ray = -1;
            }
        }
    }
    flush_counts<COUNT>(counters, nn, nt);
}

// The closest_ex query's result staged in its rise_hip_hit slot until
// ex_records turns it into the record (keeps the record's double arithmetic
// out of the traversal loop's registers).
struct StagedHit {
    int32_t id;
    unsigned leaf;
    double t;
};
static_assert(sizeof(StagedHit) <= sizeof(rise_hip_hit), "a staged hit must fit in a hit record's slot");

// The intersection record of each ray from its staged hit, in place: the
// ray state is rebuilt from the ray as the traversal built it.
template <int A, bool F64>
__global__ __launch_bounds__(kBlock) void ex_records(const Tri48* __restrict__ tris, const float* __restrict__ rays,
                                                     int64_t n, rise_hip_hit* __restrict__ hits,
                                                     const double* __restrict__ tri64,
                                                     const double* __restrict__ attr)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const StagedHit sh = *reinterpret_cast<const StagedHit*>(hits + i);
    RayStateT<A> r;
    float ox, oy, oz, dx, dy, dz;
    load_ray(rays, n, i, ox, oy, oz, dx, dy, dz);
    init_ray(r, ox, oy, oz, dx, dy, dz);
    rise_hip_hit h;
    h.prim_id = sh.id;
    if (sh.id >= 0) {
        const Tri48* tri = tris + sh.leaf;
        double ab[2], nrm[3], pt[3], sn[3], uv[2];
        hit_attributes(r, tri, sh.t, ab, nrm, pt, F64 ? tri64 + 9 * (size_t)sh.leaf : nullptr,
                       attr ? attr + kAttrDoubles * (size_t)sh.leaf : nullptr, sn, uv);
        h.mesh = float_to_bits(tri->v1[3]);
        h.range = world_range(r, sh.t);
        h.alpha = ab[0];
        h.beta = ab[1];
        for (int k = 0; k < 3; ++k) {
            h.point[k] = (float)pt[k];
            h.geom_normal[k] = (float)nrm[k];
            h.shading_normal[k] = (float)sn[k];
        }
        h.uv[0] = (float)uv[0];
        h.uv[1] = (float)uv[1];
    } else {
        h.mesh = -1;
        h.range = INFINITY;
        h.alpha = h.beta = 0.0;
        for (int k = 0; k < 3; ++k) h.point[k] = h.geom_normal[k] = h.shading_normal[k] = 0.0f;
        h.uv[0] = h.uv[1] = 0.0f;
    }
    hits[i] = h;
}

// Closest hit with the face selection of rise_hip_intersect_closest_ex; the
// result is staged in hits for ex_records.
template <int STACK, int A, bool F64 = false>
__global__ __launch_bounds__(kBlock) void closest_ex_persistent(const NodeSlot* __restrict__ nodes,
                                                                const Tri48* __restrict__ tris,
                                                                const float* __restrict__ rays, int64_t n,
                                                                unsigned faces, rise_hip_hit* __restrict__ hits,
                                                                unsigned long long* next,
                                                                const double* __restrict__ tri64,
                                                                const double* __restrict__ attr)
{
    __shared__ uint2 lds[STACK > 0 ? STACK * kBlock : 1];
    LdsStack<STACK> lst{lds + threadIdx.x, kBlock, 0};
    PrivStack pst;
    pst.sp = 0;
    RayStateT<A> r;
    ClosestState s;
    long long ray = -1;
    unsigned nn = 0, nt = 0;
    const int64_t grid_threads = (int64_t)gridDim.x * blockDim.x;
    for (bool first_pass = true;; first_pass = false) {
        const long long got = first_pass ? first_item(n) : refill(ray == -1, ray >= 0, next, n, grid_threads);
        if (ray == -1 && got != -1) {
            ray = got;
            if (ray >= 0) {
                float ox, oy, oz, dx, dy, dz;
                load_ray(rays, n, ray, ox, oy, oz, dx, dy, dz);
                init_ray(r, ox, oy, oz, dx, dy, dz);
                closest_begin(r, s);
                s.faces = faces;
                lst.sp = 0;
                pst.sp = 0;
            }
        }
        if (__ballot(ray >= 0) == 0ull && __ballot(ray == -1) == 0ull) break;
        if (ray >= 0) {
            const bool more = STACK > 0 ? closest_step<false, true, F64>(r, nodes, tris, lst, s, nn, nt, tri64)
                                        : closest_step<false, true, F64>(r, nodes, tris, pst, s, nn, nt, tri64);
            if (!more) {
                *reinterpret_cast<StagedHit*>(hits + ray) = StagedHit{s.best_id, s.best_leaf, s.best_t};
                ray = -1;
            }
        }
    }
}

// Any hit with the face selection of rise_hip_intersect_any_ex.
template <int STACK, int A, bool F64 = false>
__global__ __launch_bounds__(kBlock) void any_ex_persistent(const NodeSlot* __restrict__ nodes,
                                                            const Tri48* __restrict__ tris,
                                                            const float* __restrict__ rays,
                                                            const float* __restrict__ tmax, int64_t n,
                                                            unsigned faces, uint8_t* __restrict__ occluded,
                                                            unsigned long long* next,
                                                            const double* __restrict__ tri64)
{
    __shared__ uint2 lds[STACK > 0 ? STACK * kBlock : 1];
    LdsStack<STACK> lst{lds + threadIdx.x, kBlock, 0};
    PrivStack pst;
    pst.sp = 0;
    RayStateT<A> r;
    AnyState s;
    long long ray = -1;
    unsigned nn = 0, nt = 0;
    const int64_t grid_threads = (int64_t)gridDim.x * blockDim.x;
    for (bool first_pass = true;; first_pass = false) {
        const long long got = first_pass ? first_item(n) : refill(ray == -1, ray >= 0, next, n, grid_threads);
        if (ray == -1 && got != -1) {
            ray = got;
            if (ray >= 0) {
                float ox, oy, oz, dx, dy, dz;
                load_ray(rays, n, ray, ox, oy, oz, dx, dy, dz);
                init_ray(r, ox, oy, oz, dx, dy, dz);
                const double how_far = tmax ? (double)tmax[ray] : kRiseInfinity;
                any_begin(r, scale_how_far(r, how_far), s);
                s.faces = faces;
                lst.sp = 0;
                pst.sp = 0;
            }
        }
        if (__ballot(ray >= 0) == 0ull && __ballot(ray == -1) == 0ull) break;
        if (ray >= 0) {
            const bool more = STACK > 0 ? any_step<false, true, F64>(r, nodes, tris, lst, s, nn, nt, tri64)
                                        : any_step<false, true, F64>(r, nodes, tris, pst, s, nn, nt, tri64);
            if (!more) {
                occluded[ray] = s.hit ? 1 : 0;
                ray = -1;
            }
        }
    }
}

__global__ void reset_counter(unsigned long long* c) { *c = 0ull; }

// Waves resident at once for a persistent kernel: CUs x blocks per CU.
template <class K>
unsigned persistent_grid(K kernel, int64_t n)
{
    static int cus = 0;
    if (cus == 0) {
        int dev = 0;
        hipGetDevice(&dev);
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (cus <= 0) cus = 256;
    }
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, 0) != hipSuccess || per_cu <= 0)
        per_cu = 4;
    const int64_t want = (int64_t)cus * per_cu;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(want, (int64_t)grid_for(n)));
}

}  // namespace

// The node-group stack holds at most one entry per level below the root, so
// the smallest LDS stack that fits the tree keeps occupancy highest (LDS is
// what limits waves per SIMD here).
hipError_t launch_closest(const TraceScene& s, const float* rays, int64_t n, int32_t* prim_id, float* t,
                          unsigned long long* counters, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    const int need = s.depth - 1;
    const dim3 grid(grid_for(n)), block(kBlock);
#if RISE_PERSISTENT
    hipLaunchKernelGGL(reset_counter, dim3(1), dim3(1), 0, stream, s.fetch_counter);
#define RISE_LAUNCH_CLOSEST_F(C, S, A, F)                                                                   \
    hipLaunchKernelGGL((closest_persistent<C, S, A, F>), dim3(persistent_grid(closest_persistent<C, S, A, F>, n)), \
                       block, 0, stream, s.nodes, s.tris, rays, n, prim_id, t, counters, s.fetch_counter, s.tri64)
#define RISE_LAUNCH_CLOSEST(C, S, A)                   \
    do {                                               \
        if (s.tri64) RISE_LAUNCH_CLOSEST_F(C, S, A, true); \
        else RISE_LAUNCH_CLOSEST_F(C, S, A, false);     \
    } while (0)
#else
    if (s.tri64) return hipErrorNotSupported;  // the one-ray-per-thread kernels certify on floats only
#define RISE_LAUNCH_CLOSEST(C, S, A)                                                                          \
    hipLaunchKernelGGL((closest_kernel<C, S, A>), grid, block, 0, stream, s.nodes, s.tris, rays, n, prim_id, t, \
                       counters)
#endif
#define RISE_LAUNCH_CLOSEST_A(C, S)                                             \
    do {                                                                         \
        if (s.arith == kArithX86Fma) RISE_LAUNCH_CLOSEST(C, S, kArithX86Fma);    \
        else RISE_LAUNCH_CLOSEST(C, S, kArithIeee);                              \
    } while (0)
    if (counters) {
        if (need <= 8) RISE_LAUNCH_CLOSEST_A(true, 8);
        else if (need <= 16) RISE_LAUNCH_CLOSEST_A(true, 16);
        else RISE_LAUNCH_CLOSEST_A(true, 0);
    } else {
        if (need <= 8) RISE_LAUNCH_CLOSEST_A(false, 8);
        else if (need <= 16) RISE_LAUNCH_CLOSEST_A(false, 16);
        else RISE_LAUNCH_CLOSEST_A(false, 0);
    }
#undef RISE_LAUNCH_CLOSEST_A
#undef RISE_LAUNCH_CLOSEST
#undef RISE_LAUNCH_CLOSEST_F
    return hipGetLastError();
}

hipError_t launch_any(const TraceScene& s, const float* rays, const float* tmax, int64_t n, uint8_t* occluded,
                      unsigned long long* counters, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    const int need = s.depth - 1;
    const dim3 grid(grid_for(n)), block(kBlock);
#if RISE_PERSISTENT
    hipLaunchKernelGGL(reset_counter, dim3(1), dim3(1), 0, stream, s.fetch_counter);
#define RISE_LAUNCH_ANY_F(C, S, A, F)                                                                        \
    hipLaunchKernelGGL((any_persistent<C, S, A, F>), dim3(persistent_grid(any_persistent<C, S, A, F>, n)), block, \
                       0, stream, s.nodes, s.tris, rays, tmax, n, occluded, counters, s.fetch_counter, s.tri64)
#define RISE_LAUNCH_ANY(C, S, A)                   \
    do {                                           \
        if (s.tri64) RISE_LAUNCH_ANY_F(C, S, A, true); \
        else RISE_LAUNCH_ANY_F(C, S, A, false);     \
    } while (0)
#else
    if (s.tri64) return hipErrorNotSupported;  // the one-ray-per-thread kernels certify on floats only
#define RISE_LAUNCH_ANY(C, S, A)                                                                             \
    hipLaunchKernelGGL((any_kernel<C, S, A>), grid, block, 0, stream, s.nodes, s.tris, rays, tmax, n, occluded, \
                       counters)
#endif
#define RISE_LAUNCH_ANY_A(C, S)                                             \
    do {                                                                     \
        if (s.arith == kArithX86Fma) RISE_LAUNCH_ANY(C, S, kArithX86Fma);    \
        else RISE_LAUNCH_ANY(C, S, kArithIeee);                              \
    } while (0)
    if (counters) {
        if (need <= 8) RISE_LAUNCH_ANY_A(true, 8);
        else if (need <= 16) RISE_LAUNCH_ANY_A(true, 16);
        else RISE_LAUNCH_ANY_A(true, 0);
    } else {
        if (need <= 8) RISE_LAUNCH_ANY_A(false, 8);
        else if (need <= 16) RISE_LAUNCH_ANY_A(false, 16);
        else RISE_LAUNCH_ANY_A(false, 0);
    }
#undef RISE_LAUNCH_ANY_A
#undef RISE_LAUNCH_ANY
#undef RISE_LAUNCH_ANY_F
    return hipGetLastError();
}

hipError_t launch_closest_ex(const TraceScene& s, const float* rays, int64_t n, unsigned faces, rise_hip_hit* hits,
                             hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    const int need = s.depth - 1;
    const dim3 block(kBlock);
    hipLaunchKernelGGL(reset_counter, dim3(1), dim3(1), 0, stream, s.fetch_counter);
#define RISE_LAUNCH_EX_F(S, A, F)                                                                           \
    do {                                                                                                    \
        hipLaunchKernelGGL((closest_ex_persistent<S, A, F>),                                                \
                           dim3(persistent_grid(closest_ex_persistent<S, A, F>, n)), block, 0, stream, s.nodes, \
                           s.tris, rays, n, faces, hits, s.fetch_counter, s.tri64, s.attr);                 \
        hipLaunchKernelGGL((ex_records<A, F>), dim3(grid_for(n)), block, 0, stream, s.tris, rays, n, hits,  \
                           s.tri64, s.attr);                                                                \
    } while (0)
#define RISE_LAUNCH_EX(S, A)                   \
    do {                                       \
        if (s.tri64) RISE_LAUNCH_EX_F(S, A, true); \
        else RISE_LAUNCH_EX_F(S, A, false);     \
    } while (0)
#define RISE_LAUNCH_EX_A(S)                                          \
    do {                                                             \
        if (s.arith == kArithX86Fma) RISE_LAUNCH_EX(S, kArithX86Fma); \
        else RISE_LAUNCH_EX(S, kArithIeee);                          \
    } while (0)
    if (need <= 8) RISE_LAUNCH_EX_A(8);
    else if (need <= 16) RISE_LAUNCH_EX_A(16);
    else RISE_LAUNCH_EX_A(0);
#undef RISE_LAUNCH_EX_A
#undef RISE_LAUNCH_EX
#undef RISE_LAUNCH_EX_F
    return hipGetLastError();
}

hipError_t launch_any_ex(const TraceScene& s, const float* rays, const float* tmax, int64_t n, unsigned faces,
                         uint8_t* occluded, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    const int need = s.depth - 1;
    const dim3 block(kBlock);
    hipLaunchKernelGGL(reset_counter, dim3(1), dim3(1), 0, stream, s.fetch_counter);
#define RISE_LAUNCH_EX_F(S, A, F)                                                                          \
    hipLaunchKernelGGL((any_ex_persistent<S, A, F>), dim3(persistent_grid(any_ex_persistent<S, A, F>, n)), block, \
                       0, stream, s.nodes, s.tris, rays, tmax, n, faces, occluded, s.fetch_counter, s.tri64)
#define RISE_LAUNCH_EX(S, A)                   \
    do {                                       \
        if (s.tri64) RISE_LAUNCH_EX_F(S, A, true); \
        else RISE_LAUNCH_EX_F(S, A, false);     \
    } while (0)
#define RISE_LAUNCH_EX_A(S)                                          \
    do {                                                             \
        if (s.arith == kArithX86Fma) RISE_LAUNCH_EX(S, kArithX86Fma); \
        else RISE_LAUNCH_EX(S, kArithIeee);                          \
    } while (0)
    if (need <= 8) RISE_LAUNCH_EX_A(8);
    else if (need <= 16) RISE_LAUNCH_EX_A(16);
    else RISE_LAUNCH_EX_A(0);
#undef RISE_LAUNCH_EX_A
#undef RISE_LAUNCH_EX
#undef RISE_LAUNCH_EX_F
    return hipGetLastError();
}

// Scratch of the allocation-free closest-hit query
// (rise_hip_intersect_closest_noalloc): word 0 is the persistent kernel's
// ray-fetch counter, cleared by a kernel at the start of every call.  A fixed
// 256 bytes per device whatever the ray count, held until the process ends.
// Host code never passes `scratch` itself to a kernel: it is a host-side
// stand-in; noalloc_scratch() returns the device's own address.
constexpr int kScratchWords = 32;
static __device__ __align__(256) unsigned long long scratch[kScratchWords];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
unsigned long long* g_scratch_ptr[kMaxScratchDevices] = {};
}  // namespace

hipError_t noalloc_scratch(unsigned long long** out)
{
    *out = nullptr;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = static_cast<unsigned long long*>(p);
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

}  // namespace rise_hip
