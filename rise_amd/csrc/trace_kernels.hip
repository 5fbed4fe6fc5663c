// trace_kernels.hip -- batch closest-hit and any-hit kernels over a ray SoA.
//
// These replace the per-ray IObjectManager::IntersectRay / IntersectShadowRay
// calls (Interfaces/IObjectManager.h:38-68, Managers/ObjectManager.cpp:270-373)
// with one launch over a whole batch.  One lane per ray; the node-group stack
// lives in LDS (16 entries per lane) when the tree is shallow enough, else in
// a private array.
#include "trace_core.h"
#include "trace_kernels.h"

namespace rise_hip {

namespace {

constexpr int kBlock = 128;
constexpr int kLdsStack = 16;

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

template <bool COUNT, bool LDS>
__global__ __launch_bounds__(kBlock) void closest_kernel(const Node8* __restrict__ nodes,
                                                         const Tri48* __restrict__ tris,
                                                         const float* __restrict__ rays, int64_t n,
                                                         int32_t* __restrict__ prim_id, float* __restrict__ t_out,
                                                         unsigned long long* counters)
{
    __shared__ uint2 lds[LDS ? kLdsStack * kBlock : 1];
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    float ox, oy, oz, dx, dy, dz;
    load_ray(rays, n, i, ox, oy, oz, dx, dy, dz);
    RayState r;
    init_ray(r, ox, oy, oz, dx, dy, dz);
    unsigned nn = 0, nt = 0;
    double best;
    unsigned leaf;
    int id;
    if (LDS) {
        LdsStack<kLdsStack> st{lds + threadIdx.x, kBlock, 0};
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

template <bool COUNT, bool LDS>
__global__ __launch_bounds__(kBlock) void any_kernel(const Node8* __restrict__ nodes, const Tri48* __restrict__ tris,
                                                     const float* __restrict__ rays,
                                                     const float* __restrict__ tmax, int64_t n,
                                                     uint8_t* __restrict__ occluded, unsigned long long* counters)
{
    __shared__ uint2 lds[LDS ? kLdsStack * kBlock : 1];
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    float ox, oy, oz, dx, dy, dz;
    load_ray(rays, n, i, ox, oy, oz, dx, dy, dz);
    RayState r;
    init_ray(r, ox, oy, oz, dx, dy, dz);
    const double how_far = tmax ? (double)tmax[i] : kRiseInfinity;
    const double how_far2 = scale_how_far(r, how_far);
    unsigned nn = 0, nt = 0;
    bool occ;
    if (LDS) {
        LdsStack<kLdsStack> st{lds + threadIdx.x, kBlock, 0};
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

}  // namespace

hipError_t launch_closest(const TraceScene& s, const float* rays, int64_t n, int32_t* prim_id, float* t,
                          unsigned long long* counters, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    // The node-group stack holds at most one entry per level below the root.
    const bool lds = s.depth - 1 <= kLdsStack;
    const dim3 grid(grid_for(n)), block(kBlock);
    if (counters) {
        if (lds)
            hipLaunchKernelGGL((closest_kernel<true, true>), grid, block, 0, stream, s.nodes, s.tris, rays, n, prim_id, t,
                               counters);
        else
            hipLaunchKernelGGL((closest_kernel<true, false>), grid, block, 0, stream, s.nodes, s.tris, rays, n, prim_id,
                               t, counters);
    } else {
        if (lds)
            hipLaunchKernelGGL((closest_kernel<false, true>), grid, block, 0, stream, s.nodes, s.tris, rays, n, prim_id,
                               t, counters);
        else
            hipLaunchKernelGGL((closest_kernel<false, false>), grid, block, 0, stream, s.nodes, s.tris, rays, n,
                               prim_id, t, counters);
    }
    return hipGetLastError();
}

hipError_t launch_any(const TraceScene& s, const float* rays, const float* tmax, int64_t n, uint8_t* occluded,
                      unsigned long long* counters, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    const bool lds = s.depth - 1 <= kLdsStack;
    const dim3 grid(grid_for(n)), block(kBlock);
    if (counters) {
        if (lds)
            hipLaunchKernelGGL((any_kernel<true, true>), grid, block, 0, stream, s.nodes, s.tris, rays, tmax, n, occluded,
                               counters);
        else
            hipLaunchKernelGGL((any_kernel<true, false>), grid, block, 0, stream, s.nodes, s.tris, rays, tmax, n,
                               occluded, counters);
    } else {
        if (lds)
            hipLaunchKernelGGL((any_kernel<false, true>), grid, block, 0, stream, s.nodes, s.tris, rays, tmax, n,
                               occluded, counters);
        else
            hipLaunchKernelGGL((any_kernel<false, false>), grid, block, 0, stream, s.nodes, s.tris, rays, tmax, n,
                               occluded, counters);
    }
    return hipGetLastError();
}

}  // namespace rise_hip
