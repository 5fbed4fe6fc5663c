// object_kernels.hip -- batch closest-hit and any-hit kernels over a scene of
// objects with transforms (object_trace.h).  One lane per ray, the node-group
// stack in LDS when the trees are shallow enough, else in a private array;
// every lane runs the world tree, then the
// top-level tree over the transformed objects and the trees of the objects
// it reaches.
#include "object_kernels.h"

#include "object_trace.h"

namespace rise_hip {

namespace {

constexpr int kObjBlock = 128;

__device__ inline void load_ray6(const float* __restrict__ rays, int64_t n, int64_t i, float ray[6])
{
#pragma unroll
    for (int k = 0; k < 6; ++k) ray[k] = rays[(int64_t)k * n + i];
}

// The node-group stack: LDS entries per lane when the scene's trees fit
// (kObjLds), else a private array.
constexpr int kObjLds = 24;

template <int STACK>
struct StackOf {
    using type = ObjLdsStack<kObjLds>;
};
template <>
struct StackOf<0> {
    using type = ObjStack;
};

template <int STACK>
__device__ inline typename StackOf<STACK>::type make_stack(uint2* lds)
{
    if constexpr (STACK > 0) {
        return ObjLdsStack<kObjLds>{lds + threadIdx.x, kObjBlock, 0, 0, false};
    } else {
        ObjStack st;
        st.sp = st.base = 0;
        st.overflow = false;
        return st;
    }
}

// REC: write rise_hip_hit records; else prim_id and t (float world range).
template <int A, bool REC, int STACK>
__global__ __launch_bounds__(kObjBlock) void closest_objects(ObjectSceneView sc, const float* __restrict__ rays,
                                                             int64_t n, unsigned faces, int32_t* __restrict__ prim_id,
                                                             float* __restrict__ t_out,
                                                             rise_hip_hit* __restrict__ hits)
{
    __shared__ uint2 lds[STACK > 0 ? kObjLds * kObjBlock : 1];
    const int64_t i = (int64_t)blockIdx.x * kObjBlock + threadIdx.x;
    if (i >= n) return;
    float ray[6];
    load_ray6(rays, n, i, ray);
    auto st = make_stack<STACK>(lds);
    ObjectHit h;
    objects_closest<A>(sc, ray, faces, st, h);
    if (REC) {
        rise_hip_hit rec;
        objects_record<A>(sc, ray, h, rec);
        hits[i] = rec;
    } else {
        prim_id[i] = h.prim_id;
        t_out[i] = h.prim_id >= 0 ? (float)h.range : INFINITY;
    }
}

template <int A, int STACK>
__global__ __launch_bounds__(kObjBlock) void any_objects(ObjectSceneView sc, const float* __restrict__ rays,
                                                         const float* __restrict__ tmax, int64_t n, unsigned faces,
                                                         uint8_t* __restrict__ occluded)
{
    __shared__ uint2 lds[STACK > 0 ? kObjLds * kObjBlock : 1];
    const int64_t i = (int64_t)blockIdx.x * kObjBlock + threadIdx.x;
    if (i >= n) return;
    float ray[6];
    load_ray6(rays, n, i, ray);
    auto st = make_stack<STACK>(lds);
    const double how_far = tmax ? (double)tmax[i] : kRiseInfinity;
    occluded[i] = objects_any<A>(sc, ray, how_far, faces, st) ? 1 : 0;
}

inline unsigned obj_grid(int64_t n) { return (unsigned)((n + kObjBlock - 1) / kObjBlock); }

}  // namespace

hipError_t launch_closest_objects(const ObjectSceneView& sc, int arith, int stack_need, const float* rays, int64_t n,
                                  unsigned faces, int32_t* prim_id, float* t, rise_hip_hit* hits, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    const unsigned g = obj_grid(n);
#define RISE_OBJ_CLOSEST(A, S)                                                                                   \
    do {                                                                                                         \
        if (hits) closest_objects<A, true, S><<<g, kObjBlock, 0, stream>>>(sc, rays, n, faces, prim_id, t, hits); \
        else closest_objects<A, false, S><<<g, kObjBlock, 0, stream>>>(sc, rays, n, faces, prim_id, t, hits);     \
    } while (0)
    const bool lds = stack_need <= kObjLds;
    if (arith == kArithX86Fma) {
        if (lds) RISE_OBJ_CLOSEST(kArithX86Fma, 1);
        else RISE_OBJ_CLOSEST(kArithX86Fma, 0);
    } else {
        if (lds) RISE_OBJ_CLOSEST(kArithIeee, 1);
        else RISE_OBJ_CLOSEST(kArithIeee, 0);
    }
#undef RISE_OBJ_CLOSEST
    return hipGetLastError();
}

hipError_t launch_any_objects(const ObjectSceneView& sc, int arith, int stack_need, const float* rays,
                              const float* tmax, int64_t n, unsigned faces, uint8_t* occluded, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    const unsigned g = obj_grid(n);
    const bool lds = stack_need <= kObjLds;
    if (arith == kArithX86Fma) {
        if (lds) any_objects<kArithX86Fma, 1><<<g, kObjBlock, 0, stream>>>(sc, rays, tmax, n, faces, occluded);
        else any_objects<kArithX86Fma, 0><<<g, kObjBlock, 0, stream>>>(sc, rays, tmax, n, faces, occluded);
    } else {
        if (lds) any_objects<kArithIeee, 1><<<g, kObjBlock, 0, stream>>>(sc, rays, tmax, n, faces, occluded);
        else any_objects<kArithIeee, 0><<<g, kObjBlock, 0, stream>>>(sc, rays, tmax, n, faces, occluded);
    }
    return hipGetLastError();
}

}  // namespace rise_hip
