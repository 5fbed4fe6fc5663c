// object_kernels.hip -- batch closest-hit and any-hit kernels over a scene of
// objects with transforms (object_trace.h).  One lane per ray, the node-group
// stack in a private array; every lane runs the world tree, then the
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

// REC: write rise_hip_hit records; else prim_id and t (float world range).
template <int A, bool REC>
__global__ __launch_bounds__(kObjBlock) void closest_objects(ObjectSceneView sc, const float* __restrict__ rays,
                                                             int64_t n, unsigned faces, int32_t* __restrict__ prim_id,
                                                             float* __restrict__ t_out,
                                                             rise_hip_hit* __restrict__ hits)
{
    const int64_t i = (int64_t)blockIdx.x * kObjBlock + threadIdx.x;
    if (i >= n) return;
    float ray[6];
    load_ray6(rays, n, i, ray);
    ObjStack st;
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

template <int A>
__global__ __launch_bounds__(kObjBlock) void any_objects(ObjectSceneView sc, const float* __restrict__ rays,
                                                         const float* __restrict__ tmax, int64_t n, unsigned faces,
                                                         uint8_t* __restrict__ occluded)
{
    const int64_t i = (int64_t)blockIdx.x * kObjBlock + threadIdx.x;
    if (i >= n) return;
    float ray[6];
    load_ray6(rays, n, i, ray);
    ObjStack st;
    const double how_far = tmax ? (double)tmax[i] : kRiseInfinity;
    occluded[i] = objects_any<A>(sc, ray, how_far, faces, st) ? 1 : 0;
}

inline unsigned obj_grid(int64_t n) { return (unsigned)((n + kObjBlock - 1) / kObjBlock); }

}  // namespace

hipError_t launch_closest_objects(const ObjectSceneView& sc, int arith, const float* rays, int64_t n, unsigned faces,
                                  int32_t* prim_id, float* t, rise_hip_hit* hits, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    const unsigned g = obj_grid(n);
    if (arith == kArithX86Fma) {
        if (hits) closest_objects<kArithX86Fma, true><<<g, kObjBlock, 0, stream>>>(sc, rays, n, faces, prim_id, t, hits);
        else closest_objects<kArithX86Fma, false><<<g, kObjBlock, 0, stream>>>(sc, rays, n, faces, prim_id, t, hits);
    } else {
        if (hits) closest_objects<kArithIeee, true><<<g, kObjBlock, 0, stream>>>(sc, rays, n, faces, prim_id, t, hits);
        else closest_objects<kArithIeee, false><<<g, kObjBlock, 0, stream>>>(sc, rays, n, faces, prim_id, t, hits);
    }
    return hipGetLastError();
}

hipError_t launch_any_objects(const ObjectSceneView& sc, int arith, const float* rays, const float* tmax, int64_t n,
                              unsigned faces, uint8_t* occluded, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    const unsigned g = obj_grid(n);
    if (arith == kArithX86Fma) any_objects<kArithX86Fma><<<g, kObjBlock, 0, stream>>>(sc, rays, tmax, n, faces, occluded);
    else any_objects<kArithIeee><<<g, kObjBlock, 0, stream>>>(sc, rays, tmax, n, faces, occluded);
    return hipGetLastError();
}

}  // namespace rise_hip
