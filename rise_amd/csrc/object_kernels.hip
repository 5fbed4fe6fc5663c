// object_kernels.hip -- batch closest-hit and any-hit kernels over a scene of
// objects with transforms (object_trace.h).  Persistent waves, one ray per
// lane at a time: each lane steps its query (the world tree, then the
// top-level tree over the transformed objects and the trees of the objects it
// reaches) and takes a new ray when it ends.  The node-group stack is in LDS
// when the trees are shallow enough, else in a private array.
#include "object_kernels.h"

#include "object_trace.h"
#include "wave_fetch.h"

#include <algorithm>

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

__global__ void reset_fetch(unsigned long long* c) { *c = 0ull; }

// A query's result staged in its rise_hip_hit slot until object_records
// turns it into the record (keeps the record's double arithmetic out of the
// traversal loop's registers).
static_assert(sizeof(ObjectHit) <= sizeof(rise_hip_hit) && alignof(ObjectHit) <= alignof(rise_hip_hit),
              "an ObjectHit must fit in a hit record's slot");

// Persistent: the grid is sized to the resident waves, each lane's first ray
// is assigned statically and later ones are fetched wave-wide (refill) as
// lanes finish.  REC: stage the ObjectHit in hits[ray] for object_records;
// else write prim_id and t (float world range).
template <int A, bool REC, int STACK>
__global__ __launch_bounds__(kObjBlock) void closest_objects(ObjectSceneView sc, const float* __restrict__ rays,
                                                             int64_t n, unsigned faces, int32_t* __restrict__ prim_id,
                                                             float* __restrict__ t_out,
                                                             rise_hip_hit* __restrict__ hits,
                                                             unsigned long long* next)
{
    __shared__ uint2 lds[STACK > 0 ? kObjLds * kObjBlock : 1];
    auto st = make_stack<STACK>(lds);
    ObjClosestQuery<A> Q;
    long long ray = -1;  // -1 idle, -2 exhausted
    const int64_t grid_threads = (int64_t)gridDim.x * blockDim.x;
    for (bool first_pass = true;; first_pass = false) {
        const long long got = first_pass ? first_item(n) : refill(ray == -1, ray >= 0, next, n, grid_threads);
        if (ray == -1 && got != -1) {
            ray = got;
            if (ray >= 0) {
                float r6[6];
                load_ray6(rays, n, ray, r6);
                objects_closest_begin(sc, r6, faces, st, Q);
            }
        }
        if (__ballot(ray >= 0) == 0ull && __ballot(ray == -1) == 0ull) break;
        if (ray >= 0 && !objects_closest_step(sc, st, Q)) {
            if (REC) {
                *reinterpret_cast<ObjectHit*>(hits + ray) = Q.h;
            } else {
                prim_id[ray] = Q.h.prim_id;
                t_out[ray] = Q.h.prim_id >= 0 ? (float)Q.h.range : INFINITY;
            }
            ray = -1;
        }
    }
}

// The intersection record of each ray from its staged ObjectHit, in place.
template <int A>
__global__ __launch_bounds__(kObjBlock) void object_records(ObjectSceneView sc, const float* __restrict__ rays,
                                                            int64_t n, rise_hip_hit* __restrict__ hits)
{
    const int64_t i = (int64_t)blockIdx.x * kObjBlock + threadIdx.x;
    if (i >= n) return;
    const ObjectHit h = *reinterpret_cast<const ObjectHit*>(hits + i);
    float r6[6];
    load_ray6(rays, n, i, r6);
    rise_hip_hit rec;
    objects_record<A>(sc, r6, h, rec);
    hits[i] = rec;
}

template <int A, int STACK>
__global__ __launch_bounds__(kObjBlock) void any_objects(ObjectSceneView sc, const float* __restrict__ rays,
                                                         const float* __restrict__ tmax, int64_t n, unsigned faces,
                                                         uint8_t* __restrict__ occluded, unsigned long long* next)
{
    __shared__ uint2 lds[STACK > 0 ? kObjLds * kObjBlock : 1];
    auto st = make_stack<STACK>(lds);
    ObjAnyQuery<A> Q;
    long long ray = -1;
    const int64_t grid_threads = (int64_t)gridDim.x * blockDim.x;
    for (bool first_pass = true;; first_pass = false) {
        const long long got = first_pass ? first_item(n) : refill(ray == -1, ray >= 0, next, n, grid_threads);
        if (ray == -1 && got != -1) {
            ray = got;
            if (ray >= 0) {
                float r6[6];
                load_ray6(rays, n, ray, r6);
                objects_any_begin(sc, r6, tmax ? (double)tmax[ray] : kRiseInfinity, faces, st, Q);
            }
        }
        if (__ballot(ray >= 0) == 0ull && __ballot(ray == -1) == 0ull) break;
        if (ray >= 0 && !objects_any_step(sc, st, Q)) {
            occluded[ray] = Q.hit ? 1 : 0;
            ray = -1;
        }
    }
}

inline unsigned obj_grid(int64_t n) { return (unsigned)((n + kObjBlock - 1) / kObjBlock); }

// Workgroups that fill the device once (resident blocks per CU x CUs), or
// fewer when there are fewer rays.
template <class K>
unsigned resident_grid(K kernel, int64_t n)
{
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kObjBlock, 0) != hipSuccess || per_cu <= 0)
        per_cu = 2;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((int64_t)cus * per_cu, (int64_t)obj_grid(n)));
}

}  // namespace

hipError_t launch_closest_objects(const ObjectSceneView& sc, int arith, int stack_need, const float* rays, int64_t n,
                                  unsigned faces, int32_t* prim_id, float* t, rise_hip_hit* hits,
                                  unsigned long long* next, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    reset_fetch<<<1, 1, 0, stream>>>(next);
#define RISE_OBJ_CLOSEST_K(A, R, S)                                                                          \
    closest_objects<A, R, S><<<resident_grid(closest_objects<A, R, S>, n), kObjBlock, 0, stream>>>(sc, rays, n, \
                                                                                                  faces, prim_id, \
                                                                                                  t, hits, next)
#define RISE_OBJ_CLOSEST(A, S)                                                                   \
    do {                                                                                         \
        if (hits) {                                                                              \
            RISE_OBJ_CLOSEST_K(A, true, S);                                                      \
            object_records<A><<<obj_grid(n), kObjBlock, 0, stream>>>(sc, rays, n, hits);          \
        } else {                                                                                 \
            RISE_OBJ_CLOSEST_K(A, false, S);                                                     \
        }                                                                                        \
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
#undef RISE_OBJ_CLOSEST_K
    return hipGetLastError();
}

hipError_t launch_any_objects(const ObjectSceneView& sc, int arith, int stack_need, const float* rays,
                              const float* tmax, int64_t n, unsigned faces, uint8_t* occluded,
                              unsigned long long* next, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    reset_fetch<<<1, 1, 0, stream>>>(next);
#define RISE_OBJ_ANY(A, S)                                                                               \
    any_objects<A, S><<<resident_grid(any_objects<A, S>, n), kObjBlock, 0, stream>>>(sc, rays, tmax, n, faces, \
                                                                                      occluded, next)
    const bool lds = stack_need <= kObjLds;
    if (arith == kArithX86Fma) {
        if (lds) RISE_OBJ_ANY(kArithX86Fma, 1);
        else RISE_OBJ_ANY(kArithX86Fma, 0);
    } else {
        if (lds) RISE_OBJ_ANY(kArithIeee, 1);
        else RISE_OBJ_ANY(kArithIeee, 0);
    }
#undef RISE_OBJ_ANY
    return hipGetLastError();
}

}  // namespace rise_hip
