// trace_kernels.h -- host-side launchers for the traversal kernels.
#pragma once

#include "bvh8.h"
#include "rise_hip.h"

#include <hip/hip_runtime_api.h>

#include <cstdint>

namespace rise_hip {

struct TraceScene {
    const NodeSlot* nodes;
    const Tri48* tris;
    int depth;  // wide-node levels; selects the stack variant
    unsigned long long* fetch_counter;  // device counter for dynamic ray fetch (one per in-flight call)
    int arith;  // rise_hip_arith: kArithIeee or kArithX86Fma (trace_core.h), same values
    const double* tri64;  // leaf records' double vertices (9 each) for the certifier, or NULL
    const double* attr;   // leaf records' normals and texture coordinates (kAttrDoubles each), or NULL
};

// Closest hit: rays SoA (ox, oy, oz, dx, dy, dz), n each.  counters: NULL, or
// two uint64 (nodes visited, triangles tested) accumulated over all rays.
hipError_t launch_closest(const TraceScene& s, const float* rays, int64_t n, int32_t* prim_id, float* t,
                          unsigned long long* counters, hipStream_t stream);

// Any hit within tmax (world distance per ray; NULL = unbounded).
hipError_t launch_any(const TraceScene& s, const float* rays, const float* tmax, int64_t n, uint8_t* occluded,
                      unsigned long long* counters, hipStream_t stream);

// Closest hit with a face selection (bit 0 front, bit 1 back) and hit records;
// any hit with a face selection.
hipError_t launch_closest_ex(const TraceScene& s, const float* rays, int64_t n, unsigned faces, rise_hip_hit* hits,
                             hipStream_t stream);
hipError_t launch_any_ex(const TraceScene& s, const float* rays, const float* tmax, int64_t n, unsigned faces,
                         uint8_t* occluded, hipStream_t stream);

// Device address (on the current device) of the allocation-free query's
// scratch counter; looked up once per device with hipGetSymbolAddress.
hipError_t noalloc_scratch(unsigned long long** out);

}  // namespace rise_hip
