// object_kernels.h -- host-side launchers for the object-scene kernels.
#pragma once

#include "rise_hip.h"

#include <hip/hip_runtime_api.h>

#include <cstdint>

namespace rise_hip {

struct ObjectSceneView;

// Closest hit over objects: hits (records) when non-NULL, else prim_id and t.
// arith: kArithIeee or kArithX86Fma.
// stack_need: node-group stack entries a query needs (the scene's depths).
// next: a device counter for the rays' dynamic fetch, reset by the launch
// (one per call in flight).
hipError_t launch_closest_objects(const ObjectSceneView& sc, int arith, int stack_need, const float* rays, int64_t n,
                                  unsigned faces, int32_t* prim_id, float* t, rise_hip_hit* hits,
                                  unsigned long long* next, hipStream_t stream);

// Any hit over objects within tmax (world dHowFar per ray; NULL = unbounded).
hipError_t launch_any_objects(const ObjectSceneView& sc, int arith, int stack_need, const float* rays,
                              const float* tmax, int64_t n, unsigned faces, uint8_t* occluded,
                              unsigned long long* next, hipStream_t stream);

}  // namespace rise_hip
