// object_types.h -- device layout of a scene of objects with transforms,
// shared by the host assembly (object_scene.cpp) and the kernels
// (object_trace.h).
#pragma once

#include "bvh8.h"

#include <cstdint>

namespace rise_hip {

// One transformed object on the device.
struct DevObject {
    double to_world[16];   // m_mxFinalTrans (Matrix4 order)
    double to_object[16];  // m_mxInvFinalTrans
    double lo[3], hi[3];   // world box of the object's mesh, padded (never smaller than the exact one)
    int32_t blas;          // index of the object's mesh tree
    int32_t object;        // the object's index in the caller's list
};

struct BlasView {
    const NodeSlot* nodes;
    const Tri48* tris;
};

struct ObjectSceneView {
    const NodeSlot* world_nodes;  // NULL when no object has identity matrices
    const Tri48* world_tris;
    const DevObject* objects;
    int32_t n_objects;
    const BlasView* blas;
};

}  // namespace rise_hip
