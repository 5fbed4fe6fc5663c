// object_types.h -- device layout of a scene of objects with transforms,
// shared by the host assembly (object_scene.cpp) and the kernels
// (object_trace.h).
#pragma once

#include "bvh8.h"

#include <cstdint>

namespace rise_hip {

// Node-group stack entries of an object query: one traversal of the world
// tree, or the top-level tree plus one object's tree on top of it (the
// upload checks that their depths fit).
constexpr int kObjStack = 96;

// One transformed object on the device.
struct DevObject {
    double to_world[16];   // m_mxFinalTrans (Matrix4 order)
    double to_object[16];  // m_mxInvFinalTrans
    double lo[3], hi[3];   // world box of the object's mesh, padded (never smaller than the exact one)
    double slack;          // bound on the rounding of a world position through M^-1 then M, per unit of
                           // coordinate size (2^-40 times the condition number of the linear part)
    int32_t blas;          // index of the object's mesh tree
    int32_t object;        // the object's index in the caller's list
};

struct BlasView {
    const NodeSlot* nodes;
    const Tri48* tris;
    const double* tri64;  // the triangles' double vertices, 9 per leaf record (leaf order)
    const double* attr;   // per-corner normals and texture coordinates, kAttrDoubles per leaf record, or NULL
};

struct ObjectSceneView {
    const NodeSlot* world_nodes;  // NULL when no object has identity matrices
    const Tri48* world_tris;
    const double* world_tri64;    // the world triangles' double vertices, 9 per leaf record
    const double* world_attr;     // their per-corner normals and texture coordinates, or NULL
    const DevObject* objects;
    int32_t n_objects;
    const BlasView* blas;
    // Top-level tree over the transformed objects' world boxes, in the same
    // node layout; each leaf record's v0.w is an index into `objects`.
    const NodeSlot* top_nodes;
    const Tri48* top_leaves;
};

}  // namespace rise_hip
