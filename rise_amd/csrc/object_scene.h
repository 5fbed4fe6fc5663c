// object_scene.h -- host-side assembly of a scene of objects with transforms
// (rise_hip_scene_upload_objects and its host validation entry point).
#pragma once

#include "bvh_build.h"
#include "object_types.h"
#include "rise_hip.h"

#include <string>
#include <vector>

namespace rise_hip {

struct ObjectSceneHost {
    WideBVH world;                  // identity objects, world space; v1.w = object index
    int64_t world_tris = 0;
    std::vector<WideBVH> blas;      // one per mesh used by a transformed object, object space
    std::vector<DevObject> objects; // the transformed objects, in the caller's order
    WideBVH top;                    // over the transformed objects' world boxes (leaf v0.w = index into objects)
    int64_t total_tris = 0;         // triangles over all trees
};

// Validates the arguments and builds every tree.  Identity objects (both
// matrices exactly the identity) are merged into `world`; each other object
// refers to the tree of its mesh.  Triangle ids: the mesh's tri_ids, or the
// triangle's index within its mesh.
rise_hip_status build_object_scene(const rise_hip_mesh_desc* meshes, size_t n_meshes,
                                   const rise_hip_object_instance* objects, size_t n_objects, ObjectSceneHost& out);

}  // namespace rise_hip
