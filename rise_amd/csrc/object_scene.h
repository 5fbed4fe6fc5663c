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
    std::vector<double> world_tri64;  // its leaf records' double vertices (9 each, leaf order)
    std::vector<double> world_attr;   // their normals and texture coordinates (kAttrDoubles each), or empty
    int64_t world_tris = 0;
    std::vector<WideBVH> blas;      // one per mesh used by a transformed object, object space
    std::vector<std::vector<double>> blas_tri64;
    std::vector<std::vector<double>> blas_attr;  // empty for a mesh without normals or coordinates
    std::vector<DevObject> objects; // the transformed objects, in the caller's order
    WideBVH top;                    // over the transformed objects' world boxes (leaf v0.w = index into objects)
    int stack_need = 0;             // node-group stack entries a query needs at most
    int64_t total_tris = 0;         // triangles over all trees
};

// Validates the arguments and builds every tree.  Identity objects (both
// matrices exactly the identity) are merged into `world`; each other object
// refers to the tree of its mesh.  Triangle ids: the mesh's tri_ids, or the
// triangle's index within its mesh.
// A mesh's float vertices (as given, or its double vertices rounded to float,
// as the reference's filter takes them) and double vertices (as given, or the
// floats widened), n_tris * 9 each, appended.
void mesh_vertices(const rise_hip_mesh_desc& m, std::vector<float>& f, std::vector<double>& d);

// A mesh's per-corner vertex normals and texture coordinates (kAttrDoubles
// per triangle, zeros where the mesh has none), appended.
void mesh_attrs(const rise_hip_mesh_desc& m, std::vector<double>& a);

// The leaf-record flags a mesh contributes (kTriDoubleSided, kTriNormals,
// kTriCoords).
uint32_t mesh_tri_flags(const rise_hip_mesh_desc& m);

// Per-leaf-record rows of `stride` doubles gathered from rows indexed by the
// running index the tree was built over (the record's v0.w before ids are
// assigned).
std::vector<double> leaf_rows(const WideBVH& bvh, const std::vector<double>& rows, int stride);

// The leaf records' double vertices: record i takes v64[9 p .. 9 p + 8] where
// p is the running index the tree was built over (the record's v0.w before
// ids are assigned).
std::vector<double> leaf_vertices(const WideBVH& bvh, const std::vector<double>& v64);

rise_hip_status build_object_scene(const rise_hip_mesh_desc* meshes, size_t n_meshes,
                                   const rise_hip_object_instance* objects, size_t n_objects, ObjectSceneHost& out);

}  // namespace rise_hip
