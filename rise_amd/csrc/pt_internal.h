// pt_internal.h -- library-internal definition of a path-tracing scene.
#pragma once

#include "pt_device.h"

#include <cstdint>
#include <vector>

struct rise_hip_batch;

struct rise_hip_pt_scene {
    int device = 0;
    std::vector<void*> allocs;      // device tables owned by the scene
    rise_hip::PtSceneView view{};
    int64_t n_tris = 0;
    bool has_glossy = false;        // a material scatters glossily (counts against max_glossy_bounce)
    bool has_transmission = false;  // a material transmits (counts against max_transmission_bounce)
    unsigned material_kinds = 0;    // bit k: a material of kind k (rise_hip_material_type) is present
    bool has_media = false;         // some object holds a participating medium
    rise_hip_batch* solo = nullptr;  // cached batch for rise_hip_render, rebuilt when the request changes
    std::vector<rise_hip::DevDirLight> dir_lights;  // rise_hip_pt_scene_set_directional_lights
    bool has_sun = false;           // the sky's matched sun light
    rise_hip::DevDirLight sun{};
};

// Frees the scene's cached single-request batch (pt_render.hip).
void release_solo_batch(rise_hip_pt_scene* s);

namespace rise_hip {
struct Bvh2Tree;
// rise_hip_pt_scene_create_with_media with, optionally, a binary tree per
// object (object_trees[o] over that object's triangles, NULL: built), as a
// .risemesh file's BVH2 cache provides (pt_scene.hip).
rise_hip_status pt_scene_create(const rise_hip_object_desc* objects, size_t n_objects,
                                const rise_hip_material_desc* materials, size_t n_materials,
                                const rise_hip_spd_desc* spds, size_t n_spds, const rise_hip_medium_desc* media,
                                size_t n_media, const rise_hip_object_media* object_media,
                                const Bvh2Tree* const* object_trees, rise_hip_pt_scene** out_scene);
}  // namespace rise_hip
