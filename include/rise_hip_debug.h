/*
 * rise_hip_debug.h -- validation entry points that need no GPU.
 *
 * These have no counterpart in the reference.  They exist for callers that
 * check the library on a machine without a GPU (the CPU test suite): they
 * build the same 8-wide BVH the GPU path uploads and run the same traversal
 * and certifier code (trace_core.h is compiled for host and device) on host
 * threads.  They are validation tools, not a fallback: the GPU entry points
 * never call them.
 */
#ifndef RISE_HIP_DEBUG_H
#define RISE_HIP_DEBUG_H

#include "rise_hip.h"
#include "rise_hip_risemesh.h"

#ifdef __cplusplus
extern "C" {
#endif

/* mode bit 0 clear: closest hit (fills prim_id, t); set: any hit (fills
 * occluded; tmax may be NULL).  mode bit 1 set: RISE_HIP_ARITH_X86_FMA
 * arithmetic, else RISE_HIP_ARITH_IEEE.  All buffers are host memory;
 * rays_soa as in rise_hip_intersect_closest.  counters (optional, 2 uint64)
 * receive nodes visited and leaf triangles tested; stats (optional) the
 * build statistics. */
rise_hip_status rise_hip_debug_trace_host(const float* tris, int64_t n_tris, const float* rays_soa,
                                          const float* tmax, int64_t n, int32_t mode, int32_t* prim_id,
                                          float* t, uint8_t* occluded, uint64_t* counters,
                                          rise_hip_scene_stats* stats);

/* The same over a .risemesh file's triangles (ids = file order), the 8-wide
 * tree collapsed from the file's BVH2 cache when it is valid
 * (stats->bvh_source = 1), else built. */
rise_hip_status rise_hip_debug_trace_host_risemesh(const rise_hip_risemesh* mesh, const float* rays_soa,
                                                   const float* tmax, int64_t n, int32_t mode, int32_t* prim_id,
                                                   float* t, uint8_t* occluded, uint64_t* counters,
                                                   rise_hip_scene_stats* stats);

/* rise_hip_debug_trace_host with the face selection and hit records of
 * rise_hip_intersect_closest_ex / rise_hip_intersect_any_ex (mode bits as
 * above; faces: rise_hip_faces).  hits (closest) or occluded (any): host, n
 * entries.  The triangles form one mesh (mesh index 0). */
rise_hip_status rise_hip_debug_trace_ex_host(const float* tris, int64_t n_tris, const float* rays_soa,
                                             const float* tmax, int64_t n, int32_t mode, int32_t faces,
                                             rise_hip_hit* hits, uint8_t* occluded);

/* The kernels' FP32 slab test of one 8-wide node (80 bytes, the layout the
 * builder writes) against one ray (ox, oy, oz, dx, dy, dz), on the host:
 * bit s of *accept set when slot s is occupied and its box is taken for the
 * parameter range up to `best` (a distance along the normalized direction;
 * >= 1.7976931348623158e308 for no limit). */
rise_hip_status rise_hip_debug_node_accept_host(const void* node, const float ray[6], double best,
                                                uint32_t* accept);

/* rise_hip_scene_upload_objects followed by rise_hip_intersect_closest_ex
 * (mode bit 0 clear: fills hits) or rise_hip_intersect_any_ex (bit 0 set:
 * fills occluded; tmax may be NULL), on the host: the same trees and the same
 * per-ray code.  mode bit 1 selects RISE_HIP_ARITH_X86_FMA.  All buffers are
 * host memory. */
rise_hip_status rise_hip_debug_objects_host(const rise_hip_mesh_desc* meshes, size_t n_meshes,
                                            const rise_hip_object_instance* objects, size_t n_objects,
                                            const float* rays_soa, const float* tmax, int64_t n, int32_t mode,
                                            int32_t faces, rise_hip_hit* hits, uint8_t* occluded);

/* The kernels' double slab test of one 8-wide node against the object-space
 * ray that the world ray (ox, oy, oz, dx, dy, dz) becomes under to_object
 * (16 doubles, Matrix4 order), on the host; `best` is a distance along the
 * object-space normalized direction. */
rise_hip_status rise_hip_debug_node_accept_obj_host(const void* node, const float ray[6], const double to_object[16],
                                                    double best, uint32_t* accept);

#ifdef __cplusplus
}
#endif

#endif /* RISE_HIP_DEBUG_H */
