/*
 * rise_hip_risemesh.h -- RISE's binary mesh files (.risemesh, signature
 * "RISETMGI", versions 1-5) with their BVH2 cache block.
 *
 * Replaces TriangleMeshGeometryIndexed::Serialize / Deserialize
 * (Geometry/TriangleMeshGeometryIndexed.cpp:586-1000) and the BVH2 block of
 * BVH::Serialize / Deserialize (Acceleration/BVH.h:338-423), as loaded by the
 * risemesh_geometry chunk (Parsers/ChunkParserRegistry.cpp:5137,
 * Job::AddRISEMeshTriangleMeshGeometry).  Little-endian, as the reference's
 * MemoryBuffer writes on x86-64 and ARM64.
 *
 * A file's BVH2 cache, when present and valid, replaces the binned-SAH stage
 * of the 8-wide build (its boxes are refit to the triangles and it is
 * collapsed as usual), as the reference skips its SAH rebuild; an invalid
 * cache is ignored and the tree is built, as the reference rebuilds.
 * Legacy v2/v3 files that carry a BSP/octree block are refused.
 * Host code; no GPU is needed except for rise_hip_risemesh_upload.
 */
#ifndef RISE_HIP_RISEMESH_H
#define RISE_HIP_RISEMESH_H

#include "rise_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rise_hip_risemesh rise_hip_risemesh;

typedef struct rise_hip_risemesh_info {
    int32_t version;         /* 1..5 */
    int32_t face_normals;    /* bUseFaceNormals */
    int32_t double_sided;    /* bDoubleSided */
    int32_t has_bvh_cache;   /* a BVH2 block was present */
    int32_t bvh_cache_valid; /* ... and passed the structural checks */
    int32_t reserved;
    int64_t n_points;
    int64_t n_normals;
    int64_t n_coords;
    int64_t n_colors;
    int64_t n_tris;
    int64_t n_bvh_nodes;     /* nodes of the BVH2 block (0 without one) */
} rise_hip_risemesh_info;

/* Reads and checks a .risemesh file.  Index out of range, truncation or an
 * unknown signature/version fail with a message. */
rise_hip_status rise_hip_risemesh_read(const char* path, rise_hip_risemesh** out_mesh);

rise_hip_status rise_hip_risemesh_get_info(const rise_hip_risemesh* mesh, rise_hip_risemesh_info* out_info);

/* The triangles as 9 floats each (v0, v1, v2; the file's doubles rounded to
 * float), in file order: triangle i has id i in ray queries. */
rise_hip_status rise_hip_risemesh_get_triangles(const rise_hip_risemesh* mesh, float* out_tris);

/* Writes a version-5 file: vertices deduplicated (exact matches), optional
 * per-corner texture coordinates (6 floats per triangle; NULL: none), no
 * normals or colors, and, if with_bvh_cache, the BVH2 block of the binned-SAH
 * build (the reference loads it and skips its own SAH build). */
rise_hip_status rise_hip_risemesh_write(const char* path, const float* tris, int64_t n_tris, const float* uvs,
                                        int32_t double_sided, int32_t with_bvh_cache);

/* The ray-cast scene of the mesh (as rise_hip_scene_upload of its triangles,
 * ids = file order), built from the file's BVH2 cache when it is valid;
 * stats.bvh_source then reads 1. */
rise_hip_status rise_hip_risemesh_upload(const rise_hip_risemesh* mesh, rise_hip_scene** out_scene);

rise_hip_status rise_hip_risemesh_release(rise_hip_risemesh* mesh);

/* Scene cache: the resident ray-cast scene's built 8-wide tree (nodes and
 * triangles, copied back from the device) saved to `path`, and read back and
 * uploaded with no build at all (stats.bvh_source = 2).  No counterpart in the
 * reference, whose .risemesh cache holds the binary tree and re-derives its
 * 4-wide tree on load; for callers that re-create large resident scenes
 * (servers restarting, scenes evicted and re-admitted).  Layout: "RHBVH8\0\0",
 * uint32 version 1, uint32 node bytes (80), uint32 depth, uint32 reserved,
 * uint64 node count, uint64 triangle count, the nodes, the 48-byte
 * triangles.  A file is checked before upload: sizes, every child and
 * triangle range in bounds, children after their parent, depth within the
 * traversal stack. */
rise_hip_status rise_hip_scene_cache_write(const rise_hip_scene* scene, const char* path);
rise_hip_status rise_hip_scene_cache_read(const char* path, rise_hip_scene** out_scene);

#ifdef __cplusplus
}
#endif

#endif /* RISE_HIP_RISEMESH_H */
