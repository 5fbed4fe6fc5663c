/*
 * rise_hip.h -- C ABI of the MI355X (gfx950) implementation of RISE's
 * ray-cast and spectral path-tracing hot path.
 *
 * Conventions
 *   - Every entry point returns rise_hip_status; RISE_HIP_OK is 0.  On
 *     failure, rise_hip_last_error() returns a message for the calling
 *     thread.  No C++ exception crosses this boundary.  (RISE reports
 *     failure through bool returns plus GlobalLog(); see
 *     src/Library/RISE_API.h factory functions.)
 *   - Objects are opaque handles created by *_create / *_upload and freed by
 *     the matching *_release.  The library owns every device buffer it
 *     allocates until release.
 *   - Buffers named "device" are device (HBM) pointers; "host" are host
 *     pointers.  Work is enqueued on the caller's stream and is asynchronous
 *     unless stated otherwise.  Calls on one scene from several threads are
 *     safe; a scene must not be released while work using it is in flight.
 *
 * Reference interfaces replaced (paths relative to src/Library/ of RISE):
 *   - Interfaces/IObjectManager.h:38-68 (IntersectRay, IntersectShadowRay,
 *     PrepareForRendering): rise_hip_scene_upload + rise_hip_intersect_*.
 *   - Interfaces/IRasterizer.h:70 (RasterizeScene) with the per-pixel hook
 *     PixelBasedRasterizerHelper::IntegratePixel
 *     (Rendering/PathTracingSpectralRasterizer.h:48-58): rise_hip_render.
 */
#ifndef RISE_HIP_H
#define RISE_HIP_H

#include <stddef.h>
#include <stdint.h>

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum rise_hip_status {
    RISE_HIP_OK = 0,
    RISE_HIP_ERR_INVALID_ARGUMENT = 1,
    RISE_HIP_ERR_OUT_OF_MEMORY = 2,
    RISE_HIP_ERR_DEVICE = 3,
    RISE_HIP_ERR_BUILD = 4,
    RISE_HIP_ERR_NOT_SUPPORTED = 5
} rise_hip_status;

/* Message describing the last failure on the calling thread ("" if none). */
const char* rise_hip_last_error(void);

/* Library version, "major.minor.patch". */
const char* rise_hip_version(void);

/* ------------------------------------------------------------------------ */
/* Scenes (geometry resident in HBM)                                        */
/* ------------------------------------------------------------------------ */

typedef struct rise_hip_scene rise_hip_scene;

/* One triangle mesh in world space (identity object transform).
 * vertices:     host, n_tris * 9 floats (v0.xyz, v1.xyz, v2.xyz per triangle).
 * n_tris:       triangle count.
 * tri_ids:      host, optional n_tris ids reported for hits (NULL: the running
 *               index over all meshes of the scene, starting at 0).
 * vertices_f64: host, optional n_tris * 9 doubles, the same layout: the
 *               mesh's vertices as RISE holds them (Point3, double).  When
 *               given, `vertices` may be NULL: the double certifier and the
 *               face test use these, the float pre-filter and the tree their
 *               float rounding, as the reference's mesh does
 *               (...Specializations.h:32-47, :141-142).
 * flags:        RISE_HIP_MESH_DOUBLE_SIDED: the mesh is double-sided
 *               (TriangleMeshGeometryIndexed's bDoubleSided,
 *               TriangleMeshGeometryIndexed.cpp:151-205): the face selection
 *               of the queries does not apply to it, and a hit record's
 *               normals are turned toward the ray's origin.
 * normals:      host, optional n_tris * 9 doubles: each triangle's three
 *               vertex normals (the mesh's pNormals of its corners); a hit
 *               record's shading normal interpolates them
 *               (...Specializations.h:194-196), else it is the face normal.
 * coords:       host, optional n_tris * 6 doubles: each triangle's three
 *               texture coordinates (pCoords); a hit record's uv
 *               interpolates them (:214-216), else it is 0. */
#define RISE_HIP_MESH_DOUBLE_SIDED 1u
typedef struct rise_hip_mesh_desc {
    const float* vertices;
    int64_t n_tris;
    const int32_t* tri_ids;
    const double* vertices_f64;
    uint32_t flags;
    const double* normals;
    const double* coords;
} rise_hip_mesh_desc;

typedef struct rise_hip_scene_stats {
    int64_t n_tris;
    int64_t n_nodes;        /* 8-wide nodes */
    int64_t n_bvh2_nodes;   /* nodes of the binary build before collapse */
    int32_t depth;          /* wide-node levels */
    int32_t node_bytes;     /* 80 */
    int32_t tri_bytes;      /* 48 */
    int32_t bvh_source;     /* 0: built here; 1: from a .risemesh BVH2 cache; 2: from a scene cache */
    double build_ms;        /* host build time */
    double sah_cost;        /* SAH cost of the binary build / root area */
    int64_t device_bytes;   /* HBM held by the scene */
} rise_hip_scene_stats;

/* Builds the acceleration structure on the host and uploads it to the
 * current device.  Replaces ObjectManager::PrepareForRendering /
 * TriangleMeshGeometryIndexed::DoneIndexedTriangles
 * (Managers/ObjectManager.cpp:404-423, Geometry/TriangleMeshGeometryIndexed.cpp:444-540).
 * Synchronous. */
rise_hip_status rise_hip_scene_upload(const rise_hip_mesh_desc* meshes, size_t n_meshes,
                                      rise_hip_scene** out_scene);

/* One object of an ObjectManager: a mesh placed by its transform
 * (Objects/Object.h; Utilities/Transformable.cpp:438-449).  Matrices are
 * RISE's Matrix4 in member order _00, _01, ..., _33 (row-vector convention:
 * a point p maps to p.x row 0 + p.y row 1 + p.z row 2 + row 3,
 * Utilities/Math3D/PointsOps.h:126-133); the fourth column is not read. */
typedef struct rise_hip_object_instance {
    int32_t mesh;          /* index into the meshes of rise_hip_scene_upload_objects */
    int32_t reserved;      /* must be 0 */
    double to_world[16];   /* m_mxFinalTrans */
    double to_object[16];  /* m_mxInvFinalTrans (Matrix4Ops::Inverse of to_world) */
} rise_hip_object_instance;

/* Uploads a scene of objects: meshes in their own (object) space, each placed
 * by one or more objects.  Ray queries on the scene follow
 * ObjectManager::IntersectRay / IntersectShadowRay over the objects
 * (Managers/ObjectManager.cpp:270-373) with Object::IntersectRay /
 * IntersectRay_IntersectionOnly per object (Objects/Object.cpp:392-712):
 * the world ray is taken into the object's frame in double (origin by
 * to_object, direction by to_object's linear part, then normalized), the
 * mesh is intersected there, and the hit point, backed off 1e-12, goes back
 * to world space by to_world; the world distance from the ray origin to that
 * point is the hit's range, and the strictly smallest range wins in object
 * order.  A shadow ray's dHowFar is scaled by |to_object d| per object.
 * Objects whose two matrices are exactly the identity share one world-space
 * tree; every other object is intersected in its own frame through its
 * mesh's object-space tree (built once per mesh).  Triangle ids: the mesh's
 * tri_ids, or the triangle's index within its mesh.  On this scene,
 * rise_hip_hit.mesh holds the index of the object hit, and t /
 * rise_hip_hit.range the world range.  The _counted queries and the scene
 * cache are not supported on a scene with transformed objects
 * (RISE_HIP_ERR_NOT_SUPPORTED).  Synchronous. */
rise_hip_status rise_hip_scene_upload_objects(const rise_hip_mesh_desc* meshes, size_t n_meshes,
                                              const rise_hip_object_instance* objects, size_t n_objects,
                                              rise_hip_scene** out_scene);

rise_hip_status rise_hip_scene_release(rise_hip_scene* scene);

rise_hip_status rise_hip_scene_get_stats(const rise_hip_scene* scene, rise_hip_scene_stats* out_stats);

/* Arithmetic of the reference build whose hits the ray queries reproduce bit
 * for bit.  The hit decision is the reference's leaf test
 * (Acceleration/BVH.h:1237-1301): the float pre-filter
 * BVH::MollerTrumboreFloat (BVH.h:1111-1167) on the float normalized ray,
 * then the double certifier RayTriangleIntersection
 * (Intersection/RayTriangleIntersection.cpp:21-68) on the ray normalized by
 * Vector3Ops::Normalize (Object.cpp:414-425).
 *   RISE_HIP_ARITH_IEEE: every operation rounded on its own, in the order the
 *     reference's source writes it (the default).
 *   RISE_HIP_ARITH_X86_FMA: as the reference's Linux build compiles it
 *     (g++ -O3 -march=x86-64-v3, products contracted into fused multiply-adds:
 *     a*b - c*d -> fma(a, b, -(c*d)), a*b + c*d + e*f -> fma(e, f, fma(a, b, c*d))).
 * Applies to rise_hip_intersect_* calls made after it returns.  No
 * counterpart in the reference (a property of how it is compiled). */
typedef enum rise_hip_arith {
    RISE_HIP_ARITH_IEEE = 0,
    RISE_HIP_ARITH_X86_FMA = 1
} rise_hip_arith;

rise_hip_status rise_hip_scene_set_arithmetic(rise_hip_scene* scene, rise_hip_arith arith);

/* ------------------------------------------------------------------------ */
/* Batch ray queries                                                        */
/* ------------------------------------------------------------------------ */

/* Closest hit for n rays.  Replaces IObjectManager::IntersectRay
 * (Interfaces/IObjectManager.h:38; Managers/ObjectManager.cpp:270-318) for
 * front- and back-face hits.
 * rays_soa: device, 6*n floats: ox[n], oy[n], oz[n], dx[n], dy[n], dz[n]
 *           (world space; d need not be unit length).
 * prim_id:  device, n int32: closest triangle id, -1 on a miss; exact ties
 *           in the double-precision hit distance go to the lowest id.
 * t:        device, n floats: world distance to the hit (Object.cpp:641-643),
 *           +inf on a miss. */
rise_hip_status rise_hip_intersect_closest(rise_hip_scene* scene, const float* rays_soa, int64_t n,
                                           int32_t* prim_id, float* t, hipStream_t stream);

/* Any hit within tmax (world distance, the reference's dHowFar) for n rays.
 * Replaces IObjectManager::IntersectShadowRay (IObjectManager.h:46;
 * ObjectManager.cpp:320-373).  tmax: device, n floats, or NULL for
 * RISE_INFINITY.  occluded: device, n bytes, 1 if occluded else 0. */
rise_hip_status rise_hip_intersect_any(rise_hip_scene* scene, const float* rays_soa, const float* tmax,
                                       int64_t n, uint8_t* occluded, hipStream_t stream);

/* Face selection of IObjectManager::IntersectRay / IntersectShadowRay
 * (bHitFrontFaces, bHitBackFaces; Interfaces/IObjectManager.h:37-50): a
 * triangle is front-facing when its face normal (v1 - v0) x (v2 - v0) points
 * against the ray direction (TriangleMeshGeometryIndexedSpecializations.h:
 * 141-159); edge-on triangles count as both.  The selection applies to every
 * mesh of the scene (a RISE mesh marked double-sided ignores it). */
typedef enum rise_hip_faces {
    RISE_HIP_FACES_NONE = 0,
    RISE_HIP_FACES_FRONT = 1,
    RISE_HIP_FACES_BACK = 2,
    RISE_HIP_FACES_BOTH = 3
} rise_hip_faces;

/* The intersection record RISE's RayIntersectionGeometric holds after
 * Object::IntersectRay (Intersection/RayIntersectionGeometric.h:89-188;
 * Objects/Object.cpp:528-649): 80 bytes. */
typedef struct rise_hip_hit {
    int32_t prim_id;        /* closest triangle id (as rise_hip_intersect_closest), -1 on a miss */
    int32_t mesh;           /* index of the triangle's mesh in rise_hip_scene_upload's list, -1 on a miss */
    double range;           /* world distance (ri.range after Object.cpp:641-643), +inf on a miss */
    double alpha, beta;     /* barycentrics of the double-precision certifier (RayTriangleIntersection.cpp) */
    float point[3];         /* world hit point o + dn (t - 1e-12) (ri.ptIntersection) */
    float geom_normal[3];   /* unit face normal on the shading normal's side (ri.vGeomNormal) */
    float shading_normal[3]; /* unit shading normal (ri.vNormal): the vertex normals interpolated, or the
                                face normal */
    float uv[2];            /* texture coordinates interpolated (ri.ptCoord), 0 without them */
} rise_hip_hit;

/* rise_hip_intersect_closest with a face selection (faces: rise_hip_faces)
 * and the full intersection record per ray.  hits: device, n records.  The
 * hit decision is the same as rise_hip_intersect_closest's (bit for bit) when
 * faces is RISE_HIP_FACES_BOTH.  Replaces IObjectManager::IntersectRay
 * (Interfaces/IObjectManager.h:37-43) with its face flags. */
rise_hip_status rise_hip_intersect_closest_ex(rise_hip_scene* scene, const float* rays_soa, int64_t n, int32_t faces,
                                              rise_hip_hit* hits, hipStream_t stream);

/* rise_hip_intersect_any with a face selection.  Replaces
 * IObjectManager::IntersectShadowRay (Interfaces/IObjectManager.h:45-50) with
 * its face flags. */
rise_hip_status rise_hip_intersect_any_ex(rise_hip_scene* scene, const float* rays_soa, const float* tmax, int64_t n,
                                          int32_t faces, uint8_t* occluded, hipStream_t stream);

/* rise_hip_intersect_closest that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, results and errors as rise_hip_intersect_closest; the
 * persistent kernel's ray-fetch counter lives in a file-scope device array of
 * 256 bytes (one per GPU, held until the process ends) instead of the scene's
 * counter slots.  The counter is one 8-byte word whatever n is, so no ray
 * count exceeds the scratch.  Every call on one GPU shares that array: calls
 * that overlap on one GPU (two streams, or two host threads) race on it, and
 * the caller must serialize them. */
rise_hip_status rise_hip_intersect_closest_noalloc(rise_hip_scene* scene, const float* rays_soa, int64_t n,
                                                   int32_t* prim_id, float* t, hipStream_t stream);

/* As rise_hip_intersect_closest / _any, and also accumulates into
 * counters[0] the 8-wide nodes visited and into counters[1] the triangles
 * certified (counters: device, 2 uint64, not cleared by the call).  Used to
 * state the bytes a ray fetches; slower than the plain entry points.
 * No counterpart in the reference (its profiling counters are compile-gated,
 * Utilities/Profiling.h:163). */
rise_hip_status rise_hip_intersect_closest_counted(rise_hip_scene* scene, const float* rays_soa, int64_t n,
                                                   int32_t* prim_id, float* t, uint64_t* counters,
                                                   hipStream_t stream);
rise_hip_status rise_hip_intersect_any_counted(rise_hip_scene* scene, const float* rays_soa, const float* tmax,
                                               int64_t n, uint8_t* occluded, uint64_t* counters,
                                               hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* RISE_HIP_H */
