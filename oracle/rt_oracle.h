/*
 * rt_oracle.h -- CPU oracle for RISE's ray-cast hot path (TEST INFRASTRUCTURE ONLY).
 *
 * This is a plain-C, FP64 restatement of the reference's closest-hit and
 * any-hit semantics, used by the test-suite, smoke() and the CPU-baseline
 * leg of bench.py as the checker.  Nothing in the product path may link or
 * call it.
 *
 * Semantics restated (reference paths relative to src/Library/):
 *   - Ray/triangle: Intersection/RayTriangleIntersection.cpp:21-68
 *     (Moller-Trumbore, NEARZERO = 1e-12 from Utilities/Math3D/Constants.h:34).
 *   - Leaf commit: Geometry/TriangleMeshGeometryIndexedSpecializations.h:116-172
 *     (edges in double, strict-closer guard).  The oracle adds the rule
 *     "exact-t ties go to the lowest triangle id" (the reference resolves
 *     ties by traversal order).
 *   - Any-hit: ...Specializations.h:476-519 (hit iff NEARZERO < t < dHowFar).
 *   - Leaf pre-filter: BVH::MollerTrumboreFloat (Acceleration/BVH.h:1111-1167)
 *     on BuildFastFilter's float records (BVH.h:431-467) and PrepRayFloat's
 *     float ray (BVH.h:1212-1225), applied before the certifier as
 *     Bvh4Leaf_* do (BVH.h:1237-1301).
 *   - Object layer for an identity transform: Objects/Object.cpp:414-425
 *     (direction normalized in double via Vector3Ops::Normalize,
 *     Utilities/Math3D/VectorsOps.h:229-238) and Object.cpp:641-643
 *     (world range = |origin + dir*(t - 1e-12) - origin|), and the
 *     dHowFar scaling of Object.cpp:480-493.
 *
 * All arithmetic is compiled with -ffp-contract=off so that every operation
 * is an individually rounded IEEE-754 double operation in the order the
 * reference writes it.
 *
 * Pinning: every function here is checked in tests/test_oracle_rt.py against
 * the reference's own KATs (tests/RayTriangleIntersectionTest.cpp), the
 * reference's BVH4 path compiled from its sources (oracle/_ref), and
 * closed-form hits on axis-aligned triangles.  The filter is pinned by rays
 * aimed within 3e-5 (barycentric) of sliver-triangle edges, where it rejects
 * hits the certifier would accept: the oracle must give the reference's ids
 * on every such ray (live against oracle/_ref, and the committed fixture
 * tests/golden/raycast_filter.npz), and without the filter it must not.
 */
#ifndef RISE_RT_ORACLE_H
#define RISE_RT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RISE_ORACLE_NEARZERO 1e-12
#define RISE_ORACLE_INFINITY 1.7976931348623158e+308

/* opts bits of the brute-force queries. */
#define RISE_ORACLE_NO_FILTER 1 /* the double certifier alone, no float pre-filter */
#define RISE_ORACLE_X86_FMA 2   /* the reference's Linux build's contraction (g++ -O3
                                   -march=x86-64-v3): a*b - c*d -> fma(a, b, -(c*d)),
                                   a*b + c*d + e*f -> fma(e, f, fma(a, b, c*d)) */
#define RISE_ORACLE_NO_FRONT 4  /* skip front faces (bHitFrontFaces false) */
#define RISE_ORACLE_NO_BACK 8   /* skip back faces (bHitBackFaces false) */
#define RISE_ORACLE_OBJ_PREHIT 16 /* object queries: model Object::IntersectRay's box pre-test
                                     (Object.cpp:513-527, :695-707) on the mesh's vertex bounds */
/* Per-mesh flags of the object queries. */
#define RISE_ORACLE_MESH_DOUBLE_SIDED 1 /* the mesh is double-sided */
#define RISE_ORACLE_MESH_NORMALS 2      /* the mesh has per-corner vertex normals */
#define RISE_ORACLE_MESH_COORDS 4       /* the mesh has per-corner texture coordinates */
#define RISE_ORACLE_DOUBLE_SIDED 32 /* every triangle belongs to a double-sided mesh
                                       (TriangleMeshGeometryIndexed.cpp:151-205): face flags do not
                                       apply and the record's normal is turned toward the ray */

/* RayTriangleIntersection (RayTriangleIntersection.cpp:21-68).  Returns 1 on
 * hit and fills t/alpha/beta.  edgeA = v1 - v0, edgeB = v2 - v0. */
int rise_oracle_ray_triangle(const double o[3], const double d[3],
                             const double v0[3], const double eA[3], const double eB[3],
                             double* t, double* alpha, double* beta);

/* Vector3Ops::Normalize (VectorsOps.h:229-238). */
void rise_oracle_normalize(const double v[3], double out[3]);

/* BVH::MollerTrumboreFloat (BVH.h:1111-1167).  Returns 1 when the triangle
 * (float p0, e1 = v1 - v0, e2 = v2 - v0) passes for the float ray (o, d)
 * against the float range `best`. */
int rise_oracle_filter_float(const float o[3], const float d[3], const float p0[3], const float e1[3],
                             const float e2[3], float best);

/* Closest hit over all triangles, brute force.
 * tris: n_tris * 9 floats (v0.xyz, v1.xyz, v2.xyz).
 * rays: SoA, 6 arrays of n_rays floats each (ox, oy, oz, dx, dy, dz).
 * Output: prim_id (-1 on miss) and t = world distance (float), +inf on miss.
 * Optional t64 (may be NULL) receives the double world distance.
 * Optional flags (may be NULL): 1 where the result depends on traversal
 * order (the winner fails the filter's distance test against its own t).
 * opts: RISE_ORACLE_* bits.  n_threads <= 0 means one per online CPU. */
void rise_oracle_closest_brute(const float* tris, int64_t n_tris,
                               const float* rays, int64_t n_rays,
                               int32_t* prim_id, float* t, double* t64, uint8_t* flags,
                               int opts, int n_threads);

/* Any-hit over all triangles, brute force.  tmax: per-ray world distance
 * limit (dHowFar); NULL means RISE_INFINITY.  occluded[i] = 1 iff some
 * triangle has NEARZERO < t < dHowFar2 (Object.cpp:480-493 scaling). */
void rise_oracle_any_brute(const float* tris, int64_t n_tris,
                           const float* rays, const float* tmax, int64_t n_rays,
                           uint8_t* occluded, int opts, int n_threads);

/* rise_oracle_closest_brute with the intersection record: alpha, beta (the
 * certifier's barycentrics), the unit face normal as RayIntersectionGeometric
 * ends up holding it for a mesh without vertex normals under an identity
 * transform (Normalize of the face normal, normalized again by
 * Object::IntersectRay) and the world hit point o + dn (t - 1e-12).
 * normal and point: 3 doubles per ray (zero on a miss). */
void rise_oracle_closest_attr(const float* tris, int64_t n_tris, const float* rays, int64_t n_rays,
                              int32_t* prim_id, double* range, double* alpha, double* beta, double* normal,
                              double* point, int opts, int n_threads);

/* rise_oracle_closest_attr and rise_oracle_any_brute on double vertices
 * (n_tris * 9 doubles), as a mesh holding Point3 vertices intersects them:
 * the certifier and the face test on the doubles, the float filter on their
 * float casts (...Specializations.h:32-47, :141-142). */
void rise_oracle_closest_attr_f64(const double* tris, int64_t n_tris, const float* rays, int64_t n_rays,
                                  int32_t* prim_id, double* range, double* alpha, double* beta, double* normal,
                                  double* point, int opts, int n_threads);
void rise_oracle_any_brute_f64(const double* tris, int64_t n_tris, const float* rays, const float* tmax,
                               int64_t n_rays, uint8_t* occluded, int opts, int n_threads);

/* One object: a mesh placed by Matrix4 transforms (members _00.._33 in order;
 * row-vector convention, translation in _30.._32).  Same layout as the
 * library's rise_hip_object_instance. */
typedef struct rise_oracle_object {
    int32_t mesh;
    int32_t reserved;
    double to_world[16];  /* m_mxFinalTrans */
    double to_object[16]; /* m_mxInvFinalTrans */
} rise_oracle_object;

/* ObjectManager::IntersectRay's linear loop over objects
 * (Managers/ObjectManager.cpp:290-310) with Object::IntersectRay per object
 * (Objects/Object.cpp:414-425 frame change, :531-533 normal, :637-643 world
 * point and range), brute force over each object's mesh.
 * tris: all meshes' triangles (9 doubles each, the vertices as the meshes
 * hold them; float meshes widened exactly), mesh m being
 * [mesh_off[m], mesh_off[m+1]); mesh_flags: NULL, or one byte per mesh of
 * RISE_ORACLE_MESH_* bits; normals (9 per triangle) and coords (6 per
 * triangle): per-corner vertex normals and texture coordinates aligned with
 * tris, read for meshes flagged so (NULL when none is).  Extra outputs
 * (NULL or given): shading (3 per ray, the world shading normal, ri.vNormal)
 * and uv (2 per ray, ri.ptCoord; 0 without coordinates).  Outputs per ray: prim_id (index within its
 * mesh, -1 on a miss), object (index, -1), range (world distance, +inf);
 * optional (NULL or all given): alpha, beta, normal (3), point (3, world);
 * optional flags: 1 where the winner fails the filter against its own t
 * (traversal-order dependent).  The box pre-test of Object.cpp:513-527 is
 * modelled only with RISE_ORACLE_OBJ_PREHIT in opts.  For a unit world
 * direction it drops only objects whose hits lie beyond the running best
 * (hits within 1e-12 of their box's entry distance aside); for |d| < 1 the
 * running best is scaled by |M^-1 d| = |d| |M^-1 d/|d||, short of its
 * object-space image, and it can drop a closer object. */
void rise_oracle_objects_closest(const double* tris, const int64_t* mesh_off, int n_meshes,
                                 const uint8_t* mesh_flags, const double* normals, const double* coords,
                                 const rise_oracle_object* objs, int n_objs, const float* rays, int64_t n_rays,
                                 int32_t* prim_id, int32_t* object, double* range, double* alpha, double* beta,
                                 double* normal, double* point, double* shading, double* uv, uint8_t* flags,
                                 int opts, int n_threads);

/* ObjectManager::IntersectShadowRay's loop (ObjectManager.cpp:320-373) with
 * Object::IntersectRay_IntersectionOnly per object (Object.cpp:652-712):
 * dHowFar scaled by |M^-1 d| per object.  tmax: NULL = RISE_INFINITY. */
void rise_oracle_objects_any(const double* tris, const int64_t* mesh_off, int n_meshes, const uint8_t* mesh_flags,
                             const rise_oracle_object* objs, int n_objs, const float* rays, const float* tmax,
                             int64_t n_rays, uint8_t* occluded, int opts, int n_threads);

#ifdef __cplusplus
}
#endif

#endif
