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
 * closed-form hits on axis-aligned triangles.
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

/* RayTriangleIntersection (RayTriangleIntersection.cpp:21-68).  Returns 1 on
 * hit and fills t/alpha/beta.  edgeA = v1 - v0, edgeB = v2 - v0. */
int rise_oracle_ray_triangle(const double o[3], const double d[3],
                             const double v0[3], const double eA[3], const double eB[3],
                             double* t, double* alpha, double* beta);

/* Vector3Ops::Normalize (VectorsOps.h:229-238). */
void rise_oracle_normalize(const double v[3], double out[3]);

/* Closest hit over all triangles, brute force.
 * tris: n_tris * 9 floats (v0.xyz, v1.xyz, v2.xyz).
 * rays: SoA, 6 arrays of n_rays floats each (ox, oy, oz, dx, dy, dz).
 * Output: prim_id (-1 on miss) and t = world distance (float), +inf on miss.
 * Optional t64 (may be NULL) receives the double world distance.
 * n_threads <= 0 means one per online CPU. */
void rise_oracle_closest_brute(const float* tris, int64_t n_tris,
                               const float* rays, int64_t n_rays,
                               int32_t* prim_id, float* t, double* t64, int n_threads);

/* Any-hit over all triangles, brute force.  tmax: per-ray world distance
 * limit (dHowFar); NULL means RISE_INFINITY.  occluded[i] = 1 iff some
 * triangle has NEARZERO < t < dHowFar2 (Object.cpp:480-493 scaling). */
void rise_oracle_any_brute(const float* tris, int64_t n_tris,
                           const float* rays, const float* tmax, int64_t n_rays,
                           uint8_t* occluded, int n_threads);

#ifdef __cplusplus
}
#endif

#endif
