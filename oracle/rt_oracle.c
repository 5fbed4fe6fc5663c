/*
 * rt_oracle.c -- FP64 CPU oracle for RISE's ray-cast semantics.
 * TEST INFRASTRUCTURE ONLY: used by tests/, smoke() and bench.py's CPU
 * baseline leg as the checker; the product path never links it.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off).  See rt_oracle.h for
 * the reference lines each function restates and how each is pinned.
 */
#include "rt_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#define NZ RISE_ORACLE_NEARZERO

/* The two arithmetic models (RISE_ORACLE_X86_FMA in opts): every operation
 * rounded on its own, or the contraction g++ -O3 -march=x86-64-v3 applies to
 * the reference (a*b - c*d -> fma(a, b, -(c*d)); a*b + c*d + e*f ->
 * fma(e, f, fma(a, b, c*d)); read from its compiled BVH::IntersectRay4,
 * RayTriangleIntersection and Vector3Ops code).  This file is built with
 * -ffp-contract=off, so only the explicit fma calls fuse. */
static inline double sp64(int x, double a, double b, double c, double d)
{
    return x ? fma(a, b, -(c * d)) : a * b - c * d;
}
static inline double d364(int x, double a0, double b0, double a1, double b1, double a2, double b2)
{
    return x ? fma(a2, b2, fma(a0, b0, a1 * b1)) : a0 * b0 + a1 * b1 + a2 * b2;
}
static inline float sp32(int x, float a, float b, float c, float d)
{
    return x ? fmaf(a, b, -(c * d)) : a * b - c * d;
}
static inline float d332(int x, float a0, float b0, float a1, float b1, float a2, float b2)
{
    return x ? fmaf(a2, b2, fmaf(a0, b0, a1 * b1)) : a0 * b0 + a1 * b1 + a2 * b2;
}

/* The face normal eA x eB of RayElementIntersection
 * (TriangleMeshGeometryIndexedSpecializations.h:141-159, :196-205).  The
 * Linux build contracts its components differently from the certifier's
 * cross product (read from the compiled code): x and y as
 * fma(-(a.z), b.y, a.y*b.z) and fma(-(a.x), b.z, a.z*b.x), z as
 * fma(a.x, b.y, -(a.y*b.x)).  Float vertices make every product exact, so
 * only double vertices tell the forms apart. */
static void face_cross(int x, const double a[3], const double b[3], double n[3])
{
    if (x) {
        n[0] = fma(-a[2], b[1], a[1] * b[2]);
        n[1] = fma(-a[0], b[2], a[2] * b[0]);
        n[2] = fma(a[0], b[1], -(a[1] * b[0]));
    } else {
        n[0] = a[1] * b[2] - a[2] * b[1];
        n[1] = a[2] * b[0] - a[0] * b[2];
        n[2] = a[0] * b[1] - a[1] * b[0];
    }
}

/* p0 + (p1 - p0) a + (p2 - p0) b, one component of the vertex-normal and
 * texture-coordinate interpolation (...Specializations.h:194-216), left to
 * right; the Linux build contracts it as fma(p2 - p0, b, fma(p1 - p0, a, p0)). */
static inline double interp3(int x, double p0, double p1, double p2, double a, double b)
{
    const double d1 = p1 - p0, d2 = p2 - p0;
    return x ? fma(d2, b, fma(d1, a, p0)) : (p0 + d1 * a) + d2 * b;
}

/* The texture coordinate's interpolation (...Specializations.h:214-216):
 * p0 + ((p1 - p0) a + (p2 - p0) b), the offset summed first
 * (Point2Ops::mkPoint2 of a Vector2 sum); the Linux build contracts the sum
 * as fma(p2 - p0, b, (p1 - p0) a). */
static inline double interp_uv(int x, double p0, double p1, double p2, double a, double b)
{
    const double d1 = p1 - p0, d2 = p2 - p0;
    return x ? p0 + fma(d2, b, d1 * a) : p0 + (d1 * a + d2 * b);
}

/* Vector3Ops::Normalize (Utilities/Math3D/VectorsOps.h:229-238):
 * mag = sqrt(x*x + y*y + z*z); if mag > 0, v * (1/mag). */
static void normalize_x(int x, const double v[3], double out[3]);
void rise_oracle_normalize(const double v[3], double out[3]) { normalize_x(0, v, out); }

static void normalize_x(int x, const double v[3], double out[3])
{
    const double mag = sqrt(d364(x, v[0], v[0], v[1], v[1], v[2], v[2]));
    if (mag > 0.0) {
        const double ov = 1.0 / mag;
        out[0] = v[0] * ov;
        out[1] = v[1] * ov;
        out[2] = v[2] * ov;
    } else {
        out[0] = v[0];
        out[1] = v[1];
        out[2] = v[2];
    }
}

/* RayTriangleIntersection (Intersection/RayTriangleIntersection.cpp:21-68).
 * Cross = (a.y*b.z-a.z*b.y, a.z*b.x-a.x*b.z, a.x*b.y-a.y*b.x) and
 * Dot = a.x*b.x + a.y*b.y + a.z*b.z (VectorsOps.h:191-200), evaluated left
 * to right. */
static int ray_triangle_x(int x, const double o[3], const double d[3], const double v0[3], const double eA[3],
                          const double eB[3], double* t, double* alpha, double* beta)
{
    const double p0 = sp64(x, d[1], eB[2], d[2], eB[1]);
    const double p1 = sp64(x, d[2], eB[0], d[0], eB[2]);
    const double p2 = sp64(x, d[0], eB[1], d[1], eB[0]);
    const double det = d364(x, eA[0], p0, eA[1], p1, eA[2], p2);
    if (det > -NZ && det < NZ)
        return 0;
    const double oodet = 1.0 / det;
    const double t0 = o[0] - v0[0];
    const double t1 = o[1] - v0[1];
    const double t2 = o[2] - v0[2];
    const double a = d364(x, t0, p0, t1, p1, t2, p2) * oodet;
    if ((a < 0.0 - NZ) || (a > 1.0 + NZ))
        return 0;
    const double q0 = sp64(x, t1, eA[2], t2, eA[1]);
    const double q1 = sp64(x, t2, eA[0], t0, eA[2]);
    const double q2 = sp64(x, t0, eA[1], t1, eA[0]);
    const double b = d364(x, d[0], q0, d[1], q1, d[2], q2) * oodet;
    if ((b < 0.0 - NZ) || ((a + b) > 1.0 + NZ))
        return 0;
    const double r = d364(x, eB[0], q0, eB[1], q1, eB[2], q2) * oodet;
    if (r >= NZ) {
        *t = r;
        if (alpha) *alpha = a;
        if (beta) *beta = b;
        return 1;
    }
    return 0;
}

int rise_oracle_ray_triangle(const double o[3], const double d[3],
                             const double v0[3], const double eA[3], const double eB[3],
                             double* t, double* alpha, double* beta)
{
    return ray_triangle_x(0, o, d, v0, eA, eB, t, alpha, beta);
}

/* BVH::MollerTrumboreFloat (Acceleration/BVH.h:1111-1167): the float
 * pre-filter every leaf triangle passes before the double certifier
 * (Bvh4Leaf_*, BVH.h:1237-1301).  p0, e1 = v1 - v0, e2 = v2 - v0 are float
 * (BuildFastFilter, BVH.h:431-467); o and d are the float origin and the
 * float normalized direction (PrepRayFloat, BVH.h:1212-1225); best is
 * (float) of the current range.  Every operation is one IEEE float
 * operation in the reference's order (this file is built without
 * contraction).  Returns 1 when the triangle may be hit. */
static int filter_x(int x, const float o[3], const float d[3], const float p0[3], const float e1[3],
                    const float e2[3], float best)
{
    const float eps = 1e-5f;
    const float pv0 = sp32(x, d[1], e2[2], d[2], e2[1]);
    const float pv1 = sp32(x, d[2], e2[0], d[0], e2[2]);
    const float pv2 = sp32(x, d[0], e2[1], d[1], e2[0]);
    const float det = d332(x, e1[0], pv0, e1[1], pv1, e1[2], pv2);
    if (fabsf(det) < 1e-20f)
        return 0;
    const float inv = 1.0f / det;
    const float tv0 = o[0] - p0[0], tv1 = o[1] - p0[1], tv2 = o[2] - p0[2];
    const float u = d332(x, tv0, pv0, tv1, pv1, tv2, pv2) * inv;
    if (u < -eps || u > 1.0f + eps)
        return 0;
    const float qv0 = sp32(x, tv1, e1[2], tv2, e1[1]);
    const float qv1 = sp32(x, tv2, e1[0], tv0, e1[2]);
    const float qv2 = sp32(x, tv0, e1[1], tv1, e1[0]);
    const float v = d332(x, d[0], qv0, d[1], qv1, d[2], qv2) * inv;
    if (v < -eps || u + v > 1.0f + eps)
        return 0;
    const float t = d332(x, e2[0], qv0, e2[1], qv1, e2[2], qv2) * inv;
    if (t < -eps || t > best + eps)
        return 0;
    return 1;
}

int rise_oracle_filter_float(const float o[3], const float d[3], const float p0[3], const float e1[3],
                             const float e2[3], float best)
{
    return filter_x(0, o, d, p0, e1, e2, best);
}

typedef struct {
    double v0[3], eA[3], eB[3];
    float p0[3], e1[3], e2[3];
} tri64;

/* Edges in double from the (float-valued) vertices, as
 * TriangleMeshGeometryIndexedSpecializations.h:141-142 computes them, and
 * the float filter record of BuildFastFilter (BVH.h:457-459). */
static tri64* expand_tris(const float* tris, int64_t n)
{
    tri64* out = (tri64*)malloc(sizeof(tri64) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; ++i) {
        const float* f = tris + 9 * i;
        for (int k = 0; k < 3; ++k) {
            const double a = f[k], b = f[3 + k], c = f[6 + k];
            out[i].v0[k] = a;
            out[i].eA[k] = b - a;
            out[i].eB[k] = c - a;
            out[i].p0[k] = f[k];
            out[i].e1[k] = f[3 + k] - f[k];
            out[i].e2[k] = f[6 + k] - f[k];
        }
    }
    return out;
}

/* The same from double vertices, as a mesh that holds them does: the
 * certifier's v0 and edges in double from the doubles, the filter's record
 * from their float casts (GetFloatTriangleVertices,
 * TriangleMeshGeometryIndexedSpecializations.h:32-47). */
static tri64* expand_tris_f64(const double* tris, int64_t n)
{
    tri64* out = (tri64*)malloc(sizeof(tri64) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; ++i) {
        const double* f = tris + 9 * i;
        for (int k = 0; k < 3; ++k) {
            const double a = f[k], b = f[3 + k], c = f[6 + k];
            const float af = (float)a, bf = (float)b, cf = (float)c;
            out[i].v0[k] = a;
            out[i].eA[k] = b - a;
            out[i].eB[k] = c - a;
            out[i].p0[k] = af;
            out[i].e1[k] = bf - af;
            out[i].e2[k] = cf - af;
        }
    }
    return out;
}

typedef struct {
    const tri64* tris;
    int64_t n_tris;
    const float* rays;
    const float* tmax;
    int64_t n_rays;
    int32_t* prim_id;
    float* t;
    double* t64;
    uint8_t* flags;
    uint8_t* occluded;
    double *alpha, *beta, *normal, *point;
    int opts;
    int64_t next;
    pthread_mutex_t lock;
} job;

static int64_t grab(job* j, int64_t chunk, int64_t* end)
{
    pthread_mutex_lock(&j->lock);
    const int64_t s = j->next;
    j->next += chunk;
    pthread_mutex_unlock(&j->lock);
    *end = s + chunk < j->n_rays ? s + chunk : j->n_rays;
    return s;
}

static void load_ray(const float* rays, int64_t n, int64_t i, double o[3], double d[3])
{
    o[0] = rays[0 * n + i];
    o[1] = rays[1 * n + i];
    o[2] = rays[2 * n + i];
    d[0] = rays[3 * n + i];
    d[1] = rays[4 * n + i];
    d[2] = rays[5 * n + i];
}

/* PrepRayFloat (BVH.h:1212-1225): the float origin and float direction of
 * the object-space ray (normalized in double first). */
static void float_ray(const double o[3], const double dn[3], float fo[3], float fd[3])
{
    for (int k = 0; k < 3; ++k) {
        fo[k] = (float)o[k];
        fd[k] = (float)dn[k];
    }
}

/* Face selection of TriangleMeshGeometryIndexed::RayElementIntersection
 * (...Specializations.h:141-159): vFaceNormal = Cross(vEdgeA, vEdgeB) against
 * the normalized object-space direction; Dot < 0 is a front face, skipped
 * when bHitFrontFaces is false, Dot > 0 a back face, skipped when
 * bHitBackFaces is false.  Returns 1 when the triangle is considered. */
static int face_selected(int x, int opts, const tri64* t, const double dn[3])
{
    if (!(opts & (RISE_ORACLE_NO_FRONT | RISE_ORACLE_NO_BACK))) return 1;
    if (opts & RISE_ORACLE_DOUBLE_SIDED) return 1; /* TriangleMeshGeometryIndexed.cpp:169 */
    double n[3];
    face_cross(x, t->eA, t->eB, n);
    const double dd = d364(x, n[0], dn[0], n[1], dn[1], n[2], dn[2]);
    if ((opts & RISE_ORACLE_NO_FRONT) && dd < 0.0) return 0;
    if ((opts & RISE_ORACLE_NO_BACK) && dd > 0.0) return 0;
    return 1;
}

static void* closest_worker(void* arg)
{
    job* j = (job*)arg;
    for (;;) {
        int64_t e;
        const int64_t s = grab(j, 64, &e);
        if (s >= j->n_rays) break;
        for (int64_t i = s; i < e; ++i) {
            double o[3], dw[3], dn[3];
            load_ray(j->rays, j->n_rays, i, o, dw);
            /* Object.cpp:414-425: object-space direction normalized in
             * double (identity transform leaves o and d unchanged). */
            const int x = (j->opts & RISE_ORACLE_X86_FMA) != 0;
            const int nof = (j->opts & RISE_ORACLE_NO_FILTER) != 0;
            normalize_x(x, dw, dn);
            float fo[3], fd[3];
            float_ray(o, dn, fo, fd);
            double best = RISE_ORACLE_INFINITY;
            int64_t best_id = -1;
            for (int64_t k = 0; k < j->n_tris; ++k) {
                double tt;
                /* The filter's distance test runs against +inf here; the
                 * winner's own test is checked below. */
                if (!nof && !filter_x(x, fo, fd, j->tris[k].p0, j->tris[k].e1, j->tris[k].e2, INFINITY))
                    continue;
                if (!face_selected(x, j->opts, &j->tris[k], dn)) continue;
                if (ray_triangle_x(x, o, dn, j->tris[k].v0, j->tris[k].eA, j->tris[k].eB, &tt, NULL, NULL)) {
                    /* Strict-closer commit (...Specializations.h:172); k
                     * ascends, so an exact tie keeps the lowest id. */
                    if (tt < best) {
                        best = tt;
                        best_id = k;
                    }
                }
            }
            if (j->flags) {
                /* In a traversal the filter compares the float t with the best
                 * range found so far, which is never below the winner's own
                 * t.  The winner therefore passes in every traversal order iff
                 * it passes against its own t; otherwise the result depends on
                 * the order and the ray is flagged. */
                const tri64* w = best_id >= 0 ? &j->tris[best_id] : NULL;
                j->flags[i] = (w && !nof && !filter_x(x, fo, fd, w->p0, w->e1, w->e2, (float)best)) ? 1 : 0;
            }
            if (best_id >= 0) {
                /* Object.cpp:641-643: back off by SURFACE_INTERSEC_ERROR,
                 * then the world range |p - origin| (PointAtLength =
                 * origin + dir*s, Utilities/Ray.h:91-94). */
                const double s2 = best - 1e-12;
                const double px = x ? fma(dn[0], s2, o[0]) : o[0] + dn[0] * s2;
                const double py = x ? fma(dn[1], s2, o[1]) : o[1] + dn[1] * s2;
                const double pz = x ? fma(dn[2], s2, o[2]) : o[2] + dn[2] * s2;
                const double vx = px - o[0], vy = py - o[1], vz = pz - o[2];
                const double range = sqrt(d364(x, vx, vx, vy, vy, vz, vz));
                j->prim_id[i] = (int32_t)best_id;
                if (j->t) j->t[i] = (float)range;
                if (j->t64) j->t64[i] = range;
                if (j->alpha) {
                    /* RayElementIntersection (...Specializations.h:172-205) and
                     * Object::IntersectRay (Object.cpp:531-533, 637-640). */
                    const tri64* w = &j->tris[best_id];
                    double tt, a, b;
                    ray_triangle_x(x, o, dn, w->v0, w->eA, w->eB, &tt, &a, &b);
                    j->alpha[i] = a;
                    j->beta[i] = b;
                    double fn[3];
                    face_cross(x, w->eA, w->eB, fn);
                    double n1[3];
                    normalize_x(x, fn, n1);
                    /* TriangleMeshGeometryIndexed.cpp:170-188 */
                    if ((j->opts & RISE_ORACLE_DOUBLE_SIDED) && d364(x, n1[0], dn[0], n1[1], dn[1], n1[2], dn[2]) > 0.0)
                        for (int c = 0; c < 3; ++c) n1[c] = -n1[c];
                    normalize_x(x, n1, j->normal + 3 * i);
                    j->point[3 * i + 0] = px;
                    j->point[3 * i + 1] = py;
                    j->point[3 * i + 2] = pz;
                }
            } else {
                j->prim_id[i] = -1;
                if (j->t) j->t[i] = INFINITY;
                if (j->t64) j->t64[i] = INFINITY;
                if (j->alpha) {
                    j->alpha[i] = j->beta[i] = 0.0;
                    for (int k = 0; k < 3; ++k) j->normal[3 * i + k] = j->point[3 * i + k] = 0.0;
                }
            }
        }
    }
    return NULL;
}

static void* any_worker(void* arg)
{
    job* j = (job*)arg;
    for (;;) {
        int64_t e;
        const int64_t s = grab(j, 64, &e);
        if (s >= j->n_rays) break;
        for (int64_t i = s; i < e; ++i) {
            double o[3], dw[3], dn[3];
            load_ray(j->rays, j->n_rays, i, o, dw);
            const int x = (j->opts & RISE_ORACLE_X86_FMA) != 0;
            const int nof = (j->opts & RISE_ORACLE_NO_FILTER) != 0;
            normalize_x(x, dw, dn);
            /* Object.cpp:472-493: factor = |M^-1 d| (identity -> |d|);
             * dHowFar is scaled unless it is the RISE_INFINITY sentinel and
             * factor >= 1. */
            const double mag = sqrt(d364(x, dw[0], dw[0], dw[1], dw[1], dw[2], dw[2]));
            const double factor = (mag > NZ) ? mag : 1.0;
            const double how_far = j->tmax ? (double)j->tmax[i] : RISE_ORACLE_INFINITY;
            double how_far2 = how_far;
            if ((how_far != RISE_ORACLE_INFINITY) || (factor < 1.0))
                how_far2 = factor * how_far;
            float fo[3], fd[3];
            float_ray(o, dn, fo, fd);
            /* Bvh4Leaf_IntersectionOnly (BVH.h:1278-1301): the filter runs
             * against (float)dHowFar2. */
            const float fbest = (float)how_far2;
            uint8_t occ = 0;
            for (int64_t k = 0; k < j->n_tris && !occ; ++k) {
                double tt;
                if (!nof && !filter_x(x, fo, fd, j->tris[k].p0, j->tris[k].e1, j->tris[k].e2, fbest))
                    continue;
                if (!face_selected(x, j->opts, &j->tris[k], dn)) continue;
                /* ...Specializations.h:517-519 */
                if (ray_triangle_x(x, o, dn, j->tris[k].v0, j->tris[k].eA, j->tris[k].eB, &tt, NULL, NULL) &&
                    tt > NZ && tt < how_far2)
                    occ = 1;
            }
            j->occluded[i] = occ;
        }
    }
    return NULL;
}

static void run(job* j, void* (*fn)(void*), int n_threads)
{
    if (n_threads <= 0) {
        long c = sysconf(_SC_NPROCESSORS_ONLN);
        n_threads = c > 0 ? (int)c : 1;
    }
    if (n_threads > 256) n_threads = 256;
    pthread_t th[256];
    pthread_mutex_init(&j->lock, NULL);
    j->next = 0;
    for (int i = 0; i < n_threads; ++i) pthread_create(&th[i], NULL, fn, j);
    for (int i = 0; i < n_threads; ++i) pthread_join(th[i], NULL);
    pthread_mutex_destroy(&j->lock);
}

void rise_oracle_closest_brute(const float* tris, int64_t n_tris,
                               const float* rays, int64_t n_rays,
                               int32_t* prim_id, float* t, double* t64, uint8_t* flags,
                               int opts, int n_threads)
{
    job j;
    memset(&j, 0, sizeof(j));
    tri64* tt = expand_tris(tris, n_tris);
    j.tris = tt;
    j.n_tris = n_tris;
    j.rays = rays;
    j.n_rays = n_rays;
    j.prim_id = prim_id;
    j.t = t;
    j.t64 = t64;
    j.flags = flags;
    j.opts = opts;
    run(&j, closest_worker, n_threads);
    free(tt);
}

void rise_oracle_any_brute(const float* tris, int64_t n_tris,
                           const float* rays, const float* tmax, int64_t n_rays,
                           uint8_t* occluded, int opts, int n_threads)
{
    job j;
    memset(&j, 0, sizeof(j));
    tri64* tt = expand_tris(tris, n_tris);
    j.tris = tt;
    j.n_tris = n_tris;
    j.rays = rays;
    j.tmax = tmax;
    j.n_rays = n_rays;
    j.occluded = occluded;
    j.opts = opts;
    run(&j, any_worker, n_threads);
    free(tt);
}

void rise_oracle_closest_attr_f64(const double* tris, int64_t n_tris, const float* rays, int64_t n_rays,
                                  int32_t* prim_id, double* range, double* alpha, double* beta, double* normal,
                                  double* point, int opts, int n_threads)
{
    job j;
    memset(&j, 0, sizeof(j));
    tri64* tt = expand_tris_f64(tris, n_tris);
    j.tris = tt;
    j.n_tris = n_tris;
    j.rays = rays;
    j.n_rays = n_rays;
    j.prim_id = prim_id;
    j.t64 = range;
    j.alpha = alpha;
    j.beta = beta;
    j.normal = normal;
    j.point = point;
    j.opts = opts;
    run(&j, closest_worker, n_threads);
    free(tt);
}

void rise_oracle_any_brute_f64(const double* tris, int64_t n_tris, const float* rays, const float* tmax,
                               int64_t n_rays, uint8_t* occluded, int opts, int n_threads)
{
    job j;
    memset(&j, 0, sizeof(j));
    tri64* tt = expand_tris_f64(tris, n_tris);
    j.tris = tt;
    j.n_tris = n_tris;
    j.rays = rays;
    j.tmax = tmax;
    j.n_rays = n_rays;
    j.occluded = occluded;
    j.opts = opts;
    run(&j, any_worker, n_threads);
    free(tt);
}

void rise_oracle_closest_attr(const float* tris, int64_t n_tris, const float* rays, int64_t n_rays,
                              int32_t* prim_id, double* range, double* alpha, double* beta, double* normal,
                              double* point, int opts, int n_threads)
{
    job j;
    memset(&j, 0, sizeof(j));
    tri64* tt = expand_tris(tris, n_tris);
    j.tris = tt;
    j.n_tris = n_tris;
    j.rays = rays;
    j.n_rays = n_rays;
    j.prim_id = prim_id;
    j.t64 = range;
    j.alpha = alpha;
    j.beta = beta;
    j.normal = normal;
    j.point = point;
    j.opts = opts;
    run(&j, closest_worker, n_threads);
    free(tt);
}

/* ------------------------------------------------------------------------ */
/* Scenes of objects with transforms                                        */
/* ------------------------------------------------------------------------ */

/* Point3Ops::Transform (Utilities/Math3D/PointsOps.h:126-133) and
 * Vector3Ops::Transform (VectorsOps.h:241-248): Matrix4 members in order
 * _00.._33, m[4 i + j] = m._ij, out_j = m._0j x + m._1j y + m._2j z (+ m._3j
 * for a point), evaluated left to right. */
static void xpoint(int x, const double* m, const double p[3], double out[3])
{
    for (int j = 0; j < 3; ++j) out[j] = d364(x, m[j], p[0], m[4 + j], p[1], m[8 + j], p[2]) + m[12 + j];
}
static void xvector(int x, const double* m, const double v[3], double out[3])
{
    for (int j = 0; j < 3; ++j) out[j] = d364(x, m[j], v[0], m[4 + j], v[1], m[8 + j], v[2]);
}

typedef struct {
    const tri64* tris;          /* all meshes, concatenated */
    const int64_t* mesh_off;    /* n_meshes + 1 offsets into tris */
    const uint8_t* mesh_flags;  /* per mesh (RISE_ORACLE_MESH_*), or NULL */
    const double* normals;      /* per-corner vertex normals, 9 per triangle, or NULL */
    const double* coords;       /* per-corner texture coordinates, 6 per triangle, or NULL */
    double *shading, *uv;
    const rise_oracle_object* objs;
    int n_objs;
    const float* rays;
    const float* tmax;
    int64_t n_rays;
    int32_t *prim_id, *object;
    double *range, *alpha, *beta, *normal, *point;
    uint8_t* flags;
    uint8_t* occluded;
    int opts;
    int64_t next;
    pthread_mutex_t lock;
} ojob;

static int64_t ograb(ojob* j, int64_t chunk, int64_t* end)
{
    pthread_mutex_lock(&j->lock);
    const int64_t s = j->next;
    j->next += chunk;
    pthread_mutex_unlock(&j->lock);
    *end = s + chunk < j->n_rays ? s + chunk : j->n_rays;
    return s;
}

/* Object::IntersectRay's frame change (Object.cpp:414-425, :472-485):
 * o' = M^-1 o, dl = M^-1 d (d as given), factor = |dl| (1 when not above
 * NEARZERO), dn' = Normalize(dl). */
static void object_ray(int x, const rise_oracle_object* ob, const double o[3], const double d[3], double op[3],
                       double dn[3], double* factor)
{
    double dl[3];
    xpoint(x, ob->to_object, o, op);
    xvector(x, ob->to_object, d, dl);
    const double mag = sqrt(d364(x, dl[0], dl[0], dl[1], dl[1], dl[2], dl[2]));
    *factor = mag > NZ ? mag : 1.0;
    normalize_x(x, dl, dn);
}

/* RayBoxIntersection (Intersection/RayBoxIntersection.cpp) on the ray
 * (o, d) with Ray::RecomputeInvDir's inverse (Utilities/Ray.h:36-45: 1/d,
 * or +-1e30 for |d| <= 1e-30).  Returns bHit; range / range2 as dRange /
 * dRange2 (entry and exit, or exit and the negative entry when the origin
 * is inside). */
static int ray_box(const double o[3], const double d[3], const double lo[3], const double hi[3], double* range,
                   double* range2)
{
    double inv[3];
    int sg[3];
    for (int k = 0; k < 3; ++k) {
        inv[k] = (d[k] > 1e-30 || d[k] < -1e-30) ? 1.0 / d[k] : (d[k] >= 0 ? 1e30 : -1e30);
        sg[k] = inv[k] < 0 ? 1 : 0;
    }
    const double bx[2] = {lo[0], hi[0]}, by[2] = {lo[1], hi[1]}, bz[2] = {lo[2], hi[2]};
    double tmin = (bx[sg[0]] - o[0]) * inv[0];
    double tmax = (bx[1 - sg[0]] - o[0]) * inv[0];
    const double tymin = (by[sg[1]] - o[1]) * inv[1];
    const double tymax = (by[1 - sg[1]] - o[1]) * inv[1];
    if (tmin > tymax || tymin > tmax) return 0;
    if (tymin > tmin) tmin = tymin;
    if (tymax < tmax) tmax = tymax;
    const double tzmin = (bz[sg[2]] - o[2]) * inv[2];
    const double tzmax = (bz[1 - sg[2]] - o[2]) * inv[2];
    if (tmin > tzmax || tzmin > tmax) return 0;
    if (tzmin > tmin) tmin = tzmin;
    if (tzmax < tmax) tmax = tzmax;
    if (tmin < 0 && tmax < 0) return 0;
    if (tmin > 0) {
        *range = tmin;
        *range2 = tmax;
    } else {
        *range = tmax;
        *range2 = tmin;
    }
    return 1;
}

/* The mesh's GenerateBoundingBox: its BVH's overall box, the bounds of all
 * its vertices (TriangleMeshGeometryIndexed.cpp:503-517, 576-583). */
static void mesh_box(const tri64* tris, int64_t b0, int64_t b1, double lo[3], double hi[3])
{
    for (int k = 0; k < 3; ++k) {
        lo[k] = RISE_ORACLE_INFINITY;
        hi[k] = -RISE_ORACLE_INFINITY;
    }
    for (int64_t q = b0; q < b1; ++q)
        for (int k = 0; k < 3; ++k) {
            const double v[3] = {tris[q].v0[k], tris[q].v0[k] + tris[q].eA[k], tris[q].v0[k] + tris[q].eB[k]};
            for (int c = 0; c < 3; ++c) {
                if (v[c] < lo[k]) lo[k] = v[c];
                if (v[c] > hi[k]) hi[k] = v[c];
            }
        }
}

static void* objects_closest_worker(void* arg)
{
    ojob* j = (ojob*)arg;
    const int x = (j->opts & RISE_ORACLE_X86_FMA) != 0;
    const int nof = (j->opts & RISE_ORACLE_NO_FILTER) != 0;
    for (;;) {
        int64_t e;
        const int64_t s = ograb(j, 16, &e);
        if (s >= j->n_rays) break;
        for (int64_t i = s; i < e; ++i) {
            double o[3], d[3];
            load_ray(j->rays, j->n_rays, i, o, d);
            /* ObjectManager::IntersectRay's loop (ObjectManager.cpp:290-310):
             * each object's own closest hit, kept when its world range is
             * strictly below the best so far. */
            double best_w = RISE_ORACLE_INFINITY;
            int32_t best_obj = -1, best_prim = -1;
            double best_pw[3] = {0, 0, 0}, best_a = 0, best_b = 0, best_n[3] = {0, 0, 0};
            double best_s[3] = {0, 0, 0}, best_uv[2] = {0, 0};
            uint8_t flag = 0;
            for (int k = 0; k < j->n_objs; ++k) {
                const rise_oracle_object* ob = &j->objs[k];
                const int mflags = j->mesh_flags ? j->mesh_flags[ob->mesh] : 0;
                const int mopts = j->opts | ((mflags & RISE_ORACLE_MESH_DOUBLE_SIDED) ? RISE_ORACLE_DOUBLE_SIDED : 0);
                double op[3], dn[3], factor;
                object_ray(x, ob, o, d, op, dn, &factor);
                float fo[3], fd[3];
                float_ray(op, dn, fo, fd);
                const int64_t b0 = j->mesh_off[ob->mesh], b1 = j->mesh_off[ob->mesh + 1];
                if (j->opts & RISE_ORACLE_OBJ_PREHIT) {
                    /* Object.cpp:472-527: the running best scaled by factor,
                     * then the box pre-test on the object-space ray. */
                    double how_far2 = best_w, lo[3], hi[3], r1, r2;
                    if ((best_w != RISE_ORACLE_INFINITY) || (factor < 1.0)) how_far2 = factor * best_w;
                    mesh_box(j->tris, b0, b1, lo, hi);
                    if (!ray_box(op, dn, lo, hi, &r1, &r2)) continue;
                    if (!(r2 < 0) && r1 > how_far2) continue;
                }
                double bt = RISE_ORACLE_INFINITY;
                int64_t bk = -1;
                for (int64_t q = b0; q < b1; ++q) {
                    double tt;
                    const tri64* t = &j->tris[q];
                    if (!nof && !filter_x(x, fo, fd, t->p0, t->e1, t->e2, INFINITY)) continue;
                    if (!face_selected(x, mopts, t, dn)) continue;
                    if (ray_triangle_x(x, op, dn, t->v0, t->eA, t->eB, &tt, NULL, NULL) && tt < bt) {
                        bt = tt;
                        bk = q;
                    }
                }
                if (bk < 0) continue;
                /* Object.cpp:637-643: back off, to world space, world range. */
                const double s2 = bt - 1e-12;
                double po[3], pw[3];
                for (int c = 0; c < 3; ++c) po[c] = x ? fma(dn[c], s2, op[c]) : op[c] + dn[c] * s2;
                xpoint(x, ob->to_world, po, pw);
                const double vx = pw[0] - o[0], vy = pw[1] - o[1], vz = pw[2] - o[2];
                const double w = sqrt(d364(x, vx, vx, vy, vy, vz, vz));
                if (!(w < best_w)) continue;
                const tri64* t = &j->tris[bk];
                best_w = w;
                best_obj = k;
                best_prim = (int32_t)(bk - b0);
                for (int c = 0; c < 3; ++c) best_pw[c] = pw[c];
                flag = (!nof && !filter_x(x, fo, fd, t->p0, t->e1, t->e2, (float)bt)) ? 1 : 0;
                if (j->alpha) {
                    /* Barycentrics of the certifier; the shading normal (the
                     * vertex normals interpolated, or the face normal) and the
                     * face normal normalized and turned to the shading one's
                     * side, the texture coordinates interpolated
                     * (...Specializations.h:191-216); both normals turned
                     * toward the ray on a double-sided mesh
                     * (TriangleMeshGeometryIndexed.cpp:170-188), then taken
                     * by M^-T and normalized (Object.cpp:531-533, with
                     * m_mxInvTranspose = Transpose(M^-1), :827). */
                    double tt, a, b;
                    ray_triangle_x(x, op, dn, t->v0, t->eA, t->eB, &tt, &a, &b);
                    best_a = a;
                    best_b = b;
                    double fn[3], sn[3], gn[3];
                    face_cross(x, t->eA, t->eB, fn);
                    if ((mflags & RISE_ORACLE_MESH_NORMALS) && j->normals) {
                        const double* nn = j->normals + 9 * bk;
                        for (int c = 0; c < 3; ++c) sn[c] = interp3(x, nn[c], nn[3 + c], nn[6 + c], a, b);
                    } else {
                        for (int c = 0; c < 3; ++c) sn[c] = fn[c];
                    }
                    normalize_x(x, fn, gn);
                    if (d364(x, gn[0], sn[0], gn[1], sn[1], gn[2], sn[2]) < 0.0)
                        for (int c = 0; c < 3; ++c) gn[c] = -gn[c];
                    if ((mflags & RISE_ORACLE_MESH_COORDS) && j->coords) {
                        const double* cc = j->coords + 6 * bk;
                        for (int c = 0; c < 2; ++c) best_uv[c] = interp_uv(x, cc[c], cc[2 + c], cc[4 + c], a, b);
                    } else {
                        best_uv[0] = best_uv[1] = 0.0;
                    }
                    if (mopts & RISE_ORACLE_DOUBLE_SIDED) {
                        if (d364(x, sn[0], dn[0], sn[1], dn[1], sn[2], dn[2]) > 0.0)
                            for (int c = 0; c < 3; ++c) sn[c] = -sn[c];
                        if (d364(x, gn[0], dn[0], gn[1], dn[1], gn[2], dn[2]) > 0.0)
                            for (int c = 0; c < 3; ++c) gn[c] = -gn[c];
                    }
                    const double* mi = ob->to_object;
                    double nw[3], sw[3];
                    for (int c = 0; c < 3; ++c) {
                        nw[c] = d364(x, mi[4 * c], gn[0], mi[4 * c + 1], gn[1], mi[4 * c + 2], gn[2]);
                        sw[c] = d364(x, mi[4 * c], sn[0], mi[4 * c + 1], sn[1], mi[4 * c + 2], sn[2]);
                    }
                    normalize_x(x, nw, best_n);
                    normalize_x(x, sw, best_s);
                }
            }
            j->prim_id[i] = best_prim;
            j->object[i] = best_obj;
            j->range[i] = best_obj >= 0 ? best_w : INFINITY;
            if (j->flags) j->flags[i] = flag;
            if (j->alpha) {
                j->alpha[i] = best_a;
                j->beta[i] = best_b;
                for (int c = 0; c < 3; ++c) {
                    j->normal[3 * i + c] = best_n[c];
                    j->point[3 * i + c] = best_pw[c];
                }
            }
            if (j->shading)
                for (int c = 0; c < 3; ++c) j->shading[3 * i + c] = best_s[c];
            if (j->uv)
                for (int c = 0; c < 2; ++c) j->uv[2 * i + c] = best_uv[c];
        }
    }
    return NULL;
}

static void* objects_any_worker(void* arg)
{
    ojob* j = (ojob*)arg;
    const int x = (j->opts & RISE_ORACLE_X86_FMA) != 0;
    const int nof = (j->opts & RISE_ORACLE_NO_FILTER) != 0;
    for (;;) {
        int64_t e;
        const int64_t s = ograb(j, 16, &e);
        if (s >= j->n_rays) break;
        for (int64_t i = s; i < e; ++i) {
            double o[3], d[3];
            load_ray(j->rays, j->n_rays, i, o, d);
            const double how_far = j->tmax ? (double)j->tmax[i] : RISE_ORACLE_INFINITY;
            uint8_t occ = 0;
            /* ObjectManager::IntersectShadowRay (ObjectManager.cpp:320-373):
             * any object occluding within its own scaled limit. */
            for (int k = 0; k < j->n_objs && !occ; ++k) {
                const rise_oracle_object* ob = &j->objs[k];
                const int mflags = j->mesh_flags ? j->mesh_flags[ob->mesh] : 0;
                const int mopts = j->opts | ((mflags & RISE_ORACLE_MESH_DOUBLE_SIDED) ? RISE_ORACLE_DOUBLE_SIDED : 0);
                double op[3], dn[3], factor;
                object_ray(x, ob, o, d, op, dn, &factor);
                /* Object.cpp:688-693 */
                double how_far2 = how_far;
                if ((how_far != RISE_ORACLE_INFINITY) || (factor < 1.0)) how_far2 = factor * how_far;
                float fo[3], fd[3];
                float_ray(op, dn, fo, fd);
                const float fbest = (float)how_far2;
                const int64_t b0 = j->mesh_off[ob->mesh], b1 = j->mesh_off[ob->mesh + 1];
                if (j->opts & RISE_ORACLE_OBJ_PREHIT) {
                    /* Object.cpp:695-707: box pre-test; beyond the limit only
                     * an origin strictly inside the box goes on. */
                    double lo[3], hi[3], r1, r2;
                    mesh_box(j->tris, b0, b1, lo, hi);
                    if (!ray_box(op, dn, lo, hi, &r1, &r2)) continue;
                    if (r1 > how_far2 && !(op[0] > lo[0] && op[0] < hi[0] && op[1] > lo[1] && op[1] < hi[1] &&
                                           op[2] > lo[2] && op[2] < hi[2]))
                        continue;
                }
                for (int64_t q = b0; q < b1 && !occ; ++q) {
                    double tt;
                    const tri64* t = &j->tris[q];
                    if (!nof && !filter_x(x, fo, fd, t->p0, t->e1, t->e2, fbest)) continue;
                    if (!face_selected(x, mopts, t, dn)) continue;
                    if (ray_triangle_x(x, op, dn, t->v0, t->eA, t->eB, &tt, NULL, NULL) && tt > NZ &&
                        tt < how_far2)
                        occ = 1;
                }
            }
            j->occluded[i] = occ;
        }
    }
    return NULL;
}

static void orun(ojob* j, void* (*fn)(void*), int n_threads)
{
    if (n_threads <= 0) {
        long c = sysconf(_SC_NPROCESSORS_ONLN);
        n_threads = c > 0 ? (int)c : 1;
    }
    if (n_threads > 256) n_threads = 256;
    pthread_t th[256];
    pthread_mutex_init(&j->lock, NULL);
    j->next = 0;
    for (int i = 0; i < n_threads; ++i) pthread_create(&th[i], NULL, fn, j);
    for (int i = 0; i < n_threads; ++i) pthread_join(th[i], NULL);
    pthread_mutex_destroy(&j->lock);
}

void rise_oracle_objects_closest(const double* tris, const int64_t* mesh_off, int n_meshes,
                                 const uint8_t* mesh_flags, const double* normals, const double* coords,
                                 const rise_oracle_object* objs, int n_objs, const float* rays, int64_t n_rays,
                                 int32_t* prim_id, int32_t* object, double* range, double* alpha, double* beta,
                                 double* normal, double* point, double* shading, double* uv, uint8_t* flags,
                                 int opts, int n_threads)
{
    ojob j;
    memset(&j, 0, sizeof(j));
    tri64* tt = expand_tris_f64(tris, mesh_off[n_meshes]);
    j.tris = tt;
    j.mesh_off = mesh_off;
    j.mesh_flags = mesh_flags;
    j.normals = normals;
    j.coords = coords;
    j.shading = shading;
    j.uv = uv;
    j.objs = objs;
    j.n_objs = n_objs;
    j.rays = rays;
    j.n_rays = n_rays;
    j.prim_id = prim_id;
    j.object = object;
    j.range = range;
    j.alpha = alpha;
    j.beta = beta;
    j.normal = normal;
    j.point = point;
    j.flags = flags;
    j.opts = opts;
    orun(&j, objects_closest_worker, n_threads);
    free(tt);
}

void rise_oracle_objects_any(const double* tris, const int64_t* mesh_off, int n_meshes, const uint8_t* mesh_flags,
                             const rise_oracle_object* objs, int n_objs, const float* rays, const float* tmax,
                             int64_t n_rays, uint8_t* occluded, int opts, int n_threads)
{
    ojob j;
    memset(&j, 0, sizeof(j));
    tri64* tt = expand_tris_f64(tris, mesh_off[n_meshes]);
    j.tris = tt;
    j.mesh_off = mesh_off;
    j.mesh_flags = mesh_flags;
    j.objs = objs;
    j.n_objs = n_objs;
    j.rays = rays;
    j.tmax = tmax;
    j.n_rays = n_rays;
    j.occluded = occluded;
    j.opts = opts;
    orun(&j, objects_any_worker, n_threads);
    free(tt);
}
