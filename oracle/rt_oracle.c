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

/* Vector3Ops::Normalize (Utilities/Math3D/VectorsOps.h:229-238):
 * mag = sqrt(x*x + y*y + z*z); if mag > 0, v * (1/mag). */
void rise_oracle_normalize(const double v[3], double out[3])
{
    const double mag = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
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
int rise_oracle_ray_triangle(const double o[3], const double d[3],
                             const double v0[3], const double eA[3], const double eB[3],
                             double* t, double* alpha, double* beta)
{
    const double p0 = d[1] * eB[2] - d[2] * eB[1];
    const double p1 = d[2] * eB[0] - d[0] * eB[2];
    const double p2 = d[0] * eB[1] - d[1] * eB[0];
    const double det = eA[0] * p0 + eA[1] * p1 + eA[2] * p2;
    if (det > -NZ && det < NZ)
        return 0;
    const double oodet = 1.0 / det;
    const double t0 = o[0] - v0[0];
    const double t1 = o[1] - v0[1];
    const double t2 = o[2] - v0[2];
    const double a = (t0 * p0 + t1 * p1 + t2 * p2) * oodet;
    if ((a < 0.0 - NZ) || (a > 1.0 + NZ))
        return 0;
    const double q0 = t1 * eA[2] - t2 * eA[1];
    const double q1 = t2 * eA[0] - t0 * eA[2];
    const double q2 = t0 * eA[1] - t1 * eA[0];
    const double b = (d[0] * q0 + d[1] * q1 + d[2] * q2) * oodet;
    if ((b < 0.0 - NZ) || ((a + b) > 1.0 + NZ))
        return 0;
    const double r = (eB[0] * q0 + eB[1] * q1 + eB[2] * q2) * oodet;
    if (r >= NZ) {
        *t = r;
        if (alpha) *alpha = a;
        if (beta) *beta = b;
        return 1;
    }
    return 0;
}

typedef struct {
    double v0[3], eA[3], eB[3];
} tri64;

/* Edges in double from the (float-valued) vertices, as
 * TriangleMeshGeometryIndexedSpecializations.h:141-142 computes them. */
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
    uint8_t* occluded;
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
            rise_oracle_normalize(dw, dn);
            double best = RISE_ORACLE_INFINITY;
            int64_t best_id = -1;
            for (int64_t k = 0; k < j->n_tris; ++k) {
                double tt;
                if (rise_oracle_ray_triangle(o, dn, j->tris[k].v0, j->tris[k].eA, j->tris[k].eB,
                                             &tt, NULL, NULL)) {
                    /* Strict-closer commit (...Specializations.h:172); k
                     * ascends, so an exact tie keeps the lowest id. */
                    if (tt < best) {
                        best = tt;
                        best_id = k;
                    }
                }
            }
            if (best_id >= 0) {
                /* Object.cpp:641-643: back off by SURFACE_INTERSEC_ERROR,
                 * then the world range |p - origin| (PointAtLength =
                 * origin + dir*s, Utilities/Ray.h:91-94). */
                const double s2 = best - 1e-12;
                const double px = o[0] + dn[0] * s2;
                const double py = o[1] + dn[1] * s2;
                const double pz = o[2] + dn[2] * s2;
                const double vx = px - o[0], vy = py - o[1], vz = pz - o[2];
                const double range = sqrt(vx * vx + vy * vy + vz * vz);
                j->prim_id[i] = (int32_t)best_id;
                j->t[i] = (float)range;
                if (j->t64) j->t64[i] = range;
            } else {
                j->prim_id[i] = -1;
                j->t[i] = INFINITY;
                if (j->t64) j->t64[i] = INFINITY;
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
            rise_oracle_normalize(dw, dn);
            /* Object.cpp:472-493: factor = |M^-1 d| (identity -> |d|);
             * dHowFar is scaled unless it is the RISE_INFINITY sentinel and
             * factor >= 1. */
            const double mag = sqrt(dw[0] * dw[0] + dw[1] * dw[1] + dw[2] * dw[2]);
            const double factor = (mag > NZ) ? mag : 1.0;
            const double how_far = j->tmax ? (double)j->tmax[i] : RISE_ORACLE_INFINITY;
            double how_far2 = how_far;
            if ((how_far != RISE_ORACLE_INFINITY) || (factor < 1.0))
                how_far2 = factor * how_far;
            uint8_t occ = 0;
            for (int64_t k = 0; k < j->n_tris && !occ; ++k) {
                double tt;
                /* ...Specializations.h:517-519 */
                if (rise_oracle_ray_triangle(o, dn, j->tris[k].v0, j->tris[k].eA, j->tris[k].eB,
                                             &tt, NULL, NULL) &&
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
                               int32_t* prim_id, float* t, double* t64, int n_threads)
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
    run(&j, closest_worker, n_threads);
    free(tt);
}

void rise_oracle_any_brute(const float* tris, int64_t n_tris,
                           const float* rays, const float* tmax, int64_t n_rays,
                           uint8_t* occluded, int n_threads)
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
    run(&j, any_worker, n_threads);
    free(tt);
}
