// object_trace.h -- ray queries over a scene of objects with transforms.
//
// The reference intersects a ray with each object of its ObjectManager in
// that object's own frame (Objects/Object.cpp:392-649, :652-712): the world
// ray is taken into object space in double, the object's mesh is traversed
// with the object-space ray, and the hit comes back as a world point whose
// distance from the world origin decides, strictly smaller winning, which
// object's hit is kept (Managers/ObjectManager.cpp:68-99, :290-310).
//
// Here objects whose two matrices are both the identity are merged into one
// world-space tree (their object space is world space, and for them the
// reference's transforms are exact), and every other object keeps an
// object-space tree of its mesh, shared by all objects on that mesh.  A query
// runs the world tree with the float-origin ray state (RayStateT), then each
// transformed object whose conservative world box the ray reaches within the
// current best, with the double object-space ray state (ObjRayT).  The per-ray
// functions here are compiled for host and device.
#pragma once

#include "object_types.h"
#include "rise_hip.h"
#include "trace_core.h"

namespace rise_hip {

// Whether the world box [lo, hi] can hold a point of the ray o + s d (s >= 0)
// at a world distance below lim (kRiseInfinity: no limit).  Conservative:
// the entry distance is compared with lim widened by 1e-9, on top of the
// box's own padding.
RISE_HD inline bool world_box_reach(const double lo[3], const double hi[3], const double o[3], const double d[3],
                                    double len_d, double lim)
{
    double tn = 0.0, tf = kRiseInfinity;
    for (int a = 0; a < 3; ++a) {
        if (d[a] == 0.0) {
            if (o[a] < lo[a] || o[a] > hi[a]) return false;
            continue;
        }
        const double inv = 1.0 / d[a];
        double t1 = (lo[a] - o[a]) * inv, t2 = (hi[a] - o[a]) * inv;
        if (t1 > t2) {
            const double x = t1;
            t1 = t2;
            t2 = x;
        }
        tn = fmax(tn, t1);
        tf = fmin(tf, t2);
    }
    if (tn > tf * (1.0 + 0x1p-40)) return false;
    if (lim < kRiseInfinity && tn * len_d > lim * (1.0 + 1e-9)) return false;
    return true;
}

// Object.cpp:637-643 for a transformed object: back off 1e-12 along dn',
// take the point to world space and measure its distance from the world
// origin.  p_world receives the world point.
template <int A>
RISE_HD inline double obj_world_range(const ObjRayT<A>& q, const double* to_world, const double ow[3], double t,
                                      double p_world[3])
{
    const double s = t - kSurfaceBackoff;
    double po[3];
    for (int k = 0; k < 3; ++k) po[k] = A == kArithX86Fma ? fma(q.dn[k], s, q.o[k]) : q.o[k] + q.dn[k] * s;
    xform_point<A>(to_world, po, p_world);
    const double vx = p_world[0] - ow[0], vy = p_world[1] - ow[1], vz = p_world[2] - ow[2];
    return sqrt(dot3<A>(vx, vx, vy, vy, vz, vz));
}

// Object.cpp:480-493: the world limit dHowFar becomes factor * dHowFar in
// the object's frame, unless it is the RISE_INFINITY sentinel and factor >= 1.
RISE_HD inline double obj_how_far(double factor, double how_far)
{
    if ((how_far != kRiseInfinity) || (factor < 1.0)) return factor * how_far;
    return how_far;
}

// Result of a closest-hit query over objects.
struct ObjectHit {
    int32_t prim_id;  // caller id of the triangle, -1 on a miss
    int32_t object;   // caller index of the object
    int32_t src;      // -1: the world tree; else the index into ObjectSceneView::objects
    unsigned leaf;    // leaf-order index of the triangle in its tree
    double t;         // distance along the (object-space) normalized direction
    double range;     // world distance (Object.cpp:643)
    double p_world[3];
};

// ObjectManager::IntersectRay over the scene (linear order of the objects;
// strictly closer world range wins).  Stack: a node-group stack with
// push/pop/empty and an `sp` member.
template <int A, class Stack>
RISE_HD inline void objects_closest(const ObjectSceneView& sc, const float ray[6], unsigned faces, Stack& st,
                                    ObjectHit& h)
{
    h.prim_id = -1;
    h.object = -1;
    h.src = -1;
    h.leaf = 0;
    h.t = kRiseInfinity;
    h.range = kRiseInfinity;
    RayStateT<A> r;
    init_ray(r, ray[0], ray[1], ray[2], ray[3], ray[4], ray[5]);
    unsigned nn = 0, nt = 0;
    if (sc.world_nodes) {
        ClosestState s;
        closest_begin(r, s);
        s.faces = faces;
        st.sp = 0;
        while (closest_step<false, true>(r, sc.world_nodes, sc.world_tris, st, s, nn, nt)) {
        }
        if (s.best_id >= 0) {
            const float4* tp = reinterpret_cast<const float4*>(sc.world_tris + s.best_leaf);
            h.prim_id = s.best_id;
            h.object = float_to_bits(tp[1].w);
            h.leaf = s.best_leaf;
            h.t = s.best_t;
            h.range = world_range(r, s.best_t);
        }
    }
    const double ow[3] = {ray[0], ray[1], ray[2]};
    const double dw[3] = {ray[3], ray[4], ray[5]};
    for (int32_t k = 0; k < sc.n_objects; ++k) {
        const DevObject& ob = sc.objects[k];
        if (!world_box_reach(ob.lo, ob.hi, ow, dw, r.len_d, h.range)) continue;
        ObjRayT<A> q;
        init_obj_ray(q, ow, dw, ob.to_object);
        const BlasView bl = sc.blas[ob.blas];
        ClosestState s;
        closest_begin(q, s);
        s.faces = faces;
        st.sp = 0;
        while (closest_step<false, true>(q, bl.nodes, bl.tris, st, s, nn, nt)) {
        }
        if (s.best_id < 0) continue;
        double pw[3];
        const double w = obj_world_range(q, ob.to_world, ow, s.best_t, pw);
        if (w < h.range) {
            h.prim_id = s.best_id;
            h.object = ob.object;
            h.src = k;
            h.leaf = s.best_leaf;
            h.t = s.best_t;
            h.range = w;
            for (int c = 0; c < 3; ++c) h.p_world[c] = pw[c];
        }
    }
}

// The intersection record of a committed hit (rise_hip_hit): barycentrics of
// the certifier, the face normal normalized by the mesh
// (...Specializations.h:196-205) and, in world space, by the object's
// inverse transpose and Normalize (Object.cpp:531-533), the world point.
template <int A>
RISE_HD inline void objects_record(const ObjectSceneView& sc, const float ray[6], const ObjectHit& h,
                                   rise_hip_hit& out)
{
    out.prim_id = h.prim_id;
    out.mesh = h.object;
    if (h.prim_id < 0) {
        out.range = INFINITY;
        out.alpha = out.beta = 0.0;
        for (int k = 0; k < 3; ++k) out.point[k] = out.geom_normal[k] = 0.0f;
        return;
    }
    out.range = h.range;
    if (h.src < 0) {
        RayStateT<A> r;
        init_ray(r, ray[0], ray[1], ray[2], ray[3], ray[4], ray[5]);
        double ab[2], n[3], p[3];
        hit_attributes(r, sc.world_tris + h.leaf, h.t, ab, n, p);
        out.alpha = ab[0];
        out.beta = ab[1];
        for (int k = 0; k < 3; ++k) {
            out.point[k] = (float)p[k];
            out.geom_normal[k] = (float)n[k];
        }
        return;
    }
    const DevObject& ob = sc.objects[h.src];
    const double ow[3] = {ray[0], ray[1], ray[2]};
    const double dw[3] = {ray[3], ray[4], ray[5]};
    ObjRayT<A> q;
    init_obj_ray(q, ow, dw, ob.to_object);
    const float4* tp = reinterpret_cast<const float4*>(sc.blas[ob.blas].tris + h.leaf);
    const float4 a = tp[0], b = tp[1], c = tp[2];
    double tt = 0.0, ab[2] = {0.0, 0.0};
    (void)certify(q, a, b, c, tt, ab);
    const double eAx = (double)b.x - a.x, eAy = (double)b.y - a.y, eAz = (double)b.z - a.z;
    const double eBx = (double)c.x - a.x, eBy = (double)c.y - a.y, eBz = (double)c.z - a.z;
    const double fn[3] = {sub_prod<A>(eAy, eBz, eAz, eBy), sub_prod<A>(eAz, eBx, eAx, eBz),
                          sub_prod<A>(eAx, eBy, eAy, eBx)};
    double n1[3], nw[3], n2[3];
    normalize_model<A>(fn, n1);
    xform_normal<A>(ob.to_object, n1, nw);
    normalize_model<A>(nw, n2);
    out.alpha = ab[0];
    out.beta = ab[1];
    for (int k = 0; k < 3; ++k) {
        out.point[k] = (float)h.p_world[k];
        out.geom_normal[k] = (float)n2[k];
    }
}

// ObjectManager::IntersectShadowRay over the scene (ObjectManager.cpp:320-373
// with Object::IntersectRay_IntersectionOnly per object).  how_far: the
// world limit dHowFar.
template <int A, class Stack>
RISE_HD inline bool objects_any(const ObjectSceneView& sc, const float ray[6], double how_far, unsigned faces,
                                Stack& st)
{
    RayStateT<A> r;
    init_ray(r, ray[0], ray[1], ray[2], ray[3], ray[4], ray[5]);
    unsigned nn = 0, nt = 0;
    if (sc.world_nodes) {
        AnyState s;
        any_begin(r, scale_how_far(r, how_far), s);
        s.faces = faces;
        st.sp = 0;
        while (any_step<false, true>(r, sc.world_nodes, sc.world_tris, st, s, nn, nt)) {
        }
        if (s.hit) return true;
    }
    const double ow[3] = {ray[0], ray[1], ray[2]};
    const double dw[3] = {ray[3], ray[4], ray[5]};
    for (int32_t k = 0; k < sc.n_objects; ++k) {
        const DevObject& ob = sc.objects[k];
        ObjRayT<A> q;
        init_obj_ray(q, ow, dw, ob.to_object);
        const double how_far2 = obj_how_far(q.factor, how_far);
        // An object-space distance t along dn' is t |d| / |dl| along the
        // world ray, so the limit reaches |d| how_far2 / |dl| in world
        // space (widened by 1e-6 for the rounding of the two transforms).
        const double lim = (how_far2 < kRiseInfinity && q.mag > 0.0) ? how_far2 * r.len_d / q.mag * (1.0 + 1e-6)
                                                                     : kRiseInfinity;
        if (!world_box_reach(ob.lo, ob.hi, ow, dw, r.len_d, lim)) continue;
        const BlasView bl = sc.blas[ob.blas];
        AnyState s;
        any_begin(q, how_far2, s);
        s.faces = faces;
        st.sp = 0;
        while (any_step<false, true>(q, bl.nodes, bl.tris, st, s, nn, nt)) {
        }
        if (s.hit) return true;
    }
    return false;
}

}  // namespace rise_hip
