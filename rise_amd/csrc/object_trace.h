// object_trace.h -- ray queries over a scene of objects with transforms.
//
// The reference intersects a ray with each object of its ObjectManager in
// that object's own frame (Objects/Object.cpp:392-649, :652-712): the world
// ray is taken into object space in double, the object's mesh is traversed
// with the object-space ray, and the hit comes back as a world point whose
// distance from the world origin decides, strictly smaller winning, which
// object's hit is kept (Managers/ObjectManager.cpp:68-99, :290-310).
//
// Every triangle of an object scene keeps its double vertices beside its leaf
// record (the certifier and the face test read them; the float filter and the
// boxes use their float rounding, as the reference's BVH filter does).
// Objects whose two matrices are both the identity are merged into one
// world-space tree (their object space is world space, and for them the
// reference's transforms are exact), and every other object keeps an
// object-space tree of its mesh, shared by all objects on that mesh.  A query
// runs the world tree with the float-origin ray state (RayStateT), then walks
// a top-level tree over the transformed objects' padded world boxes near to
// far, skipping boxes entered beyond the current best; each object reached
// traverses its own tree with the double object-space ray state (ObjRayT),
// its boxes pruned at the best's object-space image (widened for rounding).
// Objects are visited out of the caller's order, so ties are settled as the
// caller's order would: an equal world range goes to the lower object index.
// The per-ray functions here are compiled for host and device.
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

// Node-group stack of an object query with a movable floor: an object's tree
// is walked on top of the top-level tree's entries.  The upload checks that
// the depths fit (ObjectSceneHost::stack_need); a push past the end is
// dropped and flagged.  In a private array of kObjStack entries, or strided
// in LDS (CAP entries per lane, `stride` lanes apart).
struct ObjStack {
    uint2 e[kObjStack];
    int sp;
    int base;
    bool overflow;
    RISE_HD inline void push(uint2 v)
    {
        if (sp < kObjStack) e[sp++] = v;
        else overflow = true;
    }
    RISE_HD inline uint2 pop() { return e[--sp]; }
    RISE_HD inline bool empty() const { return sp == base; }
};

template <int CAP>
struct ObjLdsStack {
    uint2* e;
    int stride;
    int sp;
    int base;
    bool overflow;
    RISE_HD inline void push(uint2 v)
    {
        if (sp < CAP) e[sp++ * stride] = v;
        else overflow = true;
    }
    RISE_HD inline uint2 pop() { return e[--sp * stride]; }
    RISE_HD inline bool empty() const { return sp == base; }
};

// Walks the top-level tree near to far with the world ray; visit(k) is called
// for each object record k reached and returns the new box limit, or a
// negative value to end the walk.
template <int A, class Stack, class Visit>
RISE_HD inline void top_walk(const RayStateT<A>& r, const NodeSlot* __restrict__ nodes,
                             const Tri48* __restrict__ leaves, Stack& st, float tlim, Visit&& visit)
{
    unsigned gbase = 0, ghits = 1u << (0u ^ r.oinv), gimask = 1u;
    const int floor = st.sp;
    for (;;) {
        unsigned tmask = 0, tbase = 0;
        if (ghits) {
            const unsigned k = __builtin_ctz(ghits);
            ghits &= ghits - 1u;
            const unsigned slot = k ^ r.oinv;
            const unsigned ni = gbase + __builtin_popcount(gimask & ((1u << slot) - 1u));
            if (ghits) st.push(make_uint2(gbase, ghits | (gimask << 8)));
            unsigned cb, im;
            ghits = test_node(r, nodes, ni, tlim, cb, im, tbase, tmask);
            gbase = cb;
            gimask = im;
        }
        while (tmask) {
            const unsigned j = __builtin_ctz(tmask);
            tmask &= tmask - 1u;
            tlim = visit(float_to_bits(leaves[tbase + j].v0[3]));
            if (tlim < 0.0f) {
                st.sp = floor;
                return;
            }
        }
        if (!ghits) {
            if (st.sp == floor) break;
            const uint2 g = st.pop();
            gbase = g.x;
            ghits = g.y & 0xffu;
            gimask = (g.y >> 8) & 0xffu;
        }
    }
}

// Largest absolute coordinate.
RISE_HD inline double max_abs3(const double v[3]) { return fmax(fabs(v[0]), fmax(fabs(v[1]), fabs(v[2]))); }

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

// ObjectManager::IntersectRay over the scene: each object's own closest hit
// (Object::IntersectRay), the strictly closer world range winning, ties to the
// lower object index (the caller's order).
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
    st.sp = 0;
    st.base = 0;
    st.overflow = false;
    RayStateT<A> r;
    init_ray(r, ray[0], ray[1], ray[2], ray[3], ray[4], ray[5]);
    unsigned nn = 0, nt = 0;
    if (sc.world_nodes) {
        ClosestState s;
        closest_begin(r, s);
        s.faces = faces;
        while (closest_step<false, true, true>(r, sc.world_nodes, sc.world_tris, st, s, nn, nt, sc.world_tri64)) {
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
    if (sc.n_objects == 0) return;
    const double ow[3] = {ray[0], ray[1], ray[2]};
    const double dw[3] = {ray[3], ray[4], ray[5]};
    const double omax = max_abs3(ow);
    // A hit closer than the best lies in its object's padded box within the
    // best (plus rounding) along the world ray.
    auto limit = [&]() { return box_limit(r, h.range < kRiseInfinity ? h.range * (1.0 + 1e-7) : kRiseInfinity); };
    top_walk(r, sc.top_nodes, sc.top_leaves, st, limit(), [&](int k) -> float {
        const DevObject& ob = sc.objects[k];
        if (!world_box_reach(ob.lo, ob.hi, ow, dw, r.len_d, h.range)) return limit();
        ObjRayT<A> q;
        init_obj_ray(q, ow, dw, ob.to_object);
        const BlasView bl = sc.blas[ob.blas];
        ClosestState s;
        closest_begin(q, s);
        s.faces = faces;
        if (h.range < kRiseInfinity && r.len_d > 0.0) {
            // A world range w is an object-space distance of about w |dl| / |d|;
            // the rounding of o' and of the point taken back to world space is
            // below slack (|o| + |translation| + w).
            const double tw[3] = {ob.to_world[12], ob.to_world[13], ob.to_world[14]};
            const double lim_w = h.range * (1.0 + ob.slack + 1e-9) + ob.slack * (omax + max_abs3(tw));
            s.tlim = box_limit(q, lim_w * (q.mag / r.len_d) * (1.0 + 1e-9));
        }
        const int floor = st.base;
        st.base = st.sp;
        while (closest_step<false, true, true>(q, bl.nodes, bl.tris, st, s, nn, nt, bl.tri64)) {
        }
        st.base = floor;
        if (s.best_id < 0) return limit();
        double pw[3];
        const double w = obj_world_range(q, ob.to_world, ow, s.best_t, pw);
        if (w < h.range || (w == h.range && ob.object < h.object)) {
            h.prim_id = s.best_id;
            h.object = ob.object;
            h.src = k;
            h.leaf = s.best_leaf;
            h.t = s.best_t;
            h.range = w;
            for (int c = 0; c < 3; ++c) h.p_world[c] = pw[c];
        }
        return limit();
    });
}

// Fills the record's attributes from a hit in the mesh's frame: normals taken
// by the object's inverse transpose (minv; NULL: identity) and normalized
// (Object.cpp:531-533).
template <int A>
RISE_HD inline void record_normals(const HitLocal& l, const double* minv, rise_hip_hit& out)
{
    double g[3], sh[3], gw[3], sw[3];
    if (minv) {
        xform_normal<A>(minv, l.gn, g);
        xform_normal<A>(minv, l.sn, sh);
    } else {
        for (int k = 0; k < 3; ++k) {
            g[k] = l.gn[k];
            sh[k] = l.sn[k];
        }
    }
    normalize_model<A>(g, gw);
    normalize_model<A>(sh, sw);
    out.alpha = l.ab[0];
    out.beta = l.ab[1];
    for (int k = 0; k < 3; ++k) {
        out.geom_normal[k] = (float)gw[k];
        out.shading_normal[k] = (float)sw[k];
    }
    out.uv[0] = (float)l.uv[0];
    out.uv[1] = (float)l.uv[1];
}

RISE_HD inline void record_miss(rise_hip_hit& out)
{
    out.prim_id = -1;
    out.mesh = -1;
    out.range = INFINITY;
    out.alpha = out.beta = 0.0;
    for (int k = 0; k < 3; ++k) out.point[k] = out.geom_normal[k] = out.shading_normal[k] = 0.0f;
    out.uv[0] = out.uv[1] = 0.0f;
}

// The intersection record of a committed hit (rise_hip_hit): barycentrics of
// the certifier, the geometric and shading normals and the texture
// coordinates as the mesh leaves them (hit_local), the normals taken to world
// space by the object's inverse transpose and Normalize (Object.cpp:531-533),
// the world point.
template <int A>
RISE_HD inline void objects_record(const ObjectSceneView& sc, const float ray[6], const ObjectHit& h,
                                   rise_hip_hit& out)
{
    if (h.prim_id < 0) {
        record_miss(out);
        return;
    }
    out.prim_id = h.prim_id;
    out.mesh = h.object;
    out.range = h.range;
    double v[9];
    HitLocal l;
    if (h.src < 0) {
        RayStateT<A> r;
        init_ray(r, ray[0], ray[1], ray[2], ray[3], ray[4], ray[5]);
        const Tri48* tri = sc.world_tris + h.leaf;
        triangle_vertices<RayStateT<A>>(tri, sc.world_tri64 + 9 * (size_t)h.leaf, v);
        hit_local(r, v, tri_flags(tri), sc.world_attr ? sc.world_attr + kAttrDoubles * (size_t)h.leaf : nullptr, l);
        record_normals<A>(l, nullptr, out);
        const double s = h.t - kSurfaceBackoff;
        const double d[3] = {dn0(r), dn1(r), dn2(r)};
        const double o[3] = {r.ox, r.oy, r.oz};
        for (int k = 0; k < 3; ++k) out.point[k] = (float)(A == kArithX86Fma ? fma(d[k], s, o[k]) : o[k] + d[k] * s);
        return;
    }
    const DevObject& ob = sc.objects[h.src];
    const double ow[3] = {ray[0], ray[1], ray[2]};
    const double dw[3] = {ray[3], ray[4], ray[5]};
    ObjRayT<A> q;
    init_obj_ray(q, ow, dw, ob.to_object);
    const BlasView bl = sc.blas[ob.blas];
    const Tri48* tri = bl.tris + h.leaf;
    triangle_vertices<ObjRayT<A>>(tri, bl.tri64 + 9 * (size_t)h.leaf, v);
    hit_local(q, v, tri_flags(tri), bl.attr ? bl.attr + kAttrDoubles * (size_t)h.leaf : nullptr, l);
    record_normals<A>(l, ob.to_object, out);
    for (int k = 0; k < 3; ++k) out.point[k] = (float)h.p_world[k];
}

// ObjectManager::IntersectShadowRay over the scene (ObjectManager.cpp:320-373
// with Object::IntersectRay_IntersectionOnly per object).  how_far: the
// world limit dHowFar.  The answer does not depend on the order objects are
// tried in.
template <int A, class Stack>
RISE_HD inline bool objects_any(const ObjectSceneView& sc, const float ray[6], double how_far, unsigned faces,
                                Stack& st)
{
    st.sp = 0;
    st.base = 0;
    st.overflow = false;
    RayStateT<A> r;
    init_ray(r, ray[0], ray[1], ray[2], ray[3], ray[4], ray[5]);
    unsigned nn = 0, nt = 0;
    if (sc.world_nodes) {
        AnyState s;
        any_begin(r, scale_how_far(r, how_far), s);
        s.faces = faces;
        while (any_step<false, true, true>(r, sc.world_nodes, sc.world_tris, st, s, nn, nt, sc.world_tri64)) {
        }
        if (s.hit) return true;
    }
    if (sc.n_objects == 0) return false;
    const double ow[3] = {ray[0], ray[1], ray[2]};
    const double dw[3] = {ray[3], ray[4], ray[5]};
    bool hit = false;
    // The top-level walk takes no distance limit: an object whose transform
    // nearly collapses the direction has no bounded world reach.
    top_walk(r, sc.top_nodes, sc.top_leaves, st, INFINITY, [&](int k) -> float {
        const DevObject& ob = sc.objects[k];
        ObjRayT<A> q;
        init_obj_ray(q, ow, dw, ob.to_object);
        const double how_far2 = obj_how_far(q.factor, how_far);
        // An object-space distance t along dn' is t |d| / |dl| along the
        // world ray, so the limit reaches |d| how_far2 / |dl| in world
        // space (widened by 1e-6 for the rounding of the two transforms).
        const double lim = (how_far2 < kRiseInfinity && q.mag > 0.0) ? how_far2 * r.len_d / q.mag * (1.0 + 1e-6)
                                                                     : kRiseInfinity;
        if (!world_box_reach(ob.lo, ob.hi, ow, dw, r.len_d, lim)) return INFINITY;
        const BlasView bl = sc.blas[ob.blas];
        AnyState s;
        any_begin(q, how_far2, s);
        s.faces = faces;
        const int floor = st.base;
        st.base = st.sp;
        while (any_step<false, true, true>(q, bl.nodes, bl.tris, st, s, nn, nt, bl.tri64)) {
        }
        st.sp = st.base;  // an early exit leaves the object's entries
        st.base = floor;
        hit = s.hit;
        return hit ? -1.0f : INFINITY;
    });
    return hit;
}

}  // namespace rise_hip
