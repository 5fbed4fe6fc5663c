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

// The top-level walk's state: the node group being tested and the leaves of
// the last node tested.
struct TopWalk {
    unsigned gbase, ghits, gimask;
    unsigned tbase, tmask;
    float tlim;  // box limit along the scaled world direction
};

template <int A>
RISE_HD inline void top_begin(const RayStateT<A>& r, float tlim, TopWalk& w)
{
    w.gbase = 0;
    w.ghits = 1u << (0u ^ r.oinv);  // the root, in slot 0 of a virtual group
    w.gimask = 1u;
    w.tbase = 0;
    w.tmask = 0;
    w.tlim = tlim;
}

// One step of the top-level walk, near to far with the world ray: returns the
// object record index of the next leaf reached, -1 after a node test or a
// pop, -2 once the walk has ended (the stack down to its floor).  A node's
// leaves are handed out before the next node is tested, so a limit lowered
// by one object applies to the nodes after it.
template <int A, class Stack>
RISE_HD inline int top_step(const RayStateT<A>& r, const NodeSlot* __restrict__ nodes,
                            const Tri48* __restrict__ leaves, Stack& st, TopWalk& w)
{
    if (w.tmask) {
        const unsigned j = __builtin_ctz(w.tmask);
        w.tmask &= w.tmask - 1u;
        return (int)float_to_bits(leaves[w.tbase + j].v0[3]);
    }
    if (w.ghits) {
        const unsigned k = __builtin_ctz(w.ghits);
        w.ghits &= w.ghits - 1u;
        const unsigned slot = k ^ r.oinv;
        const unsigned ni = w.gbase + __builtin_popcount(w.gimask & ((1u << slot) - 1u));
        if (w.ghits) st.push(make_uint2(w.gbase, w.ghits | (w.gimask << 8)));
        unsigned cb, im;
        w.ghits = test_node(r, nodes, ni, w.tlim, cb, im, w.tbase, w.tmask);
        w.gbase = cb;
        w.gimask = im;
        return -1;
    }
    if (st.empty()) return -2;
    const uint2 g = st.pop();
    w.gbase = g.x;
    w.ghits = g.y & 0xffu;
    w.gimask = (g.y >> 8) & 0xffu;
    return -1;
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

// Phases of an object query.
enum : int { kObjPhaseWorld = 0, kObjPhaseTop = 1, kObjPhaseObject = 2, kObjPhaseDone = 3 };

// ObjectManager::IntersectRay over the scene (each object's own closest hit,
// Object::IntersectRay; the strictly closer world range wins, ties to the
// lower object index, the caller's order) as a state machine advanced one
// traversal step at a time, so that a persistent kernel can give a lane
// whose query has ended a new ray while the other lanes of its wave go on.
// Phases: the world tree, then the top-level walk, which enters the tree of
// each object it reaches and resumes when that tree is done.
template <int A>
struct ObjClosestQuery {
    RayStateT<A> r;  // the world ray (world tree and top-level tree)
    ObjRayT<A> q;    // the ray of the object being traversed
    ClosestState s;  // the walk of the tree being traversed
    TopWalk w;
    ObjectHit h;     // the best so far
    const NodeSlot* nodes;  // the tree of the object being traversed
    const Tri48* tris;
    const double* tri64;
    unsigned faces;
    int k;  // its object record
    int phase;
};

// The top-level box limit: a hit closer than the best lies in its object's
// padded box within the best (plus rounding) along the world ray.
template <int A>
RISE_HD inline float obj_closest_limit(const ObjClosestQuery<A>& Q)
{
    return box_limit(Q.r, Q.h.range < kRiseInfinity ? Q.h.range * (1.0 + 1e-7) : kRiseInfinity);
}

template <int A>
RISE_HD inline void obj_closest_to_top(const ObjectSceneView& sc, ObjClosestQuery<A>& Q)
{
    if (sc.n_objects == 0) {
        Q.phase = kObjPhaseDone;
        return;
    }
    top_begin(Q.r, obj_closest_limit(Q), Q.w);
    Q.phase = kObjPhaseTop;
}

template <int A, class Stack>
RISE_HD inline void objects_closest_begin(const ObjectSceneView& sc, const float ray[6], unsigned faces, Stack& st,
                                          ObjClosestQuery<A>& Q)
{
    Q.h.prim_id = -1;
    Q.h.object = -1;
    Q.h.src = -1;
    Q.h.leaf = 0;
    Q.h.t = kRiseInfinity;
    Q.h.range = kRiseInfinity;
    st.sp = 0;
    st.base = 0;
    st.overflow = false;
    init_ray(Q.r, ray[0], ray[1], ray[2], ray[3], ray[4], ray[5]);
    Q.faces = faces;
    if (sc.world_nodes) {
        closest_begin(Q.r, Q.s);
        Q.s.faces = faces;
        Q.phase = kObjPhaseWorld;
    } else {
        obj_closest_to_top(sc, Q);
    }
}

// One step; false once the query has ended (the hit is in Q.h).
template <int A, class Stack>
RISE_HD inline bool objects_closest_step(const ObjectSceneView& sc, Stack& st, ObjClosestQuery<A>& Q)
{
    unsigned nn = 0, nt = 0;
    const double ow[3] = {Q.r.ox, Q.r.oy, Q.r.oz};
    if (Q.phase == kObjPhaseWorld) {
        if (closest_step<false, true, true>(Q.r, sc.world_nodes, sc.world_tris, st, Q.s, nn, nt, sc.world_tri64))
            return true;
        if (Q.s.best_id >= 0) {
            const float4* tp = reinterpret_cast<const float4*>(sc.world_tris + Q.s.best_leaf);
            Q.h.prim_id = Q.s.best_id;
            Q.h.object = float_to_bits(tp[1].w);
            Q.h.leaf = Q.s.best_leaf;
            Q.h.t = Q.s.best_t;
            Q.h.range = world_range(Q.r, Q.s.best_t);
        }
        obj_closest_to_top(sc, Q);
        return Q.phase != kObjPhaseDone;
    }
    if (Q.phase == kObjPhaseObject) {
        if (closest_step<false, true, true>(Q.q, Q.nodes, Q.tris, st, Q.s, nn, nt, Q.tri64)) return true;
        st.base = 0;  // back to the top-level walk's entries
        Q.phase = kObjPhaseTop;
        if (Q.s.best_id >= 0) {
            const DevObject& ob = sc.objects[Q.k];
            double pw[3];
            const double w = obj_world_range(Q.q, ob.to_world, ow, Q.s.best_t, pw);
            if (w < Q.h.range || (w == Q.h.range && ob.object < Q.h.object)) {
                Q.h.prim_id = Q.s.best_id;
                Q.h.object = ob.object;
                Q.h.src = Q.k;
                Q.h.leaf = Q.s.best_leaf;
                Q.h.t = Q.s.best_t;
                Q.h.range = w;
                for (int c = 0; c < 3; ++c) Q.h.p_world[c] = pw[c];
                Q.w.tlim = obj_closest_limit(Q);
            }
        }
        return true;
    }
    if (Q.phase != kObjPhaseTop) return false;
    const int k = top_step(Q.r, sc.top_nodes, sc.top_leaves, st, Q.w);
    if (k == -2) {
        Q.phase = kObjPhaseDone;
        return false;
    }
    if (k < 0) return true;
    const DevObject& ob = sc.objects[k];
    const double dw[3] = {Q.r.dx, Q.r.dy, Q.r.dz};
    if (!world_box_reach(ob.lo, ob.hi, ow, dw, Q.r.len_d, Q.h.range)) return true;
    init_obj_ray(Q.q, ow, dw, ob.to_object);
    const BlasView bl = sc.blas[ob.blas];
    Q.nodes = bl.nodes;
    Q.tris = bl.tris;
    Q.tri64 = bl.tri64;
    closest_begin(Q.q, Q.s);
    Q.s.faces = Q.faces;
    if (Q.h.range < kRiseInfinity && Q.r.len_d > 0.0) {
        // A world range w is an object-space distance of about w |dl| / |d|;
        // the rounding of o' and of the point taken back to world space is
        // below slack (|o| + |translation| + w).
        const double tw[3] = {ob.to_world[12], ob.to_world[13], ob.to_world[14]};
        const double lim_w = Q.h.range * (1.0 + ob.slack + 1e-9) + ob.slack * (max_abs3(ow) + max_abs3(tw));
        Q.s.tlim = box_limit(Q.q, lim_w * (Q.q.mag / Q.r.len_d) * (1.0 + 1e-9));
    }
    st.base = st.sp;  // the object's entries go above the top-level walk's
    Q.k = k;
    Q.phase = kObjPhaseObject;
    return true;
}

// The whole closest-hit query for one ray.
template <int A, class Stack>
RISE_HD inline void objects_closest(const ObjectSceneView& sc, const float ray[6], unsigned faces, Stack& st,
                                    ObjectHit& h)
{
    ObjClosestQuery<A> Q;
    objects_closest_begin(sc, ray, faces, st, Q);
    while (objects_closest_step(sc, st, Q)) {
    }
    h = Q.h;
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
// with Object::IntersectRay_IntersectionOnly per object) as a state machine
// like ObjClosestQuery.  how_far: the world limit dHowFar.  The answer does
// not depend on the order objects are tried in.
template <int A>
struct ObjAnyQuery {
    RayStateT<A> r;
    ObjRayT<A> q;
    AnyState s;
    TopWalk w;
    const NodeSlot* nodes;
    const Tri48* tris;
    const double* tri64;
    double how_far;
    unsigned faces;
    int phase;
    bool hit;
};

template <int A>
RISE_HD inline void obj_any_to_top(const ObjectSceneView& sc, ObjAnyQuery<A>& Q)
{
    if (sc.n_objects == 0) {
        Q.phase = kObjPhaseDone;
        return;
    }
    // No distance limit on the top-level walk: an object whose transform
    // nearly collapses the direction has no bounded world reach.
    top_begin(Q.r, INFINITY, Q.w);
    Q.phase = kObjPhaseTop;
}

template <int A, class Stack>
RISE_HD inline void objects_any_begin(const ObjectSceneView& sc, const float ray[6], double how_far, unsigned faces,
                                      Stack& st, ObjAnyQuery<A>& Q)
{
    st.sp = 0;
    st.base = 0;
    st.overflow = false;
    init_ray(Q.r, ray[0], ray[1], ray[2], ray[3], ray[4], ray[5]);
    Q.how_far = how_far;
    Q.faces = faces;
    Q.hit = false;
    if (sc.world_nodes) {
        any_begin(Q.r, scale_how_far(Q.r, how_far), Q.s);
        Q.s.faces = faces;
        Q.phase = kObjPhaseWorld;
    } else {
        obj_any_to_top(sc, Q);
    }
}

// One step; false once the query has ended (the answer is in Q.hit).
template <int A, class Stack>
RISE_HD inline bool objects_any_step(const ObjectSceneView& sc, Stack& st, ObjAnyQuery<A>& Q)
{
    unsigned nn = 0, nt = 0;
    if (Q.phase == kObjPhaseWorld) {
        if (any_step<false, true, true>(Q.r, sc.world_nodes, sc.world_tris, st, Q.s, nn, nt, sc.world_tri64))
            return true;
        if (Q.s.hit) {
            Q.hit = true;
            Q.phase = kObjPhaseDone;
            return false;
        }
        obj_any_to_top(sc, Q);
        return Q.phase != kObjPhaseDone;
    }
    if (Q.phase == kObjPhaseObject) {
        if (any_step<false, true, true>(Q.q, Q.nodes, Q.tris, st, Q.s, nn, nt, Q.tri64)) return true;
        if (Q.s.hit) {
            Q.hit = true;
            Q.phase = kObjPhaseDone;
            return false;
        }
        st.sp = st.base;  // back to the top-level walk's entries
        st.base = 0;
        Q.phase = kObjPhaseTop;
        return true;
    }
    if (Q.phase != kObjPhaseTop) return false;
    const int k = top_step(Q.r, sc.top_nodes, sc.top_leaves, st, Q.w);
    if (k == -2) {
        Q.phase = kObjPhaseDone;
        return false;
    }
    if (k < 0) return true;
    const DevObject& ob = sc.objects[k];
    const double ow[3] = {Q.r.ox, Q.r.oy, Q.r.oz};
    const double dw[3] = {Q.r.dx, Q.r.dy, Q.r.dz};
    init_obj_ray(Q.q, ow, dw, ob.to_object);
    const double how_far2 = obj_how_far(Q.q.factor, Q.how_far);
    // An object-space distance t along dn' is t |d| / |dl| along the world
    // ray, so the limit reaches |d| how_far2 / |dl| in world space (widened
    // by 1e-6 for the rounding of the two transforms).
    const double lim = (how_far2 < kRiseInfinity && Q.q.mag > 0.0) ? how_far2 * Q.r.len_d / Q.q.mag * (1.0 + 1e-6)
                                                                   : kRiseInfinity;
    if (!world_box_reach(ob.lo, ob.hi, ow, dw, Q.r.len_d, lim)) return true;
    const BlasView bl = sc.blas[ob.blas];
    Q.nodes = bl.nodes;
    Q.tris = bl.tris;
    Q.tri64 = bl.tri64;
    any_begin(Q.q, how_far2, Q.s);
    Q.s.faces = Q.faces;
    st.base = st.sp;
    Q.phase = kObjPhaseObject;
    return true;
}

// The whole any-hit query for one ray.
template <int A, class Stack>
RISE_HD inline bool objects_any(const ObjectSceneView& sc, const float ray[6], double how_far, unsigned faces,
                                Stack& st)
{
    ObjAnyQuery<A> Q;
    objects_any_begin(sc, ray, how_far, faces, st, Q);
    while (objects_any_step(sc, st, Q)) {
    }
    return Q.hit;
}

}  // namespace rise_hip
