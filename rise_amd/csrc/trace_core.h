// trace_core.h -- device-side traversal of the 8-wide quantized BVH.
//
// One lane traces one ray.  Boxes are tested in FP32 with a conservative
// slab test; every triangle whose leaf box is reached is certified in FP64
// with the reference's Moller-Trumbore (Intersection/RayTriangleIntersection.cpp:21-68),
// so the committed hit is the one the reference's double-precision leaf test
// would commit.  Closest-hit keeps the smallest t, exact ties going to the
// lowest triangle id; the reference's strict-closer guard is
// Geometry/TriangleMeshGeometryIndexedSpecializations.h:172.
//
// Stack: node groups (base index, 8-bit hit mask, 8-bit internal mask), at
// most one per tree level, held in LDS (short stack, STACK entries per lane)
// or, for deep trees, in a private array.
//
// The whole file is compiled with -ffp-contract=off: double operations are
// rounded one by one in the order the reference writes them.
#pragma once

#include "bvh8.h"

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

#define RISE_HD __host__ __device__

namespace rise_hip {

RISE_HD inline float bits_to_float(unsigned u) { return __builtin_bit_cast(float, u); }
RISE_HD inline int float_to_bits(float f) { return __builtin_bit_cast(int, f); }

constexpr double kNearZero = 1e-12;                      // Utilities/Math3D/Constants.h:34
constexpr double kRiseInfinity = 1.7976931348623158e+308;  // Constants.h:33
constexpr double kSurfaceBackoff = 1e-12;                // Objects/Object.cpp:36

// 1 + 2^-19: relative widening of the FP32 slab interval.  The computed entry
// and exit distances (b - o) * (1/d) carry at most ~3 float roundings
// (relative 2^-22 each), so scaling the exit distance by this factor (and the
// running best by it again) never rejects a box the exact ray enters.
#define RISE_SLAB_PAD 1.0000019073486328f

// The certifier's direction is d normalized in double exactly as
// Vector3Ops::Normalize computes it (Object.cpp:414-425): dn_k = d_k * ov with
// ov = 1/|d| (or 1 when |d| = 0).  Only ov is kept in registers; dn and the
// double origin are rebuilt from the float inputs when a triangle is
// certified, bit-identically.
struct RayState {
    float ox, oy, oz;      // float origin (exact input)
    float dx, dy, dz;      // direction as given
    float ix, iy, iz;      // 1/d with d == 0 -> FLT_MAX (BVH.h:1212-1225)
    unsigned oinv;         // (~octant) & 7
    double ov;             // Normalize's 1/|d|
    double len_d;          // |d| in double
};

RISE_HD inline double dn0(const RayState& r) { return (double)r.dx * r.ov; }
RISE_HD inline double dn1(const RayState& r) { return (double)r.dy * r.ov; }
RISE_HD inline double dn2(const RayState& r) { return (double)r.dz * r.ov; }

RISE_HD inline float safe_inv(float d)
{
    if (d == 0.0f) return FLT_MAX;
    float r = 1.0f / d;
    if (!(fabsf(r) <= FLT_MAX)) r = copysignf(FLT_MAX, d);
    return r;
}

// Vector3Ops::Normalize (Utilities/Math3D/VectorsOps.h:229-238).
RISE_HD inline void normalize_rise(const double v[3], double out[3])
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

RISE_HD inline void init_ray(RayState& r, float ox, float oy, float oz, float dx, float dy, float dz)
{
    r.ox = ox;
    r.oy = oy;
    r.oz = oz;
    r.ix = safe_inv(dx);
    r.iy = safe_inv(dy);
    r.iz = safe_inv(dz);
    const unsigned oct = (dx < 0.0f ? 1u : 0u) | (dy < 0.0f ? 2u : 0u) | (dz < 0.0f ? 4u : 0u);
    r.oinv = (~oct) & 7u;
    r.dx = dx;
    r.dy = dy;
    r.dz = dz;
    const double d[3] = {dx, dy, dz};
    const double mag = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    r.ov = mag > 0.0 ? 1.0 / mag : 1.0;
    r.len_d = mag;
}

// FP64 Moller-Trumbore, RayTriangleIntersection.cpp:21-68 with the edges of
// ...Specializations.h:141-142 (computed in double from the vertices).
RISE_HD inline bool certify(const RayState& r, const float4& a, const float4& b, const float4& c, double& t_out)
{
    const double v0x = a.x, v0y = a.y, v0z = a.z;
    const double eAx = (double)b.x - v0x, eAy = (double)b.y - v0y, eAz = (double)b.z - v0z;
    const double eBx = (double)c.x - v0x, eBy = (double)c.y - v0y, eBz = (double)c.z - v0z;
    const double d[3] = {dn0(r), dn1(r), dn2(r)};
    const double p0 = d[1] * eBz - d[2] * eBy;
    const double p1 = d[2] * eBx - d[0] * eBz;
    const double p2 = d[0] * eBy - d[1] * eBx;
    const double det = eAx * p0 + eAy * p1 + eAz * p2;
    if (det > -kNearZero && det < kNearZero) return false;
    const double oodet = 1.0 / det;
    const double t0 = (double)r.ox - v0x, t1 = (double)r.oy - v0y, t2 = (double)r.oz - v0z;
    const double al = (t0 * p0 + t1 * p1 + t2 * p2) * oodet;
    if ((al < 0.0 - kNearZero) || (al > 1.0 + kNearZero)) return false;
    const double q0 = t1 * eAz - t2 * eAy;
    const double q1 = t2 * eAx - t0 * eAz;
    const double q2 = t0 * eAy - t1 * eAx;
    const double be = (d[0] * q0 + d[1] * q1 + d[2] * q2) * oodet;
    if ((be < 0.0 - kNearZero) || ((al + be) > 1.0 + kNearZero)) return false;
    const double tt = (eBx * q0 + eBy * q1 + eBz * q2) * oodet;
    if (tt >= kNearZero) {
        t_out = tt;
        return true;
    }
    return false;
}

// World range of a committed hit: Object.cpp:641-643 for an identity
// transform (|origin + dn * (t - 1e-12) - origin|).
RISE_HD inline double world_range(const RayState& r, double t)
{
    const double s = t - kSurfaceBackoff;
    const double ox = r.ox, oy = r.oy, oz = r.oz;
    const double px = ox + dn0(r) * s, py = oy + dn1(r) * s, pz = oz + dn2(r) * s;
    const double vx = px - ox, vy = py - oy, vz = pz - oz;
    return sqrt(vx * vx + vy * vy + vz * vz);
}

RISE_HD inline unsigned byte_of(unsigned w, int k) { return (w >> (8 * k)) & 0xffu; }

// Tests the eight children of one node.  Returns the permuted internal-hit
// mask (bit k <-> slot k ^ oinv) and fills the node's leaf-triangle mask.
RISE_HD inline unsigned test_node(const RayState& r, const NodeSlot* __restrict__ nodes, unsigned ni, float tmax,
                                     unsigned& child_base, unsigned& imask, unsigned& tri_base, unsigned& tmask)
{
    const uint4* p = reinterpret_cast<const uint4*>(nodes + ni);
    const uint4 w0 = p[0], w1 = p[1], w2 = p[2], w3 = p[3], w4 = p[4];
    const float px = bits_to_float(w0.x), py = bits_to_float(w0.y), pz = bits_to_float(w0.z);
    const int ex = (int)(signed char)(w0.w & 0xff);
    const int ey = (int)(signed char)((w0.w >> 8) & 0xff);
    const int ez = (int)(signed char)((w0.w >> 16) & 0xff);
    imask = (w0.w >> 24) & 0xffu;
    child_base = w1.x;
    tri_base = w1.y;
    const float sx = bits_to_float((unsigned)(ex + 127) << 23);
    const float sy = bits_to_float((unsigned)(ey + 127) << 23);
    const float sz = bits_to_float((unsigned)(ez + 127) << 23);
    // Child planes relative to the origin, pre-scaled: (p + q*s - o) * inv.
    unsigned hits = 0;
    tmask = 0;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const unsigned meta = byte_of(s < 4 ? w1.z : w1.w, s & 3);
        const int k = s & 3;
        const bool hi4 = s >= 4;
        const float qlx = (float)byte_of(hi4 ? w2.y : w2.x, k);
        const float qly = (float)byte_of(hi4 ? w2.w : w2.z, k);
        const float qlz = (float)byte_of(hi4 ? w3.y : w3.x, k);
        const float qhx = (float)byte_of(hi4 ? w3.w : w3.z, k);
        const float qhy = (float)byte_of(hi4 ? w4.y : w4.x, k);
        const float qhz = (float)byte_of(hi4 ? w4.w : w4.z, k);
        const float lx = fmaf(qlx, sx, px), ly = fmaf(qly, sy, py), lz = fmaf(qlz, sz, pz);
        const float hx = fmaf(qhx, sx, px), hy = fmaf(qhy, sy, py), hz = fmaf(qhz, sz, pz);
        const float t1x = (lx - r.ox) * r.ix, t2x = (hx - r.ox) * r.ix;
        const float t1y = (ly - r.oy) * r.iy, t2y = (hy - r.oy) * r.iy;
        const float t1z = (lz - r.oz) * r.iz, t2z = (hz - r.oz) * r.iz;
        const float tn = fmaxf(fmaxf(fminf(t1x, t2x), fminf(t1y, t2y)), fmaxf(fminf(t1z, t2z), 0.0f));
        const float tf = fminf(fminf(fmaxf(t1x, t2x), fmaxf(t1y, t2y)), fminf(fmaxf(t1z, t2z), tmax));
        const bool hit = (meta != 0) && (tn <= tf * RISE_SLAB_PAD);
        if (hit) {
            if ((imask >> s) & 1u) {
                hits |= 1u << (s ^ r.oinv);
            } else {
                const unsigned cnt = meta >> 5, off = meta & 31u;
                tmask |= ((1u << cnt) - 1u) << off;
            }
        }
    }
    return hits;
}

// Converts a best distance along dn into the float traversal parameter along
// d, widened (RISE_SLAB_PAD); kRiseInfinity maps to +inf.
RISE_HD inline float box_limit(const RayState& r, double best)
{
    if (!(best < kRiseInfinity)) return INFINITY;
    const double s = best / r.len_d;
    return (float)s * RISE_SLAB_PAD;
}

template <int STACK>
struct LdsStack {
    uint2* base;   // STACK rows of blockDim.x entries
    int stride;
    int sp;
    RISE_HD inline void push(uint2 v) { base[sp++ * stride] = v; }
    RISE_HD inline uint2 pop() { return base[--sp * stride]; }
    RISE_HD inline bool empty() const { return sp == 0; }
};

struct PrivStack {
    uint2 e[kMaxStack];
    int sp;
    RISE_HD inline void push(uint2 v) { e[sp++] = v; }
    RISE_HD inline uint2 pop() { return e[--sp]; }
    RISE_HD inline bool empty() const { return sp == 0; }
};

// Traversal state of one closest-hit query: the current node group (base,
// 8-bit hit mask permuted by the ray octant, 8-bit internal mask), the best
// certified hit and the box limit derived from it.
struct ClosestState {
    unsigned gbase, ghits, gimask;
    unsigned best_leaf;
    int best_id;
    float tlim;
    double best_t;
};

RISE_HD inline void closest_begin(const RayState& r, ClosestState& s)
{
    s.gbase = 0;
    s.ghits = 1u << (0u ^ r.oinv);  // the root, in slot 0 of a virtual group
    s.gimask = 1u;
    s.best_leaf = 0xffffffffu;
    s.best_id = -1;
    s.tlim = INFINITY;
    s.best_t = kRiseInfinity;
}

// One traversal step: visit the next node of the current group (near-to-far
// by octant) and certify the triangles of its leaf children; when the group
// is used up, pop the next one.  Returns false once the traversal is over.
// Closest hit: smallest t, exact ties to the lowest caller id (the
// reference's strict-closer guard, ...Specializations.h:172, plus the tie rule).
template <bool COUNT, class Stack>
RISE_HD inline bool closest_step(const RayState& r, const NodeSlot* __restrict__ nodes,
                                 const Tri48* __restrict__ tris, Stack& st, ClosestState& s, unsigned& n_nodes,
                                 unsigned& n_tris)
{
    unsigned tmask = 0, tbase = 0;
    if (s.ghits) {
        const unsigned k = __builtin_ctz(s.ghits);
        s.ghits &= s.ghits - 1u;
        const unsigned slot = k ^ r.oinv;
        const unsigned ni = s.gbase + __builtin_popcount(s.gimask & ((1u << slot) - 1u));
        if (s.ghits) st.push(make_uint2(s.gbase, s.ghits | (s.gimask << 8)));
        if (COUNT) ++n_nodes;
        unsigned cb, im;
        s.ghits = test_node(r, nodes, ni, s.tlim, cb, im, tbase, tmask);
        s.gbase = cb;
        s.gimask = im;
    }
    while (tmask) {
        const unsigned j = __builtin_ctz(tmask);
        tmask &= tmask - 1u;
        const float4* tp = reinterpret_cast<const float4*>(tris + tbase + j);
        const float4 a = tp[0], b = tp[1], c = tp[2];
        if (COUNT) ++n_tris;
        double tt;
        if (certify(r, a, b, c, tt)) {
            const int id = float_to_bits(a.w);
            if (tt < s.best_t || (tt == s.best_t && id < s.best_id)) {
                s.best_t = tt;
                s.best_id = id;
                s.best_leaf = tbase + j;
                s.tlim = box_limit(r, s.best_t);
            }
        }
    }
    if (!s.ghits) {
        if (st.empty()) return false;
        const uint2 g = st.pop();
        s.gbase = g.x;
        s.ghits = g.y & 0xffu;
        s.gimask = (g.y >> 8) & 0xffu;
    }
    return true;
}

// Closest-hit traversal.  Returns the caller id (-1 on miss), the best
// parametric distance along dn in best_t and the leaf-order triangle index
// of the hit in best_leaf.
template <bool COUNT, class Stack>
RISE_HD inline int trace_closest(const RayState& r, const NodeSlot* __restrict__ nodes,
                                 const Tri48* __restrict__ tris, Stack& st, double& best_t, unsigned& best_leaf,
                                 unsigned& n_nodes, unsigned& n_tris)
{
    ClosestState s;
    closest_begin(r, s);
    while (closest_step<COUNT>(r, nodes, tris, st, s, n_nodes, n_tris)) {
    }
    best_t = s.best_t;
    best_leaf = s.best_leaf;
    return s.best_id;
}

// Any-hit state: the current group, the box limit and the result.
struct AnyState {
    unsigned gbase, ghits, gimask;
    bool hit;
    float tlim;
    double how_far2;
};

RISE_HD inline void any_begin(const RayState& r, double how_far2, AnyState& s)
{
    s.gbase = 0;
    s.ghits = 1u << (0u ^ r.oinv);
    s.gimask = 1u;
    s.hit = false;
    s.how_far2 = how_far2;
    s.tlim = box_limit(r, how_far2);
}

// One any-hit step; true iff some triangle certifies with
// kNearZero < t < how_far2 (...Specializations.h:517-519) and casts shadows
// (bit 0 of v2.w clear; ObjectManager.cpp:119-124).  Returns false once
// the query is decided (a hit, or the tree exhausted).
template <bool COUNT, class Stack>
RISE_HD inline bool any_step(const RayState& r, const NodeSlot* __restrict__ nodes, const Tri48* __restrict__ tris,
                             Stack& st, AnyState& s, unsigned& n_nodes, unsigned& n_tris)
{
    unsigned tmask = 0, tbase = 0;
    if (s.ghits) {
        const unsigned k = __builtin_ctz(s.ghits);
        s.ghits &= s.ghits - 1u;
        const unsigned slot = k ^ r.oinv;
        const unsigned ni = s.gbase + __builtin_popcount(s.gimask & ((1u << slot) - 1u));
        if (s.ghits) st.push(make_uint2(s.gbase, s.ghits | (s.gimask << 8)));
        if (COUNT) ++n_nodes;
        unsigned cb, im;
        s.ghits = test_node(r, nodes, ni, s.tlim, cb, im, tbase, tmask);
        s.gbase = cb;
        s.gimask = im;
    }
    while (tmask) {
        const unsigned j = __builtin_ctz(tmask);
        tmask &= tmask - 1u;
        const float4* tp = reinterpret_cast<const float4*>(tris + tbase + j);
        const float4 a = tp[0], b = tp[1], c = tp[2];
        if (COUNT) ++n_tris;
        double tt;
        if (certify(r, a, b, c, tt) && tt > kNearZero && tt < s.how_far2 && !(float_to_bits(c.w) & 1u)) {
            s.hit = true;
            return false;
        }
    }
    if (!s.ghits) {
        if (st.empty()) return false;
        const uint2 g = st.pop();
        s.gbase = g.x;
        s.ghits = g.y & 0xffu;
        s.gimask = (g.y >> 8) & 0xffu;
    }
    return true;
}

template <bool COUNT, class Stack>
RISE_HD inline bool trace_any(const RayState& r, const NodeSlot* __restrict__ nodes, const Tri48* __restrict__ tris,
                              Stack& st, double how_far2, unsigned& n_nodes, unsigned& n_tris)
{
    AnyState s;
    any_begin(r, how_far2, s);
    while (any_step<COUNT>(r, nodes, tris, st, s, n_nodes, n_tris)) {
    }
    return s.hit;
}

// Object.cpp:472-493: the world limit dHowFar becomes dHowFar2 = |d| * dHowFar
// in the normalized frame, unless it is the RISE_INFINITY sentinel and |d| >= 1.
RISE_HD inline double scale_how_far(const RayState& r, double how_far)
{
    const double factor = (r.len_d > kNearZero) ? r.len_d : 1.0;
    if ((how_far != kRiseInfinity) || (factor < 1.0)) return factor * how_far;
    return how_far;
}

}  // namespace rise_hip
