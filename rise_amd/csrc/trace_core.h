// trace_core.h -- device-side traversal of the 8-wide quantized BVH.
//
// One lane traces one ray.  Boxes are tested in FP32 with a conservative
// slab test; every triangle whose leaf box is reached goes through the
// reference's FP32 Moller-Trumbore pre-filter (Acceleration/BVH.h:1111-1167,
// the same float operations in the same order) and, when the filter passes
// it, is certified in FP64 with the reference's Moller-Trumbore
// (Intersection/RayTriangleIntersection.cpp:21-68), so the committed hit is
// the one the reference's leaf (BVH.h:1237-1301) would commit.  Closest-hit
// keeps the smallest t, exact ties going to the lowest triangle id; the
// reference's strict-closer guard is
// Geometry/TriangleMeshGeometryIndexedSpecializations.h:172.
//
// The filter's distance test (t > current best + 1e-5) depends on the best
// hit found so far, hence on traversal order; it can change the committed
// hit only when the winner's float distance exceeds its double distance by
// more than 1e-5 (the oracle flags such rays; none occur on the 1M-triangle
// workload).
//
// Stack: node groups (base index, 8-bit hit mask, 8-bit internal mask), at
// most one per tree level, held in LDS (short stack, STACK entries per lane)
// or, for deep trees, in a private array.
//
// The whole file is compiled with -ffp-contract=off: float and double
// operations are rounded one by one in the order the reference writes them.
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

// 1 + 2^-19: relative widening of the FP32 slab interval.  After the absolute
// slack of each plane (slab_axis below) the computed entry and exit distances
// are within 2.0001 * 2^-24 (relative) of the exact ones, so scaling the exit
// distance by this factor (and the running best by it again) never rejects a
// box the exact ray enters.
#define RISE_SLAB_PAD 1.0000019073486328f

// The slab test runs along d * 2^k, the direction scaled by the power of two
// that puts its largest component in [1, 2).  A component whose scaled size is
// below 2^-80 (zero included) takes the inverse +-2^80 instead of 1/d: that
// shrinks its slab interval toward t = 0 only by a factor the box padding
// covers.  Every leaf box is padded by at least 1e-9 of the scene extent, so
// a ray that can hit a triangle of the box lies at least that far inside the
// box along such an axis, and its interval still reaches 1e-9 * extent * 2^80
// (> 1e15 scene extents) -- beyond any hit, since t along the scaled direction
// is at most the distance travelled along its largest axis.
constexpr float kInvClamp = 0x1p80f;

// Arithmetic of the reference build whose results are reproduced.
//   kArithIeee: every float and double operation rounded on its own, in the
//     order the reference's source writes it (its code built without
//     contraction).
//   kArithX86Fma: the reference as its Linux build compiles it (g++ -O3
//     -march=x86-64-v3, contraction on).  That build fuses the first product
//     of a*b - c*d into fma(a, b, -(c*d)) and evaluates a*b + c*d + e*f as
//     fma(e, f, fma(a, b, c*d)), in Normalize, the float pre-filter, the
//     double certifier and the world range alike (read from the compiled
//     BVH::IntersectRay4, RayTriangleIntersection and Vector3Ops code).
// Both forms are exact IEEE operations on the GPU (v_fma_f32 / v_fma_f64).
constexpr int kArithIeee = 0;
constexpr int kArithX86Fma = 1;

template <int A>
RISE_HD inline float sub_prod(float a, float b, float c, float d)
{
    if (A == kArithX86Fma) return fmaf(a, b, -(c * d));
    return a * b - c * d;
}
template <int A>
RISE_HD inline double sub_prod(double a, double b, double c, double d)
{
    if (A == kArithX86Fma) return fma(a, b, -(c * d));
    return a * b - c * d;
}
template <int A>
RISE_HD inline float dot3(float a0, float b0, float a1, float b1, float a2, float b2)
{
    if (A == kArithX86Fma) return fmaf(a2, b2, fmaf(a0, b0, a1 * b1));
    return a0 * b0 + a1 * b1 + a2 * b2;
}
template <int A>
RISE_HD inline double dot3(double a0, double b0, double a1, double b1, double a2, double b2)
{
    if (A == kArithX86Fma) return fma(a2, b2, fma(a0, b0, a1 * b1));
    return a0 * b0 + a1 * b1 + a2 * b2;
}

// The certifier's direction is d normalized in double exactly as
// Vector3Ops::Normalize computes it (Object.cpp:414-425): dn_k = d_k * ov with
// ov = 1/|d| (or 1 when |d| = 0).  Only ov is kept in registers; dn and the
// double origin are rebuilt from the float inputs when a triangle is
// certified, bit-identically.  A selects the arithmetic (kArith*).
template <int A>
struct RayStateT {
    static constexpr int kArith = A;
    float ox, oy, oz;      // float origin (exact input)
    float dx, dy, dz;      // direction as given
    float fx, fy, fz;      // (float)dn: the filter's direction (BVH.h:1212-1225)
    float ix, iy, iz;      // slab inverses 1/(d 2^k) (kInvClamp for tiny components)
    unsigned oinv;         // (~octant) & 7
    double ov;             // Normalize's 1/|d|
    double len_d;          // |d| in double
    double len_s;          // |d| 2^k: t along the scaled direction = distance / len_s
};
using RayState = RayStateT<kArithIeee>;

template <int A>
RISE_HD inline double dn0(const RayStateT<A>& r) { return (double)r.dx * r.ov; }
template <int A>
RISE_HD inline double dn1(const RayStateT<A>& r) { return (double)r.dy * r.ov; }
template <int A>
RISE_HD inline double dn2(const RayStateT<A>& r) { return (double)r.dz * r.ov; }
// The certifier's origin: the float input, exact in double.
template <int A>
RISE_HD inline double org0(const RayStateT<A>& r) { return (double)r.ox; }
template <int A>
RISE_HD inline double org1(const RayStateT<A>& r) { return (double)r.oy; }
template <int A>
RISE_HD inline double org2(const RayStateT<A>& r) { return (double)r.oz; }

// Inverse of a scaled direction component (largest component in [1, 2)):
// correctly rounded 1/d, or +-kInvClamp when |d| < 2^-80 (+ for a zero).
RISE_HD inline float slab_inv(float d)
{
    if (fabsf(d) < 0x1p-80f) return d < 0.0f ? -kInvClamp : kInvClamp;
    return 1.0f / d;
}

// Vector3Ops::Normalize (Utilities/Math3D/VectorsOps.h:229-238) under
// arithmetic model A.
template <int A>
RISE_HD inline void normalize_model(const double v[3], double out[3])
{
    const double mag = sqrt(dot3<A>(v[0], v[0], v[1], v[1], v[2], v[2]));
    if (mag > 0.0) {
        const double ov = 1.0 / mag;
        for (int k = 0; k < 3; ++k) out[k] = v[k] * ov;
    } else {
        for (int k = 0; k < 3; ++k) out[k] = v[k];
    }
}

template <int A>
RISE_HD inline void init_ray(RayStateT<A>& r, float ox, float oy, float oz, float dx, float dy, float dz)
{
    r.ox = ox;
    r.oy = oy;
    r.oz = oz;
    const float m = fmaxf(fmaxf(fabsf(dx), fabsf(dy)), fabsf(dz));
    int k = 0;
    if (m > 0.0f) (void)frexpf(m, &k);  // m = f 2^k, f in [0.5, 1)
    const int sh = 1 - k;                // d 2^sh: largest component in [1, 2)
    r.ix = slab_inv(ldexpf(dx, sh));
    r.iy = slab_inv(ldexpf(dy, sh));
    r.iz = slab_inv(ldexpf(dz, sh));
    const unsigned oct = (dx < 0.0f ? 1u : 0u) | (dy < 0.0f ? 2u : 0u) | (dz < 0.0f ? 4u : 0u);
    r.oinv = (~oct) & 7u;
    r.dx = dx;
    r.dy = dy;
    r.dz = dz;
    const double d[3] = {dx, dy, dz};
    const double mag = sqrt(dot3<A>(d[0], d[0], d[1], d[1], d[2], d[2]));
    r.ov = mag > 0.0 ? 1.0 / mag : 1.0;
    r.len_d = mag;
    r.len_s = mag > 0.0 ? ldexp(mag, sh) : 1.0;
    r.fx = (float)dn0(r);
    r.fy = (float)dn1(r);
    r.fz = (float)dn2(r);
}

// Point3Ops::Transform and Vector3Ops::Transform (Utilities/Math3D/
// PointsOps.h:126-133, VectorsOps.h:241-248) with the row-vector convention of
// Matrix4: m[4 i + j] = m._ij, the translation in row 3.
template <int A>
RISE_HD inline void xform_point(const double* m, const double p[3], double out[3])
{
    for (int j = 0; j < 3; ++j) out[j] = dot3<A>(m[j], p[0], m[4 + j], p[1], m[8 + j], p[2]) + m[12 + j];
}
template <int A>
RISE_HD inline void xform_vector(const double* m, const double v[3], double out[3])
{
    for (int j = 0; j < 3; ++j) out[j] = dot3<A>(m[j], v[0], m[4 + j], v[1], m[8 + j], v[2]);
}
// Vector3Ops::Transform by the inverse transpose (Object.cpp:531-533, with
// m_mxInvTranspose = Transpose(m_mxInvFinalTrans), Object.cpp:827), from the
// inverse itself: out_j = minv._j0 n.x + minv._j1 n.y + minv._j2 n.z.
template <int A>
RISE_HD inline void xform_normal(const double* minv, const double n[3], double out[3])
{
    for (int j = 0; j < 3; ++j) out[j] = dot3<A>(minv[4 * j], n[0], minv[4 * j + 1], n[1], minv[4 * j + 2], n[2]);
}

// Object-space ray of a transformed object, as Object::IntersectRay and
// _IntersectionOnly make it (Objects/Object.cpp:414-425, :666-674):
// o' = Transform(M^-1, o), dl = Transform(M^-1, d) for the world direction d
// as given, dn' = Normalize(dl), all in double, and factor = |dl| (1 when it
// is not above NEARZERO, :472-485).  The float pre-filter sees (float)o' and
// (float)dn' (PrepRayFloat, BVH.h:1212-1225); the certifier sees o' and dn'.
// The boxes are tested with the FP32 slab test of the float ray ((float)o',
// (float)dn' scaled as in init_ray); o' is not a float, so each axis's plane
// distances take the origin's rounding |o' - (float)o'| / d as extra slack
// (the direction's rounding is relative, covered by RISE_SLAB_PAD).
template <int A>
struct ObjRayT {
    static constexpr int kArith = A;
    float ox, oy, oz;  // (float)o'
    float fx, fy, fz;  // (float)dn'
    float ix, iy, iz;  // slab inverses of (float)dn' scaled as in init_ray
    float ex, ey, ez;  // extra slack per axis: |o' - (float)o'| |inverse| (1 + 2^-20)
    unsigned oinv;     // (~octant of (float)dn') & 7
    double o[3];       // o'
    double dn[3];      // dn'
    double len_s;      // |(float)dn'| 2^k: t along the scaled direction = distance / len_s
    double mag;        // |dl|
    double factor;     // |dl|, or 1 when |dl| <= NEARZERO
};

template <int A>
RISE_HD inline double dn0(const ObjRayT<A>& r) { return r.dn[0]; }
template <int A>
RISE_HD inline double dn1(const ObjRayT<A>& r) { return r.dn[1]; }
template <int A>
RISE_HD inline double dn2(const ObjRayT<A>& r) { return r.dn[2]; }
template <int A>
RISE_HD inline double org0(const ObjRayT<A>& r) { return r.o[0]; }
template <int A>
RISE_HD inline double org1(const ObjRayT<A>& r) { return r.o[1]; }
template <int A>
RISE_HD inline double org2(const ObjRayT<A>& r) { return r.o[2]; }

// minv: the object's m_mxInvFinalTrans (16 doubles, Matrix4 order).
template <int A>
RISE_HD inline void init_obj_ray(ObjRayT<A>& r, const double ow[3], const double dw[3], const double* minv)
{
    xform_point<A>(minv, ow, r.o);
    double dl[3];
    xform_vector<A>(minv, dw, dl);
    const double mag = sqrt(dot3<A>(dl[0], dl[0], dl[1], dl[1], dl[2], dl[2]));
    r.mag = mag;
    r.factor = mag > kNearZero ? mag : 1.0;
    normalize_model<A>(dl, r.dn);
    r.ox = (float)r.o[0];
    r.oy = (float)r.o[1];
    r.oz = (float)r.o[2];
    r.fx = (float)r.dn[0];
    r.fy = (float)r.dn[1];
    r.fz = (float)r.dn[2];
    const float m = fmaxf(fmaxf(fabsf(r.fx), fabsf(r.fy)), fabsf(r.fz));
    int k = 0;
    if (m > 0.0f) (void)frexpf(m, &k);
    const int sh = 1 - k;
    r.ix = slab_inv(ldexpf(r.fx, sh));
    r.iy = slab_inv(ldexpf(r.fy, sh));
    r.iz = slab_inv(ldexpf(r.fz, sh));
    const unsigned oct = (r.fx < 0.0f ? 1u : 0u) | (r.fy < 0.0f ? 2u : 0u) | (r.fz < 0.0f ? 4u : 0u);
    r.oinv = (~oct) & 7u;
    const double f[3] = {r.fx, r.fy, r.fz};
    const double lf = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    r.len_s = lf > 0.0 ? ldexp(lf, sh) : 1.0;
    r.ex = (float)(fabs(r.o[0] - (double)r.ox) * fabs((double)r.ix) * (1.0 + 0x1p-20));
    r.ey = (float)(fabs(r.o[1] - (double)r.oy) * fabs((double)r.iy) * (1.0 + 0x1p-20));
    r.ez = (float)(fabs(r.o[2] - (double)r.oz) * fabs((double)r.iz) * (1.0 + 0x1p-20));
}

// BVH::MollerTrumboreFloat (Acceleration/BVH.h:1111-1167) on the leaf
// triangle's float vertices: p0 = v0, e1 = v1 - v0, e2 = v2 - v0 in float
// (BuildFastFilter, BVH.h:431-467), the float origin and the float
// normalized direction (PrepRayFloat, BVH.h:1212-1225).  `best` is the
// caller's (float) current range.  Every operation is one IEEE float
// operation in the reference's order (no contraction; 1/det correctly
// rounded).  A triangle it rejects is never certified.
template <class R>
RISE_HD inline bool filter_float(const R& r, const float4& a, const float4& b, const float4& c, float best)
{
    constexpr int A = R::kArith;
    constexpr float kEps = 1e-5f;
    const float e1x = b.x - a.x, e1y = b.y - a.y, e1z = b.z - a.z;
    const float e2x = c.x - a.x, e2y = c.y - a.y, e2z = c.z - a.z;
    const float pv0 = sub_prod<A>(r.fy, e2z, r.fz, e2y);
    const float pv1 = sub_prod<A>(r.fz, e2x, r.fx, e2z);
    const float pv2 = sub_prod<A>(r.fx, e2y, r.fy, e2x);
    const float det = dot3<A>(e1x, pv0, e1y, pv1, e1z, pv2);
    if (fabsf(det) < 1e-20f) return false;
    const float inv = 1.0f / det;
    const float tv0 = r.ox - a.x, tv1 = r.oy - a.y, tv2 = r.oz - a.z;
    const float u = dot3<A>(tv0, pv0, tv1, pv1, tv2, pv2) * inv;
    if (u < -kEps || u > 1.0f + kEps) return false;
    const float qv0 = sub_prod<A>(tv1, e1z, tv2, e1y);
    const float qv1 = sub_prod<A>(tv2, e1x, tv0, e1z);
    const float qv2 = sub_prod<A>(tv0, e1y, tv1, e1x);
    const float v = dot3<A>(r.fx, qv0, r.fy, qv1, r.fz, qv2) * inv;
    if (v < -kEps || u + v > 1.0f + kEps) return false;
    const float t = dot3<A>(e2x, qv0, e2y, qv1, e2z, qv2) * inv;
    if (t < -kEps || t > best + kEps) return false;
    return true;
}

// FP64 Moller-Trumbore, RayTriangleIntersection.cpp:21-68, on v0 and the
// edges eA = v1 - v0, eB = v2 - v0 of ...Specializations.h:141-142 (computed
// in double from the vertices).
template <class R>
RISE_HD inline bool certify_edges(const R& r, double v0x, double v0y, double v0z, double eAx, double eAy, double eAz,
                                  double eBx, double eBy, double eBz, double& t_out, double* ab)
{
    constexpr int A = R::kArith;
    const double d[3] = {dn0(r), dn1(r), dn2(r)};
    const double p0 = sub_prod<A>(d[1], eBz, d[2], eBy);
    const double p1 = sub_prod<A>(d[2], eBx, d[0], eBz);
    const double p2 = sub_prod<A>(d[0], eBy, d[1], eBx);
    const double det = dot3<A>(eAx, p0, eAy, p1, eAz, p2);
    if (det > -kNearZero && det < kNearZero) return false;
    const double oodet = 1.0 / det;
    const double t0 = org0(r) - v0x, t1 = org1(r) - v0y, t2 = org2(r) - v0z;
    const double al = dot3<A>(t0, p0, t1, p1, t2, p2) * oodet;
    if ((al < 0.0 - kNearZero) || (al > 1.0 + kNearZero)) return false;
    const double q0 = sub_prod<A>(t1, eAz, t2, eAy);
    const double q1 = sub_prod<A>(t2, eAx, t0, eAz);
    const double q2 = sub_prod<A>(t0, eAy, t1, eAx);
    const double be = dot3<A>(d[0], q0, d[1], q1, d[2], q2) * oodet;
    if ((be < 0.0 - kNearZero) || ((al + be) > 1.0 + kNearZero)) return false;
    const double tt = dot3<A>(eBx, q0, eBy, q1, eBz, q2) * oodet;
    if (tt >= kNearZero) {
        t_out = tt;
        if (ab) {
            ab[0] = al;
            ab[1] = be;
        }
        return true;
    }
    return false;
}

// The certifier on the leaf record's float vertices (exact in double).
template <class R>
RISE_HD inline bool certify(const R& r, const float4& a, const float4& b, const float4& c, double& t_out,
                            double* ab = nullptr)
{
    return certify_edges(r, a.x, a.y, a.z, (double)b.x - a.x, (double)b.y - a.y, (double)b.z - a.z,
                         (double)c.x - a.x, (double)c.y - a.y, (double)c.z - a.z, t_out, ab);
}

// The certifier on a triangle's double vertices v[0..8] (v0, v1, v2), as a
// mesh with double vertices holds them (PointerTriangle, Polygon.h:142).
template <class R>
RISE_HD inline bool certify64(const R& r, const double* __restrict__ v, double& t_out, double* ab = nullptr)
{
    return certify_edges(r, v[0], v[1], v[2], v[3] - v[0], v[4] - v[1], v[5] - v[2], v[6] - v[0], v[7] - v[1],
                         v[8] - v[2], t_out, ab);
}

// The face normal eA x eB as RayElementIntersection computes it
// (...Specializations.h:141-159, :196-205).  The Linux build contracts its x
// and y components as fma(-(a.z), b.y, a.y b.z) and fma(-(a.x), b.z, a.z b.x)
// and its z component as fma(a.x, b.y, -(a.y b.x)) (read from the compiled
// code; with float vertices every product of edges is exact and the forms
// agree, double vertices tell them apart).
template <int A>
RISE_HD inline void face_cross(double ax, double ay, double az, double bx, double by, double bz, double n[3])
{
    if (A == kArithX86Fma) {
        n[0] = fma(-az, by, ay * bz);
        n[1] = fma(-ax, bz, az * bx);
        n[2] = fma(ax, by, -(ay * bx));
    } else {
        n[0] = ay * bz - az * by;
        n[1] = az * bx - ax * bz;
        n[2] = ax * by - ay * bx;
    }
}

// Face selection of TriangleMeshGeometryIndexed::RayElementIntersection
// (...Specializations.h:141-159): the face normal against the normalized
// direction; dot < 0 is a front face.  faces: bit 0 hit front faces, bit 1
// hit back faces.
template <class R>
RISE_HD inline bool face_ok_edges(const R& r, double eAx, double eAy, double eAz, double eBx, double eBy, double eBz,
                                  unsigned faces)
{
    constexpr int A = R::kArith;
    double n[3];
    face_cross<A>(eAx, eAy, eAz, eBx, eBy, eBz, n);
    const double dd = dot3<A>(n[0], dn0(r), n[1], dn1(r), n[2], dn2(r));
    if (!(faces & 1u) && dd < 0.0) return false;
    if (!(faces & 2u) && dd > 0.0) return false;
    return true;
}

template <class R>
RISE_HD inline bool face_ok(const R& r, const float4& a, const float4& b, const float4& c, unsigned faces)
{
    return face_ok_edges(r, (double)b.x - a.x, (double)b.y - a.y, (double)b.z - a.z, (double)c.x - a.x,
                         (double)c.y - a.y, (double)c.z - a.z, faces);
}

template <class R>
RISE_HD inline bool face_ok64(const R& r, const double* __restrict__ v, unsigned faces)
{
    return face_ok_edges(r, v[3] - v[0], v[4] - v[1], v[5] - v[2], v[6] - v[0], v[7] - v[1], v[8] - v[2], faces);
}

// World range of a committed hit: Object.cpp:641-643 for an identity
// transform (|origin + dn * (t - 1e-12) - origin|).
template <int A>
RISE_HD inline double world_range(const RayStateT<A>& r, double t)
{
    const double s = t - kSurfaceBackoff;
    const double ox = r.ox, oy = r.oy, oz = r.oz;
    double px, py, pz;
    if (A == kArithX86Fma) {
        px = fma(dn0(r), s, ox);
        py = fma(dn1(r), s, oy);
        pz = fma(dn2(r), s, oz);
    } else {
        px = ox + dn0(r) * s;
        py = oy + dn1(r) * s;
        pz = oz + dn2(r) * s;
    }
    const double vx = px - ox, vy = py - oy, vz = pz - oz;
    return sqrt(dot3<A>(vx, vx, vy, vy, vz, vz));
}

RISE_HD inline unsigned byte_of(unsigned w, int k) { return (w >> (8 * k)) & 0xffu; }

// One axis of a node's slab test in fused form: the plane at quantized
// coordinate q lies at distance t(q) = q s + b along the scaled direction,
// with s = 2^e / d (exact scaling of the inverse) and b = (p - o) / d.  The
// form rounds b before adding q s, so t carries an absolute error of up to
// 2.0001 * 2^-24 |b| besides its relative error (2.0001 * 2^-24 |t|): the
// near plane subtracts, the far plane adds, 2^-21 |b| + 2^-100 (which also
// covers flushed subnormals), leaving only the relative error for
// RISE_SLAB_PAD.  Should b or s be too large for q s + b to stay finite
// (|b| < 2^126 and |s| < 2^119 keep 255 |s| + |b| (1 + 2^-21) below 2^128),
// the axis constrains nothing (s = 0, planes at -inf / +inf): still
// conservative.  With the 2^80 inverse of a zero direction component that
// happens only for nodes 2^46 away or 2^39 wide.
struct SlabAxis {
    float s, bn, bf;
};

RISE_HD inline SlabAxis slab_axis(float p, int e, float o, float inv)
{
    const float b = (p - o) * inv;
    const float s = ldexpf(inv, e);
    const float slack = fmaf(fabsf(b), 0x1p-21f, 0x1p-100f);
    if (fabsf(b) < 0x1p126f && fabsf(s) < 0x1p119f) return SlabAxis{s, b - slack, b + slack};
    return SlabAxis{0.0f, -INFINITY, INFINITY};
}

// slab_axis for an origin known only to within `extra / |inv|` of o: the
// plane distances widen by `extra` on top of their own slack (the sum, rounded
// to nearest, is below 2^127 while |b| < 2^126 and extra < 2^125).
RISE_HD inline SlabAxis slab_axis_widened(float p, int e, float o, float inv, float extra)
{
    const float b = (p - o) * inv;
    const float s = ldexpf(inv, e);
    const float slack = fmaf(fabsf(b), 0x1p-21f, 0x1p-100f) + extra * (1.0f + 0x1p-22f);
    if (fabsf(b) < 0x1p126f && fabsf(s) < 0x1p119f && extra < 0x1p125f) return SlabAxis{s, b - slack, b + slack};
    return SlabAxis{0.0f, -INFINITY, INFINITY};
}

// The eight child boxes of one node against the axes' plane distances: bit s
// of the result set when slot s is occupied and its box meets [0, tmax]
// (conservatively).  The quantized planes are split by the sign of the
// inverse into near and far words per axis, so each slot costs six byte
// conversions, six FMAs and the interval reductions.
RISE_HD inline unsigned accept_slots(const SlabAxis& ax, const SlabAxis& ay, const SlabAxis& az, bool nx, bool ny,
                                     bool nz, const uint4& w1, const uint4& w2, const uint4& w3, const uint4& w4,
                                     float tmax)
{
    // Low planes: qlox = w2.x|y, qloy = w2.z|w, qloz = w3.x|y; high planes:
    // qhix = w3.z|w, qhiy = w4.x|y, qhiz = w4.z|w (slots 0-3 | 4-7).  With a
    // negative inverse the high plane is the near one.
    const unsigned nx0 = nx ? w3.z : w2.x, nx1 = nx ? w3.w : w2.y;
    const unsigned fx0 = nx ? w2.x : w3.z, fx1 = nx ? w2.y : w3.w;
    const unsigned ny0 = ny ? w4.x : w2.z, ny1 = ny ? w4.y : w2.w;
    const unsigned fy0 = ny ? w2.z : w4.x, fy1 = ny ? w2.w : w4.y;
    const unsigned nz0 = nz ? w4.z : w3.x, nz1 = nz ? w4.w : w3.y;
    const unsigned fz0 = nz ? w3.x : w4.z, fz1 = nz ? w3.y : w4.w;
    unsigned acc = 0;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const int k = s & 3;
        const bool hi4 = s >= 4;
        const unsigned meta = byte_of(s < 4 ? w1.z : w1.w, s & 3);
        const float tnx = fmaf((float)byte_of(hi4 ? nx1 : nx0, k), ax.s, ax.bn);
        const float tny = fmaf((float)byte_of(hi4 ? ny1 : ny0, k), ay.s, ay.bn);
        const float tnz = fmaf((float)byte_of(hi4 ? nz1 : nz0, k), az.s, az.bn);
        const float tfx = fmaf((float)byte_of(hi4 ? fx1 : fx0, k), ax.s, ax.bf);
        const float tfy = fmaf((float)byte_of(hi4 ? fy1 : fy0, k), ay.s, ay.bf);
        const float tfz = fmaf((float)byte_of(hi4 ? fz1 : fz0, k), az.s, az.bf);
        const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, 0.0f));
        const float tf = fminf(fminf(tfx, tfy), fminf(tfz, tmax));
        if ((meta != 0) && (tn <= tf * RISE_SLAB_PAD)) acc |= 1u << s;
    }
    return acc;
}

template <int A>
RISE_HD inline unsigned node_accept_mask(const RayStateT<A>& r, const uint4& w0, const uint4& w1, const uint4& w2,
                                         const uint4& w3, const uint4& w4, float tmax)
{
    const SlabAxis ax = slab_axis(bits_to_float(w0.x), (int)(signed char)(w0.w & 0xff), r.ox, r.ix);
    const SlabAxis ay = slab_axis(bits_to_float(w0.y), (int)(signed char)((w0.w >> 8) & 0xff), r.oy, r.iy);
    const SlabAxis az = slab_axis(bits_to_float(w0.z), (int)(signed char)((w0.w >> 16) & 0xff), r.oz, r.iz);
    return accept_slots(ax, ay, az, r.ix < 0.0f, r.iy < 0.0f, r.iz < 0.0f, w1, w2, w3, w4, tmax);
}

// The same for an object-space ray: the float ray's planes widened by the
// origin's rounding (ObjRayT).
template <int A>
RISE_HD inline unsigned node_accept_mask(const ObjRayT<A>& r, const uint4& w0, const uint4& w1, const uint4& w2,
                                         const uint4& w3, const uint4& w4, float tmax)
{
    const SlabAxis ax = slab_axis_widened(bits_to_float(w0.x), (int)(signed char)(w0.w & 0xff), r.ox, r.ix, r.ex);
    const SlabAxis ay =
        slab_axis_widened(bits_to_float(w0.y), (int)(signed char)((w0.w >> 8) & 0xff), r.oy, r.iy, r.ey);
    const SlabAxis az =
        slab_axis_widened(bits_to_float(w0.z), (int)(signed char)((w0.w >> 16) & 0xff), r.oz, r.iz, r.ez);
    return accept_slots(ax, ay, az, r.ix < 0.0f, r.iy < 0.0f, r.iz < 0.0f, w1, w2, w3, w4, tmax);
}

// Tests the eight children of one node.  Returns the permuted internal-hit
// mask (bit k <-> slot k ^ oinv) and fills the node's leaf-triangle mask.
template <class R>
RISE_HD inline unsigned test_node(const R& r, const NodeSlot* __restrict__ nodes, unsigned ni, float tmax,
                                     unsigned& child_base, unsigned& imask, unsigned& tri_base, unsigned& tmask)
{
    const uint4* p = reinterpret_cast<const uint4*>(nodes + ni);
    const uint4 w0 = p[0], w1 = p[1], w2 = p[2], w3 = p[3], w4 = p[4];
    imask = (w0.w >> 24) & 0xffu;
    child_base = w1.x;
    tri_base = w1.y;
    const unsigned acc = node_accept_mask(r, w0, w1, w2, w3, w4, tmax);
    unsigned hits = 0;
    tmask = 0;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        if ((acc >> s) & 1u) {
            if ((imask >> s) & 1u) {
                hits |= 1u << (s ^ r.oinv);
            } else {
                const unsigned meta = byte_of(s < 4 ? w1.z : w1.w, s & 3);
                const unsigned cnt = meta >> 5, off = meta & 31u;
                tmask |= ((1u << cnt) - 1u) << off;
            }
        }
    }
    return hits;
}

// Converts a best distance along dn into the float traversal parameter along
// the scaled direction, widened (RISE_SLAB_PAD); kRiseInfinity maps to +inf.
template <int A>
RISE_HD inline float box_limit(const RayStateT<A>& r, double best)
{
    if (!(best < kRiseInfinity)) return INFINITY;
    const double s = best / r.len_s;
    return (float)s * RISE_SLAB_PAD;
}

// box_limit for an object-space ray: its slab test runs along the scaled
// float direction, whose length differs from |dn'| = 1 by a relative 2^-24
// (covered by RISE_SLAB_PAD, applied as for the world ray).
template <int A>
RISE_HD inline float box_limit(const ObjRayT<A>& r, double best)
{
    if (!(best < kRiseInfinity)) return INFINITY;
    const double s = best / r.len_s;
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
    unsigned faces;  // face selection (face_ok), read by the culling variant only
};

template <class R>
RISE_HD inline void closest_begin(const R& r, ClosestState& s)
{
    s.gbase = 0;
    s.ghits = 1u << (0u ^ r.oinv);  // the root, in slot 0 of a virtual group
    s.gimask = 1u;
    s.best_leaf = 0xffffffffu;
    s.best_id = -1;
    s.tlim = INFINITY;
    s.best_t = kRiseInfinity;
    s.faces = 3u;
}

// One traversal step: visit the next node of the current group (near-to-far
// by octant) and certify the triangles of its leaf children; when the group
// is used up, pop the next one.  Returns false once the traversal is over.
// Closest hit: smallest t, exact ties to the lowest caller id (the
// reference's strict-closer guard, ...Specializations.h:172, plus the tie rule).
// CULL: apply the face selection in s.faces (face_ok) before the filter.
// F64: the triangles' double vertices are in tri64 (9 doubles per leaf
// record, leaf order) and the certifier and face test use them.
template <bool COUNT, bool CULL = false, bool F64 = false, class Stack, class R>
RISE_HD inline bool closest_step(const R& r, const NodeSlot* __restrict__ nodes,
                                 const Tri48* __restrict__ tris, Stack& st, ClosestState& s, unsigned& n_nodes,
                                 unsigned& n_tris, const double* __restrict__ tri64 = nullptr)
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
        const double* v64 = F64 ? tri64 + 9 * (size_t)(tbase + j) : nullptr;
        if (CULL && s.faces != 3u && !(float_to_bits(c.w) & kTriDoubleSided) &&
            !(F64 ? face_ok64(r, v64, s.faces) : face_ok(r, a, b, c, s.faces)))
            continue;
        double tt;
        if (filter_float(r, a, b, c, (float)s.best_t) && (F64 ? certify64(r, v64, tt) : certify(r, a, b, c, tt))) {
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
template <bool COUNT, class Stack, class R>
RISE_HD inline int trace_closest(const R& r, const NodeSlot* __restrict__ nodes,
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
    unsigned faces;  // face selection (face_ok), read by the culling variant only
    bool hit;
    float tlim;
    float fbest;  // (float)dHowFar2: the filter's range (BVH.h:1283)
    double how_far2;
};

template <class R>
RISE_HD inline void any_begin(const R& r, double how_far2, AnyState& s)
{
    s.gbase = 0;
    s.ghits = 1u << (0u ^ r.oinv);
    s.gimask = 1u;
    s.faces = 3u;
    s.hit = false;
    s.how_far2 = how_far2;
    s.fbest = (float)how_far2;
    s.tlim = box_limit(r, how_far2);
}

// One any-hit step; true iff some triangle certifies with
// kNearZero < t < how_far2 (...Specializations.h:517-519) and casts shadows
// (bit 0 of v2.w clear; ObjectManager.cpp:119-124).  Returns false once
// the query is decided (a hit, or the tree exhausted).
template <bool COUNT, bool CULL = false, bool F64 = false, class Stack, class R>
RISE_HD inline bool any_step(const R& r, const NodeSlot* __restrict__ nodes, const Tri48* __restrict__ tris,
                             Stack& st, AnyState& s, unsigned& n_nodes, unsigned& n_tris,
                             const double* __restrict__ tri64 = nullptr)
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
        const double* v64 = F64 ? tri64 + 9 * (size_t)(tbase + j) : nullptr;
        if (CULL && s.faces != 3u && !(float_to_bits(c.w) & kTriDoubleSided) &&
            !(F64 ? face_ok64(r, v64, s.faces) : face_ok(r, a, b, c, s.faces)))
            continue;
        double tt;
        if (filter_float(r, a, b, c, s.fbest) && (F64 ? certify64(r, v64, tt) : certify(r, a, b, c, tt)) &&
            tt > kNearZero && tt < s.how_far2 &&
            !(float_to_bits(c.w) & 1u)) {
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

template <bool COUNT, class Stack, class R>
RISE_HD inline bool trace_any(const R& r, const NodeSlot* __restrict__ nodes, const Tri48* __restrict__ tris,
                              Stack& st, double how_far2, unsigned& n_nodes, unsigned& n_tris)
{
    AnyState s;
    any_begin(r, how_far2, s);
    while (any_step<COUNT>(r, nodes, tris, st, s, n_nodes, n_tris)) {
    }
    return s.hit;
}

// Attributes of a committed hit on triangle `tri` at distance t along dn,
// as TriangleMeshGeometryIndexed::RayElementIntersection and
// Object::IntersectRay leave them for a mesh without vertex normals or
// texture coordinates under an identity transform: the certifier's
// barycentrics, the face normal (v1 - v0) x (v2 - v0) normalized
// (...Specializations.h:196-205) and normalized again by the object's normal
// transform (Object.cpp:531-533), and the hit point o + dn (t - 1e-12)
// (Object.cpp:637-640).  v64: the triangle's double vertices, or NULL for
// the leaf record's floats.
template <class R>
RISE_HD inline void triangle_vertices(const Tri48* __restrict__ tri, const double* __restrict__ v64, double v[9])
{
    if (v64) {
        for (int k = 0; k < 9; ++k) v[k] = v64[k];
        return;
    }
    const float4* tp = reinterpret_cast<const float4*>(tri);
    const float4 a = tp[0], b = tp[1], c = tp[2];
    const double w[9] = {a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z};
    for (int k = 0; k < 9; ++k) v[k] = w[k];
}

// p0 + (p1 - p0) a + (p2 - p0) b for a vertex-normal component
// (...Specializations.h:194-196), left to right; the Linux build contracts it
// as fma(p2 - p0, b, fma(p1 - p0, a, p0)).
template <int A>
RISE_HD inline double interp_normal(double p0, double p1, double p2, double a, double b)
{
    const double d1 = p1 - p0, d2 = p2 - p0;
    if (A == kArithX86Fma) return fma(d2, b, fma(d1, a, p0));
    return (p0 + d1 * a) + d2 * b;
}

// The same for a texture coordinate (...Specializations.h:214-216): the
// offset is summed first, p0 + ((p1 - p0) a + (p2 - p0) b); the Linux build
// contracts the sum as fma(p2 - p0, b, (p1 - p0) a).
template <int A>
RISE_HD inline double interp_uv(double p0, double p1, double p2, double a, double b)
{
    const double d1 = p1 - p0, d2 = p2 - p0;
    if (A == kArithX86Fma) return p0 + fma(d2, b, d1 * a);
    return p0 + (d1 * a + d2 * b);
}

// A committed hit in the mesh's frame, as RayElementIntersection and the
// mesh's IntersectRay leave it (...Specializations.h:191-216,
// TriangleMeshGeometryIndexed.cpp:170-188): the certifier's barycentrics; the
// shading normal, the vertex normals interpolated or else the face normal
// (v1 - v0) x (v2 - v0); the face normal normalized and turned to the shading
// normal's side; the texture coordinates interpolated (0 without them); on a
// double-sided mesh both normals turned toward the ray.
struct HitLocal {
    double ab[2];
    double gn[3];  // geometric normal (unit)
    double sn[3];  // shading normal (not normalized)
    double uv[2];
};

template <class R>
RISE_HD inline void hit_local(const R& r, const double v[9], unsigned tflags, const double* __restrict__ attr,
                              HitLocal& h)
{
    constexpr int A = R::kArith;
    double tt = 0.0;
    h.ab[0] = h.ab[1] = 0.0;
    (void)certify64(r, v, tt, h.ab);
    const double a = h.ab[0], b = h.ab[1];
    double fn[3];
    face_cross<A>(v[3] - v[0], v[4] - v[1], v[5] - v[2], v[6] - v[0], v[7] - v[1], v[8] - v[2], fn);
    if ((tflags & kTriNormals) && attr) {
        for (int c = 0; c < 3; ++c) h.sn[c] = interp_normal<A>(attr[c], attr[3 + c], attr[6 + c], a, b);
    } else {
        for (int c = 0; c < 3; ++c) h.sn[c] = fn[c];
    }
    normalize_model<A>(fn, h.gn);
    if (dot3<A>(h.gn[0], h.sn[0], h.gn[1], h.sn[1], h.gn[2], h.sn[2]) < 0.0)
        for (int c = 0; c < 3; ++c) h.gn[c] = -h.gn[c];
    if ((tflags & kTriCoords) && attr) {
        for (int c = 0; c < 2; ++c) h.uv[c] = interp_uv<A>(attr[9 + c], attr[11 + c], attr[13 + c], a, b);
    } else {
        h.uv[0] = h.uv[1] = 0.0;
    }
    if (tflags & kTriDoubleSided) {
        if (dot3<A>(h.sn[0], dn0(r), h.sn[1], dn1(r), h.sn[2], dn2(r)) > 0.0)
            for (int c = 0; c < 3; ++c) h.sn[c] = -h.sn[c];
        if (dot3<A>(h.gn[0], dn0(r), h.gn[1], dn1(r), h.gn[2], dn2(r)) > 0.0)
            for (int c = 0; c < 3; ++c) h.gn[c] = -h.gn[c];
    }
}

RISE_HD inline unsigned tri_flags(const Tri48* __restrict__ tri)
{
    return (unsigned)float_to_bits(reinterpret_cast<const float4*>(tri)[2].w);
}

template <int A>
RISE_HD inline void hit_attributes(const RayStateT<A>& r, const Tri48* __restrict__ tri, double t, double ab[2],
                                   double n_out[3], double p_out[3], const double* __restrict__ v64 = nullptr,
                                   const double* __restrict__ attr = nullptr, double* s_out = nullptr,
                                   double* uv_out = nullptr)
{
    double v[9];
    triangle_vertices<RayStateT<A>>(tri, v64, v);
    HitLocal h;
    hit_local(r, v, tri_flags(tri), attr, h);
    ab[0] = h.ab[0];
    ab[1] = h.ab[1];
    normalize_model<A>(h.gn, n_out);
    if (s_out) normalize_model<A>(h.sn, s_out);
    if (uv_out) {
        uv_out[0] = h.uv[0];
        uv_out[1] = h.uv[1];
    }
    const double s = t - kSurfaceBackoff;
    const double d[3] = {dn0(r), dn1(r), dn2(r)};
    const double o[3] = {r.ox, r.oy, r.oz};
    for (int k = 0; k < 3; ++k) p_out[k] = A == kArithX86Fma ? fma(d[k], s, o[k]) : o[k] + d[k] * s;
}

// Object.cpp:472-493: the world limit dHowFar becomes dHowFar2 = |d| * dHowFar
// in the normalized frame, unless it is the RISE_INFINITY sentinel and |d| >= 1.
template <int A>
RISE_HD inline double scale_how_far(const RayStateT<A>& r, double how_far)
{
    const double factor = (r.len_d > kNearZero) ? r.len_d : 1.0;
    if ((how_far != kRiseInfinity) || (factor < 1.0)) return factor * how_far;
    return how_far;
}

}  // namespace rise_hip
