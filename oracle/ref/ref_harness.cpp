// ref_harness.cpp -- drives RISE's own mesh intersection path, compiled from
// the reference sources where they lie (see oracle/ref/Makefile), so the
// oracle and the HIP path can be checked against the genuine code.
// TEST INFRASTRUCTURE ONLY (checker and CPU baseline); never shipped.
//
// Path exercised (reference paths relative to src/Library/):
//   TriangleMeshGeometryIndexed::DoneIndexedTriangles (Geometry/TriangleMeshGeometryIndexed.cpp:444-540)
//     -> BVH<const PointerTriangle*> SAH BVH2 build, float filter, BVH4 collapse (Acceleration/BVH.h)
//   TriangleMeshGeometryIndexed::IntersectRay / _IntersectionOnly (TriangleMeshGeometryIndexed.cpp:151-205)
//     -> BVH::IntersectRay4 -> MollerTrumboreFloat -> RayElementIntersection -> RayTriangleIntersection
//
// The Object layer (Objects/Object.cpp) cannot be built from a few sources
// (it links the whole material library), so this harness performs
// Object::IntersectRay's steps around the mesh call with the reference's own
// Math3D operations: for an identity transform, normalize the direction
// (Object.cpp:414-425) and recompute the world range after the 1e-12
// back-off (Object.cpp:641-643), and for the shadow ray scale dHowFar
// (Object.cpp:472-493); for objects with transforms, every step of
// Object::IntersectRay / _IntersectionOnly in an ObjectManager's linear loop
// (rise_ref_objects_closest / _any).
//
// Triangle ids: triangle i gets its own texture coordinate (i, 0) on all three
// corners, so the reference's interpolated ri.ptCoord.x returns i exactly.

#include "Geometry/TriangleMeshGeometryIndexed.h"
#include "Intersection/RayIntersectionGeometric.h"
#include "Intersection/RayPrimitiveIntersections.h"
#include "Utilities/GeometricUtilities.h"
#include "Utilities/MemoryBuffer.h"
#include "RasterImages/HDRWriter.h"
#include "RasterImages/PPMWriter.h"
#include "RasterImages/TGAWriter.h"
#include "Rendering/DisplayTransformWriter.h"

#include <atomic>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

using namespace RISE;

namespace {

struct RefMesh {
    Implementation::TriangleMeshGeometryIndexed* mesh;
};

// Reads the protected BVH pointer, to tell a loaded cache from a rebuild.
struct MeshProbe : public Implementation::TriangleMeshGeometryIndexed {
    MeshProbe() : Implementation::TriangleMeshGeometryIndexed(false, true) {}
    unsigned bvh2_nodes() const { return pPtrBVH ? (unsigned)pPtrBVH->numNodes() : 0u; }
};

// Persistent worker pool: RISE renders with long-lived pool threads
// (Utilities/ThreadPool.h), so per-thread state (the BVH's thread_local
// traversal stack, the mesh's thread_local mailbox) is allocated once, not
// once per batch.
class Pool {
public:
    static Pool& get()
    {
        static Pool p;
        return p;
    }

    void run(int64_t n, int n_threads, const std::function<void(int64_t)>& fn)
    {
        if (n_threads <= 0) {
            n_threads = (int)std::thread::hardware_concurrency();
            if (n_threads <= 0) n_threads = 1;
        }
        std::unique_lock<std::mutex> lk(m_);
        while ((int)th_.size() < n_threads - 1) {
            const int idx = (int)th_.size();
            th_.emplace_back([this, idx] { loop(idx); });
        }
        fn_ = &fn;
        n_ = n;
        next_ = 0;
        want_ = n_threads - 1;
        done_ = 0;
        ++gen_;
        cv_.notify_all();
        lk.unlock();
        work();
        lk.lock();
        done_cv_.wait(lk, [this] { return done_ == want_; });
        fn_ = nullptr;
    }

    ~Pool()
    {
        {
            std::lock_guard<std::mutex> lk(m_);
            stop_ = true;
            ++gen_;
        }
        cv_.notify_all();
        for (auto& t : th_) t.join();
    }

private:
    void work()
    {
        for (;;) {
            const int64_t s = next_.fetch_add(256);
            if (s >= n_) break;
            const int64_t e = s + 256 < n_ ? s + 256 : n_;
            for (int64_t i = s; i < e; ++i) (*fn_)(i);
        }
    }

    void loop(int idx)
    {
        uint64_t seen = 0;
        for (;;) {
            std::unique_lock<std::mutex> lk(m_);
            cv_.wait(lk, [&] { return gen_ != seen; });
            seen = gen_;
            if (stop_) return;
            if (idx >= want_) continue;
            lk.unlock();
            work();
            lk.lock();
            ++done_;
            done_cv_.notify_all();
        }
    }

    std::vector<std::thread> th_;
    std::mutex m_;
    std::condition_variable cv_, done_cv_;
    const std::function<void(int64_t)>* fn_ = nullptr;
    int64_t n_ = 0;
    std::atomic<int64_t> next_{0};
    int want_ = 0;
    int done_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
};

template <class F>
void parallel_rays(int64_t n, int n_threads, F&& fn)
{
    const std::function<void(int64_t)> f(fn);
    Pool::get().run(n, n_threads, f);
}

}  // namespace

extern "C" {

// Builds the reference mesh (and its BVH) over n_tris triangles given as
// 9 floats each (v0, v1, v2).  Returns an opaque handle.
void* rise_ref_mesh_create(const float* tris, int64_t n_tris)
{
    RefMesh* h = new RefMesh;
    h->mesh = new Implementation::TriangleMeshGeometryIndexed(false, true);
    h->mesh->BeginIndexedTriangles();
    for (int64_t i = 0; i < n_tris; ++i) {
        const float* f = tris + 9 * i;
        for (int k = 0; k < 3; ++k)
            h->mesh->AddVertex(Vertex((double)f[3 * k + 0], (double)f[3 * k + 1], (double)f[3 * k + 2]));
        h->mesh->AddTexCoord(TexCoord((double)i, 0.0));
        IndexedTriangle t;
        for (int k = 0; k < 3; ++k) {
            t.iVertices[k] = (unsigned int)(3 * i + k);
            t.iNormals[k] = 0;
            t.iCoords[k] = (unsigned int)i;
        }
        h->mesh->AddIndexedTriangle(t);
    }
    h->mesh->DoneIndexedTriangles();
    return h;
}

// rise_ref_mesh_create / _create_bary on double vertices (9 doubles per
// triangle), kept as the mesh's Point3 vertices; double_sided: the mesh's
// bDoubleSided.
void* rise_ref_mesh_create_f64(const double* tris, int64_t n_tris, int bary, int double_sided)
{
    RefMesh* h = new RefMesh;
    h->mesh = new Implementation::TriangleMeshGeometryIndexed(double_sided != 0, !bary);
    h->mesh->BeginIndexedTriangles();
    if (bary) {
        h->mesh->AddNormal(Normal(1.0, 0.0, 0.0));
        h->mesh->AddNormal(Normal(0.0, 1.0, 0.0));
        h->mesh->AddNormal(Normal(0.0, 0.0, 1.0));
    }
    for (int64_t i = 0; i < n_tris; ++i) {
        const double* f = tris + 9 * i;
        for (int k = 0; k < 3; ++k) h->mesh->AddVertex(Vertex(f[3 * k + 0], f[3 * k + 1], f[3 * k + 2]));
        h->mesh->AddTexCoord(TexCoord((double)i, 0.0));
        IndexedTriangle t;
        for (int k = 0; k < 3; ++k) {
            t.iVertices[k] = (unsigned int)(3 * i + k);
            t.iNormals[k] = bary ? (unsigned int)k : 0u;
            t.iCoords[k] = (unsigned int)i;
        }
        h->mesh->AddIndexedTriangle(t);
    }
    h->mesh->DoneIndexedTriangles();
    return h;
}

// A mesh of double vertices with per-corner vertex normals (NULL: none, the
// face normal shades) and texture coordinates: corner k of triangle i gets
// the normal normals[9 i + 3 k ..] and the coordinate (i, vcoords[3 i + k]),
// so the interpolated ri.ptCoord.x is the id and .y the interpolated v.
void* rise_ref_mesh_create_attr(const double* tris, const double* normals, const double* vcoords, int64_t n_tris,
                                int double_sided)
{
    RefMesh* h = new RefMesh;
    h->mesh = new Implementation::TriangleMeshGeometryIndexed(double_sided != 0, normals == nullptr);
    h->mesh->BeginIndexedTriangles();
    for (int64_t i = 0; i < n_tris; ++i) {
        const double* f = tris + 9 * i;
        for (int k = 0; k < 3; ++k) {
            h->mesh->AddVertex(Vertex(f[3 * k + 0], f[3 * k + 1], f[3 * k + 2]));
            if (normals) {
                const double* nn = normals + 9 * i + 3 * k;
                h->mesh->AddNormal(Normal(nn[0], nn[1], nn[2]));
            }
            h->mesh->AddTexCoord(TexCoord((double)i, vcoords ? vcoords[3 * i + k] : 0.0));
        }
        IndexedTriangle t;
        for (int k = 0; k < 3; ++k) {
            t.iVertices[k] = (unsigned int)(3 * i + k);
            t.iNormals[k] = normals ? (unsigned int)(3 * i + k) : 0u;
            t.iCoords[k] = (unsigned int)(3 * i + k);
        }
        h->mesh->AddIndexedTriangle(t);
    }
    h->mesh->DoneIndexedTriangles();
    return h;
}

void rise_ref_mesh_release(void* handle)
{
    RefMesh* h = (RefMesh*)handle;
    if (!h) return;
    h->mesh->release();
    delete h;
}

// Closest hit; rays SoA (ox, oy, oz, dx, dy, dz), each n floats.
void rise_ref_closest(void* handle, const float* rays, int64_t n, int32_t* prim_id, float* t,
                      double* t64, int n_threads)
{
    RefMesh* h = (RefMesh*)handle;
    parallel_rays(n, n_threads, [&](int64_t i) {
        const Point3 o(rays[0 * n + i], rays[1 * n + i], rays[2 * n + i]);
        const Vector3 d(rays[3 * n + i], rays[4 * n + i], rays[5 * n + i]);
        const Vector3 dn = Vector3Ops::Normalize(d);
        const RasterizerState rs = {0, 0};
        RayIntersectionGeometric ri(Ray(o, dn), rs);
        h->mesh->IntersectRay(ri, true, true, false);
        if (ri.bHit) {
            const Point3 p = ri.ray.PointAtLength(ri.range - 1e-12);
            const Scalar range = Vector3Ops::Magnitude(Vector3Ops::mkVector3(p, o));
            prim_id[i] = (int32_t)ri.ptCoord.x;
            t[i] = (float)range;
            if (t64) t64[i] = range;
        } else {
            prim_id[i] = -1;
            t[i] = INFINITY;
            if (t64) t64[i] = INFINITY;
        }
    });
}

// rise_ref_mesh_create with the vertex normals (1,0,0), (0,1,0), (0,0,1) on
// the three corners of every triangle: the shading normal the reference
// interpolates at a hit, n0 + (n1 - n0) a + (n2 - n0) b
// (...Specializations.h:187-191), is then (1 - a - b, a, b), whose y and z are
// the certifier's barycentrics a and b exactly.  Hits and ranges are those of
// the plain mesh.
void* rise_ref_mesh_create_bary(const float* tris, int64_t n_tris)
{
    RefMesh* h = new RefMesh;
    h->mesh = new Implementation::TriangleMeshGeometryIndexed(false, false);  // vertex normals kept
    h->mesh->BeginIndexedTriangles();
    h->mesh->AddNormal(Normal(1.0, 0.0, 0.0));
    h->mesh->AddNormal(Normal(0.0, 1.0, 0.0));
    h->mesh->AddNormal(Normal(0.0, 0.0, 1.0));
    for (int64_t i = 0; i < n_tris; ++i) {
        const float* f = tris + 9 * i;
        for (int k = 0; k < 3; ++k)
            h->mesh->AddVertex(Vertex((double)f[3 * k + 0], (double)f[3 * k + 1], (double)f[3 * k + 2]));
        h->mesh->AddTexCoord(TexCoord((double)i, 0.0));
        IndexedTriangle t;
        for (int k = 0; k < 3; ++k) {
            t.iVertices[k] = (unsigned int)(3 * i + k);
            t.iNormals[k] = (unsigned int)k;
            t.iCoords[k] = (unsigned int)i;
        }
        h->mesh->AddIndexedTriangle(t);
    }
    h->mesh->DoneIndexedTriangles();
    return h;
}

// Closest hit with the face flags and the intersection record: id, world
// range, ri.vNormal.y / .z (the barycentrics on a mesh from
// rise_ref_mesh_create_bary), the geometric normal after Object::IntersectRay's
// Normalize (Object.cpp:531-533; identity transform), and the world hit point
// (Object.cpp:637-640).  normal and point: 3 doubles per ray.
void rise_ref_closest_ex(void* handle, const float* rays, int64_t n, int front, int back, int32_t* prim_id,
                         double* range, double* alpha, double* beta, double* normal, double* point, int n_threads)
{
    RefMesh* h = (RefMesh*)handle;
    parallel_rays(n, n_threads, [&](int64_t i) {
        const Point3 o(rays[0 * n + i], rays[1 * n + i], rays[2 * n + i]);
        const Vector3 d(rays[3 * n + i], rays[4 * n + i], rays[5 * n + i]);
        const Vector3 dn = Vector3Ops::Normalize(d);
        const RasterizerState rs = {0, 0};
        RayIntersectionGeometric ri(Ray(o, dn), rs);
        h->mesh->IntersectRay(ri, front != 0, back != 0, false);
        if (ri.bHit) {
            const Point3 p = ri.ray.PointAtLength(ri.range - 1e-12);
            prim_id[i] = (int32_t)ri.ptCoord.x;
            range[i] = Vector3Ops::Magnitude(Vector3Ops::mkVector3(p, o));
            alpha[i] = ri.vNormal.y;
            beta[i] = ri.vNormal.z;
            const Vector3 g = Vector3Ops::Normalize(ri.vGeomNormal);
            normal[3 * i] = g.x;
            normal[3 * i + 1] = g.y;
            normal[3 * i + 2] = g.z;
            point[3 * i] = p.x;
            point[3 * i + 1] = p.y;
            point[3 * i + 2] = p.z;
        } else {
            prim_id[i] = -1;
            range[i] = INFINITY;
            alpha[i] = beta[i] = 0.0;
            for (int k = 0; k < 3; ++k) normal[3 * i + k] = point[3 * i + k] = 0.0;
        }
    });
}

// Any hit within tmax[i] with the face flags.
void rise_ref_any_ex(void* handle, const float* rays, const float* tmax, int64_t n, int front, int back,
                     uint8_t* occluded, int n_threads)
{
    RefMesh* h = (RefMesh*)handle;
    parallel_rays(n, n_threads, [&](int64_t i) {
        const Point3 o(rays[0 * n + i], rays[1 * n + i], rays[2 * n + i]);
        const Vector3 d(rays[3 * n + i], rays[4 * n + i], rays[5 * n + i]);
        const Scalar mag = Vector3Ops::Magnitude(d);
        const Vector3 dn = Vector3Ops::Normalize(d);
        const Scalar factor = (mag > NEARZERO) ? mag : Scalar(1.0);
        const Scalar how_far = tmax ? (Scalar)tmax[i] : RISE_INFINITY;
        Scalar how_far2 = how_far;
        if ((how_far != RISE_INFINITY) || (factor < 1.0)) how_far2 = factor * how_far;
        occluded[i] = h->mesh->IntersectRay_IntersectionOnly(Ray(o, dn), how_far2, front != 0, back != 0) ? 1 : 0;
    });
}

// Any hit within tmax[i] (world distance; NULL = RISE_INFINITY).
void rise_ref_any(void* handle, const float* rays, const float* tmax, int64_t n, uint8_t* occluded,
                  int n_threads)
{
    RefMesh* h = (RefMesh*)handle;
    parallel_rays(n, n_threads, [&](int64_t i) {
        const Point3 o(rays[0 * n + i], rays[1 * n + i], rays[2 * n + i]);
        const Vector3 d(rays[3 * n + i], rays[4 * n + i], rays[5 * n + i]);
        const Scalar mag = Vector3Ops::Magnitude(d);
        const Vector3 dn = Vector3Ops::Normalize(d);
        const Scalar factor = (mag > NEARZERO) ? mag : Scalar(1.0);
        const Scalar how_far = tmax ? (Scalar)tmax[i] : RISE_INFINITY;
        Scalar how_far2 = how_far;
        if ((how_far != RISE_INFINITY) || (factor < 1.0)) how_far2 = factor * how_far;
        occluded[i] = h->mesh->IntersectRay_IntersectionOnly(Ray(o, dn), how_far2, true, true) ? 1 : 0;
    });
}

// A scene of objects with transforms: ObjectManager::IntersectRay's linear
// loop (Managers/ObjectManager.cpp:290-310) around Object::IntersectRay
// (Objects/Object.cpp:392-649), performed here step by step with the
// reference's Math3D operations around the genuine mesh call, since the
// Object class links the whole material library: frame change
// (Object.cpp:414-425), dHowFar scaling (:472-493), the box pre-test
// (:513-527), the normal transform (:531-533) and the world point and range
// (:637-643).  meshes: handles from rise_ref_mesh_create(_bary); per object
// k: obj_mesh[k] and 16 + 16 doubles of to_world / to_object (Matrix4
// order).  prim_id is the triangle's index within its mesh; alpha / beta are
// read from ri.vNormal (meshes from rise_ref_mesh_create_bary), else 0.
void rise_ref_objects_closest(void** meshes, const int32_t* obj_mesh, const double* to_world, const double* to_object,
                              int n_objs, const float* rays, int64_t n, int front, int back, int bary,
                              int32_t* prim_id, int32_t* object, double* range, double* alpha, double* beta,
                              double* normal, double* point, double* shading, double* vcoord, int n_threads)
{
    std::vector<Matrix4> M(n_objs), Mi(n_objs), MiT(n_objs);
    std::vector<BoundingBox> box(n_objs);
    for (int k = 0; k < n_objs; ++k) {
        const double* a = to_world + 16 * k;
        const double* b = to_object + 16 * k;
        M[k] = Matrix4(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9], a[10], a[11], a[12], a[13], a[14],
                       a[15]);
        Mi[k] = Matrix4(b[0], b[1], b[2], b[3], b[4], b[5], b[6], b[7], b[8], b[9], b[10], b[11], b[12], b[13], b[14],
                        b[15]);
        MiT[k] = Matrix4Ops::Transpose(Mi[k]);
        box[k] = ((RefMesh*)meshes[obj_mesh[k]])->mesh->GenerateBoundingBox();
    }
    parallel_rays(n, n_threads, [&](int64_t i) {
        const Point3 o(rays[0 * n + i], rays[1 * n + i], rays[2 * n + i]);
        const Vector3 d(rays[3 * n + i], rays[4 * n + i], rays[5 * n + i]);
        Scalar best = RISE_INFINITY;
        prim_id[i] = -1;
        object[i] = -1;
        range[i] = INFINITY;
        alpha[i] = beta[i] = 0.0;
        for (int c = 0; c < 3; ++c) normal[3 * i + c] = point[3 * i + c] = 0.0;
        if (shading)
            for (int c = 0; c < 3; ++c) shading[3 * i + c] = 0.0;
        if (vcoord) vcoord[i] = 0.0;
        for (int k = 0; k < n_objs; ++k) {
            const Point3 op = Point3Ops::Transform(Mi[k], o);
            const Vector3 dl = Vector3Ops::Transform(Mi[k], d);
            const Scalar mag = Vector3Ops::Magnitude(dl);
            const Ray r(op, Vector3Ops::Normalize(dl));
            const Scalar factor = (mag > NEARZERO) ? mag : Scalar(1.0);
            Scalar how_far2 = best;
            if ((best != RISE_INFINITY) || (factor < 1.0)) how_far2 = factor * best;
            BOX_HIT bh;
            RayBoxIntersection(r, bh, box[k].ll, box[k].ur);
            if (!bh.bHit) continue;
            if (!(bh.dRange2 < 0) && bh.dRange > how_far2) continue;
            const RasterizerState rs = {0, 0};
            RayIntersectionGeometric ri(r, rs);
            ((RefMesh*)meshes[obj_mesh[k]])->mesh->IntersectRay(ri, front != 0, back != 0, false);
            if (!ri.bHit) continue;
            const Point3 pobj = ri.ray.PointAtLength(ri.range - 1e-12);
            const Point3 pw = Point3Ops::Transform(M[k], pobj);
            const Scalar w = Vector3Ops::Magnitude(Vector3Ops::mkVector3(pw, o));
            if (!(w < best)) continue;
            best = w;
            prim_id[i] = (int32_t)ri.ptCoord.x;
            object[i] = k;
            range[i] = w;
            alpha[i] = bary ? ri.vNormal.y : 0.0;
            beta[i] = bary ? ri.vNormal.z : 0.0;
            const Vector3 g = Vector3Ops::Normalize(Vector3Ops::Transform(MiT[k], ri.vGeomNormal));
            normal[3 * i] = g.x;
            normal[3 * i + 1] = g.y;
            normal[3 * i + 2] = g.z;
            if (shading) {
                const Vector3 sn = Vector3Ops::Normalize(Vector3Ops::Transform(MiT[k], ri.vNormal));
                shading[3 * i] = sn.x;
                shading[3 * i + 1] = sn.y;
                shading[3 * i + 2] = sn.z;
            }
            if (vcoord) vcoord[i] = ri.ptCoord.y;
            point[3 * i] = pw.x;
            point[3 * i + 1] = pw.y;
            point[3 * i + 2] = pw.z;
        }
    });
}

// ObjectManager::IntersectShadowRay's loop (ObjectManager.cpp:320-373) around
// Object::IntersectRay_IntersectionOnly (Object.cpp:652-712), as above.
void rise_ref_objects_any(void** meshes, const int32_t* obj_mesh, const double* to_object, int n_objs,
                          const float* rays, const float* tmax, int64_t n, int front, int back, uint8_t* occluded,
                          int n_threads)
{
    std::vector<Matrix4> Mi(n_objs);
    std::vector<BoundingBox> box(n_objs);
    for (int k = 0; k < n_objs; ++k) {
        const double* b = to_object + 16 * k;
        Mi[k] = Matrix4(b[0], b[1], b[2], b[3], b[4], b[5], b[6], b[7], b[8], b[9], b[10], b[11], b[12], b[13], b[14],
                        b[15]);
        box[k] = ((RefMesh*)meshes[obj_mesh[k]])->mesh->GenerateBoundingBox();
    }
    parallel_rays(n, n_threads, [&](int64_t i) {
        const Point3 o(rays[0 * n + i], rays[1 * n + i], rays[2 * n + i]);
        const Vector3 d(rays[3 * n + i], rays[4 * n + i], rays[5 * n + i]);
        const Scalar how_far = tmax ? (Scalar)tmax[i] : RISE_INFINITY;
        uint8_t occ = 0;
        for (int k = 0; k < n_objs && !occ; ++k) {
            const Point3 op = Point3Ops::Transform(Mi[k], o);
            const Vector3 dl = Vector3Ops::Transform(Mi[k], d);
            const Scalar mag = Vector3Ops::Magnitude(dl);
            const Ray r(op, Vector3Ops::Normalize(dl));
            const Scalar factor = (mag > NEARZERO) ? mag : Scalar(1.0);
            Scalar how_far2 = how_far;
            if ((how_far != RISE_INFINITY) || (factor < 1.0)) how_far2 = factor * how_far;
            BOX_HIT bh;
            RayBoxIntersection(r, bh, box[k].ll, box[k].ur);
            if (!bh.bHit) continue;
            if (bh.dRange > how_far2 && !GeometricUtilities::IsPointInsideBox(r.origin, box[k].ll, box[k].ur))
                continue;
            if (((RefMesh*)meshes[obj_mesh[k]])->mesh->IntersectRay_IntersectionOnly(r, how_far2, front != 0,
                                                                                      back != 0))
                occ = 1;
        }
        occluded[i] = occ;
    });
}

// Writes the mesh with TriangleMeshGeometryIndexed::Serialize (.risemesh v5,
// with its BVH2 cache block).  Returns 0 on success.
int rise_ref_risemesh_save(void* handle, const char* path)
{
    RefMesh* h = (RefMesh*)handle;
    Implementation::MemoryBuffer* mb = new Implementation::MemoryBuffer();
    h->mesh->Serialize(*mb);
    const bool ok = mb->DumpToFileToCursor(path);
    mb->release();
    return ok ? 0 : 1;
}

// Loads a .risemesh file with TriangleMeshGeometryIndexed::Deserialize (which
// takes the BVH2 cache when it is valid).  Triangle ids in ray queries come
// from the texture coordinate's u, as for rise_ref_mesh_create.
void* rise_ref_risemesh_load(const char* path)
{
    RefMesh* h = new RefMesh;
    h->mesh = new MeshProbe();
    Implementation::MemoryBuffer* mb = new Implementation::MemoryBuffer(path);
    h->mesh->Deserialize(*mb);
    mb->release();
    return h;
}

// Nodes of the BVH2 of a mesh from rise_ref_risemesh_load (equal to the
// file's BVH2 block when Deserialize took the cache).
unsigned rise_ref_bvh2_nodes(void* handle)
{
    return static_cast<MeshProbe*>(((RefMesh*)handle)->mesh)->bvh2_nodes();
}

// Writes an XYZ film (h x w x 3 floats, row y = raster row y) with the
// reference's own colour pipeline and writers, as FileRasterizerOutput /
// FrameEncoderBase::Encode drive them: each pixel RISEColor(RISEPel(XYZPel),
// alpha 1); for LDR formats a DisplayTransformWriter when the transform is
// not `none` or the exposure not 0.  type: 0 TGA, 1 PPM, 3 HDR; color_space:
// 0 linear Rec.709, 1 sRGB.  Returns 0 on success.
int rise_ref_write_image(const float* xyz, int w, int h, int type, int color_space, double exposure_ev,
                         int display_transform, const char* path)
{
    Implementation::MemoryBuffer* mb = new Implementation::MemoryBuffer();
    const COLOR_SPACE cs = color_space == 0 ? eColorSpace_Rec709RGB_Linear : eColorSpace_sRGB;
    IRasterImageWriter* w0 = nullptr;
    if (type == 0) w0 = new Implementation::TGAWriter(*mb, cs);
    else if (type == 1) w0 = new Implementation::PPMWriter(*mb, cs);
    else if (type == 3) w0 = new Implementation::HDRWriter(*mb, cs);
    else {
        mb->release();
        return 2;
    }
    IRasterImageWriter* wr = w0;
    Implementation::DisplayTransformWriter* dtw = nullptr;
    if (type != 3 && (display_transform != 0 || exposure_ev != 0.0)) {
        dtw = new Implementation::DisplayTransformWriter(*w0, exposure_ev, (DISPLAY_TRANSFORM)display_transform);
        wr = dtw;
    }
    wr->BeginWrite(w, h);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const float* p = xyz + 3 * ((size_t)y * w + x);
            const XYZPel X(p[0], p[1], p[2]);
            wr->WriteColor(RISEColor(RISEPel(X), 1.0), x, y);
        }
    wr->EndWrite();
    const bool ok = mb->DumpToFileToCursor(path);
    if (dtw) dtw->release();
    w0->release();
    mb->release();
    return ok ? 0 : 1;
}

// RISEPel(XYZPel): ColorUtils::XYZtoRec709RGB with its gamut clip, n pixels.
void rise_ref_xyz_to_rec709(const float* xyz, int64_t n, double* rgb)
{
    for (int64_t i = 0; i < n; ++i) {
        const RISEPel p(XYZPel(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]));
        rgb[3 * i] = p.r;
        rgb[3 * i + 1] = p.g;
        rgb[3 * i + 2] = p.b;
    }
}

}  // extern "C"
