// rise_binding.hpp -- the RISE-side binding: classes a RISE maintainer adds to
// RISE's own tree (they include RISE's headers, paths relative to
// src/Library/) so that a Job can hand its objects and its path-tracing
// renders to the MI355X library.
//
//   HipAcceleratedObjectManager   an Implementation::ObjectManager
//                                 (Managers/ObjectManager.h) that also keeps
//                                 its objects resident on the GPU and answers
//                                 batches of closest-hit / shadow queries there:
//                                 each object's mesh in its own frame with its
//                                 m_mxFinalTrans / m_mxInvFinalTrans, as
//                                 Object::IntersectRay intersects it.
//   HipPathTracingSpectralRasterizer
//                                 an IRasterizer (Interfaces/IRasterizer.h)
//                                 that renders the job's scene file on the GPU
//                                 and hands the image to the job's rasterizer
//                                 outputs as RISE's own rasterizers do.
//
// Compiled against RISE's headers by tests/test_rise_binding.py (syntax and
// types; linking needs RISE's library and librise_hip.so).  Errors follow
// RISE: bool returns and GlobalLog().
#pragma once

#include "Geometry/TriangleMeshGeometryIndexed.h"
#include "Interfaces/IEnumCallback.h"
#include "Interfaces/ILog.h"
#include "Interfaces/IObjectPriv.h"
#include "Interfaces/IRasterizer.h"
#include "Interfaces/IRasterizerOutput.h"
#include "Managers/ObjectManager.h"
#include "RasterImages/RasterImage.h"
#include "Utilities/Color/Color.h"
#include "Utilities/Reference.h"

#include "rise_hip_adapters.hpp"

#include <hip/hip_runtime_api.h>

#include <memory>
#include <string>
#include <vector>

namespace RISE {
namespace Implementation {

// ObjectManager whose PrepareForRendering also uploads every object's mesh
// (object space) and transforms to the GPU.  RISE's own per-ray
// IntersectRay / IntersectShadowRay stay the CPU ones; a wavefront integrator
// calls IntersectRays / IntersectShadowRays with whole batches in device
// memory, and ObjectOf maps a hit record's object index back to the object.
class HipAcceleratedObjectManager : public ObjectManager {
public:
    HipAcceleratedObjectManager(bool use_bvh, bool use_octree, unsigned max_objects_per_node, unsigned max_depth)
        : ObjectManager(use_bvh, use_octree, max_objects_per_node, max_depth)
    {
    }

    // ObjectManager::PrepareForRendering (Managers/ObjectManager.cpp:404-423)
    // plus the GPU upload.  Objects are taken in the manager's order (the
    // order its linear loop visits them); every visible object must carry a
    // TriangleMeshGeometryIndexed.
    void PrepareForRendering() const override
    {
        ObjectManager::PrepareForRendering();
        gpu_ok_ = Upload();
        if (!gpu_ok_) GlobalLog()->PrintEasyError("HipAcceleratedObjectManager: GPU upload failed");
    }

    // Closest hit for n rays (device SoA ox, oy, oz, dx, dy, dz) with the
    // face flags of IObjectManager::IntersectRay; hits: device, n records.
    bool IntersectRays(const float* rays_soa, int64_t n, bool front, bool back, rise_hip_hit* hits,
                       hipStream_t stream) const
    {
        return gpu_ok_ && gpu_->IntersectRay(rays_soa, n, front, back, hits, stream);
    }

    // IObjectManager::IntersectShadowRay for n rays; tmax (device) holds each
    // ray's dHowFar, NULL for RISE_INFINITY.
    bool IntersectShadowRays(const float* rays_soa, const float* tmax, int64_t n, bool front, bool back,
                             uint8_t* occluded, hipStream_t stream) const
    {
        return gpu_ok_ && gpu_->IntersectShadowRay(rays_soa, tmax, n, front, back, occluded, stream);
    }

    // The object of a hit record (rise_hip_hit::mesh), NULL for a miss.
    const IObjectPriv* ObjectOf(int32_t index) const
    {
        return (index >= 0 && (size_t)index < objects_.size()) ? objects_[(size_t)index] : nullptr;
    }

protected:
    ~HipAcceleratedObjectManager() override = default;

private:
    // Reads a mesh's protected bDoubleSided (TriangleMeshGeometryIndexed.h:72)
    // through a pointer to member taken in a derived class's scope.
    struct Sidedness : public TriangleMeshGeometryIndexed {
        static bool double_sided(const TriangleMeshGeometryIndexed& m) { return m.*(&Sidedness::bDoubleSided); }
    };

    // Collects the visible objects in the manager's order.
    struct Collect : public IEnumCallback<IObjectPriv> {
        std::vector<const IObjectPriv*>* out;
        bool operator()(const IObjectPriv& o) override
        {
            if (o.IsWorldVisible()) out->push_back(&o);
            return true;
        }
    };

    bool Upload() const
    {
        objects_.clear();
        vertices_.clear();
        Collect c;
        c.out = &objects_;
        EnumerateObjects(c);
        gpu_.reset(new rise_hip_host::HipObjectManager());
        vertices_.resize(objects_.size());
        normals_.assign(objects_.size(), {});
        coords_.assign(objects_.size(), {});
        for (size_t k = 0; k < objects_.size(); ++k) {
            const TriangleMeshGeometryIndexed* mesh =
                dynamic_cast<const TriangleMeshGeometryIndexed*>(objects_[k]->GetGeometry());
            if (!mesh) {
                GlobalLog()->PrintEasyError("HipAcceleratedObjectManager: only triangle meshes go to the GPU");
                return false;
            }
            // The mesh's triangles in its own frame, with the double
            // vertices the mesh holds (the library takes their float casts
            // for the filter, as the mesh's BVH does, BVH.h:431-467).
            // Its corners' vertex normals and texture coordinates go with
            // them when the mesh has them, for the hit records.
            std::vector<double>& v = vertices_[k];
            std::vector<double>& nrm = normals_[k];
            std::vector<double>& crd = coords_[k];
            bool has_n = true, has_c = true;
            for (const PointerTriangle& t : mesh->getFaces())
                for (int c3 = 0; c3 < 3; ++c3) {
                    has_n = has_n && t.pNormals[c3];
                    has_c = has_c && t.pCoords[c3];
                }
            for (const PointerTriangle& t : mesh->getFaces())
                for (int c3 = 0; c3 < 3; ++c3) {
                    v.push_back(t.pVertices[c3]->x);
                    v.push_back(t.pVertices[c3]->y);
                    v.push_back(t.pVertices[c3]->z);
                    if (has_n) {
                        nrm.push_back(t.pNormals[c3]->x);
                        nrm.push_back(t.pNormals[c3]->y);
                        nrm.push_back(t.pNormals[c3]->z);
                    }
                    if (has_c) {
                        crd.push_back(t.pCoords[c3]->x);
                        crd.push_back(t.pCoords[c3]->y);
                    }
                }
            const int m = gpu_->AddMeshF64(v.data(), (int64_t)(v.size() / 9), nullptr,
                                           Sidedness::double_sided(*mesh) ? RISE_HIP_MESH_DOUBLE_SIDED : 0u,
                                           has_n ? nrm.data() : nullptr, has_c ? crd.data() : nullptr);
            double tw[16], to[16];
            ToArray(objects_[k]->GetFinalTransformMatrix(), tw);
            ToArray(objects_[k]->GetFinalInverseTransformMatrix(), to);
            gpu_->AddObject(m, tw, to);
        }
        return gpu_->PrepareForRendering();
    }

    static void ToArray(const Matrix4& m, double out[16])
    {
        const double v[16] = {m._00, m._01, m._02, m._03, m._10, m._11, m._12, m._13,
                              m._20, m._21, m._22, m._23, m._30, m._31, m._32, m._33};
        for (int k = 0; k < 16; ++k) out[k] = v[k];
    }

    mutable std::unique_ptr<rise_hip_host::HipObjectManager> gpu_;
    mutable std::vector<const IObjectPriv*> objects_;
    mutable std::vector<std::vector<double>> vertices_;  // kept until the next upload
    mutable std::vector<std::vector<double>> normals_, coords_;
    mutable bool gpu_ok_ = false;
};

// The spectral path tracer as an IRasterizer: the job's scene file is read
// and uploaded once, each RasterizeScene renders it on the GPU and gives the
// image (XYZ mean per pixel, converted to RISEPel as ProgressiveFilm::Resolve
// does) to every rasterizer output, as PixelBasedRasterizerHelper does at the
// end of a render (Rendering/PixelBasedRasterizerHelper.cpp).
class HipPathTracingSpectralRasterizer : public virtual IRasterizer, public virtual Reference {
public:
    explicit HipPathTracingSpectralRasterizer(const char* scene_file) : path_(scene_file ? scene_file : "") {}

    void AddRasterizerOutput(IRasterizerOutput* ro) override
    {
        if (!ro) return;
        ro->addref();
        outputs_.push_back(ro);
    }

    void FreeRasterizerOutputs() override
    {
        for (IRasterizerOutput* ro : outputs_) ro->release();
        outputs_.clear();
    }

    void EnumerateRasterizerOutputs(IEnumCallback<IRasterizerOutput>& fn) const override
    {
        for (IRasterizerOutput* ro : outputs_) fn(*ro);
    }

    void AttachToScene(const IScene* scene) override { scene_ = scene; }
    void DetachFromScene(const IScene* scene) override
    {
        if (scene_ == scene) scene_ = nullptr;
    }

    unsigned int PredictTimeToRasterizeScene(const IScene&, const ISampling2D&, unsigned int* actual) const override
    {
        if (actual) *actual = 0;
        return 0;
    }

    void RasterizeScene(const IScene&, const Rect* rect, IRasterizeSequence*) const override
    {
        if (!loaded_) {
            loaded_ = gpu_.LoadAsciiScene(path_.c_str(), &desc_, nullptr);
            if (!loaded_) {
                GlobalLog()->PrintEasyError("HipPathTracingSpectralRasterizer: scene file not accepted");
                return;
            }
        }
        const size_t w = (size_t)desc_.width, h = (size_t)desc_.height;
        float* film = nullptr;
        if (hipMalloc((void**)&film, w * h * 3 * sizeof(float)) != hipSuccess) {
            GlobalLog()->PrintEasyError("HipPathTracingSpectralRasterizer: film allocation failed");
            return;
        }
        std::vector<float> xyz(w * h * 3);
        const bool ok = gpu_.RasterizeScene(desc_, film, nullptr) &&
                        hipMemcpy(xyz.data(), film, xyz.size() * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess;
        (void)hipFree(film);
        if (!ok) {
            GlobalLog()->PrintEasyError("HipPathTracingSpectralRasterizer: render failed");
            return;
        }
        RISERasterImage* image = new RISERasterImage((unsigned)w, (unsigned)h, RISEColor(0, 0, 0, 0));
        for (size_t y = 0; y < h; ++y)
            for (size_t x = 0; x < w; ++x) {
                const float* p = &xyz[3 * (y * w + x)];
                image->SetPEL((unsigned)x, (unsigned)y, RISEColor(RISEPel(XYZPel(p[0], p[1], p[2])), 1.0));
            }
        for (IRasterizerOutput* ro : outputs_) ro->OutputImage(*image, rect, 0);
        image->release();
    }

    void RasterizeSceneAnimation(const IScene&, const Scalar, const Scalar, const unsigned int, const bool, const bool,
                                 const Rect*, const unsigned int*, IRasterizeSequence*) const override
    {
        GlobalLog()->PrintEasyError("HipPathTracingSpectralRasterizer: animations are rendered frame by frame");
    }

    void SetProgressCallback(IProgressCallback* fn) override { progress_ = fn; }

protected:
    ~HipPathTracingSpectralRasterizer() override { FreeRasterizerOutputs(); }

private:
    std::string path_;
    std::vector<IRasterizerOutput*> outputs_;
    const IScene* scene_ = nullptr;
    IProgressCallback* progress_ = nullptr;
    mutable rise_hip_host::HipSpectralRasterizer gpu_;
    mutable rise_hip_render_desc desc_{};
    mutable bool loaded_ = false;
};

}  // namespace Implementation
}  // namespace RISE
