// rise_hip_adapters.hpp -- C++ host classes over the C ABI, shaped like the
// reference interfaces they stand in for, so RISE-side code can call the GPU
// path through familiar entry points.
//
//   HipObjectManager      ~ IObjectManager (Interfaces/IObjectManager.h:38-68):
//                           PrepareForRendering, IntersectRay, IntersectShadowRay,
//                           here over batches of rays in device memory.
//   HipSpectralRasterizer ~ IRasterizer::RasterizeScene (Interfaces/IRasterizer.h:70)
//                           for PathTracingSpectralRasterizer
//                           (Rendering/PathTracingSpectralRasterizer.h:48-58).
//
// Error convention as in RISE: methods return bool and log the reason
// (GlobalLog()->PrintEasyError in the reference; a caller-settable sink here);
// no exception leaves these classes.  Header-only; link librise_hip.so.
#pragma once

#include "rise_hip.h"
#include "rise_hip_render.h"
#include "rise_hip_risemesh.h"
#include "rise_hip_scene_file.h"
#include "rise_hip_output.h"

#include <cstdint>
#include <cstdio>
#include <functional>
#include <string>
#include <utility>
#include <vector>

namespace rise_hip_host {

using LogSink = std::function<void(const std::string&)>;

inline LogSink& log_sink()
{
    static LogSink sink = [](const std::string& m) { std::fprintf(stderr, "rise_hip: %s\n", m.c_str()); };
    return sink;
}

inline bool check(rise_hip_status s, const char* what)
{
    if (s == RISE_HIP_OK) return true;
    log_sink()(std::string(what) + " failed: " + rise_hip_last_error());
    return false;
}

// Batch ray queries over a resident triangle scene.
class HipObjectManager {
public:
    HipObjectManager() = default;
    HipObjectManager(const HipObjectManager&) = delete;
    HipObjectManager& operator=(const HipObjectManager&) = delete;
    ~HipObjectManager() { rise_hip_scene_release(scene_); }

    // Adds a mesh (n_tris * 9 floats); copied at PrepareForRendering.  Without
    // objects it is world-space geometry; with AddObject it is placed by the
    // objects that refer to it.  Returns its index.
    int AddMesh(const float* vertices, int64_t n_tris, const int32_t* tri_ids = nullptr, uint32_t flags = 0)
    {
        meshes_.push_back({vertices, n_tris, tri_ids, nullptr, flags, nullptr, nullptr});
        return (int)meshes_.size() - 1;
    }

    // The same with the mesh's double vertices (n_tris * 9), as RISE holds
    // them: the certifier uses them, the float filter their float casts.
    // normals (n_tris * 9) and coords (n_tris * 6), optional: per-corner
    // vertex normals and texture coordinates for the hit records.
    int AddMeshF64(const double* vertices, int64_t n_tris, const int32_t* tri_ids = nullptr, uint32_t flags = 0,
                   const double* normals = nullptr, const double* coords = nullptr)
    {
        meshes_.push_back({nullptr, n_tris, tri_ids, vertices, flags, normals, coords});
        return (int)meshes_.size() - 1;
    }

    // Adds an object: mesh `mesh` placed by the object's Matrix4s
    // (m_mxFinalTrans, m_mxInvFinalTrans; 16 doubles each in member order).
    // Returns its index, which hit records report in rise_hip_hit::mesh.
    int AddObject(int mesh, const double to_world[16], const double to_object[16])
    {
        rise_hip_object_instance o{};
        o.mesh = mesh;
        for (int k = 0; k < 16; ++k) {
            o.to_world[k] = to_world[k];
            o.to_object[k] = to_object[k];
        }
        objects_.push_back(o);
        return (int)objects_.size() - 1;
    }

    // ObjectManager::PrepareForRendering (Managers/ObjectManager.cpp:404-423):
    // builds the acceleration structures and uploads them.
    bool PrepareForRendering()
    {
        rise_hip_scene_release(scene_);
        scene_ = nullptr;
        if (!objects_.empty())
            return check(rise_hip_scene_upload_objects(meshes_.data(), meshes_.size(), objects_.data(),
                                                       objects_.size(), &scene_),
                         "PrepareForRendering");
        return check(rise_hip_scene_upload(meshes_.data(), meshes_.size(), &scene_), "PrepareForRendering");
    }

    // IntersectRay for a batch: rays_soa (device, 6*n floats), outputs device.
    bool IntersectRay(const float* rays_soa, int64_t n, int32_t* prim_id, float* t, hipStream_t stream) const
    {
        return check(rise_hip_intersect_closest(scene_, rays_soa, n, prim_id, t, stream), "IntersectRay");
    }

    // IntersectShadowRay for a batch; tmax (device) may be NULL.
    bool IntersectShadowRay(const float* rays_soa, const float* tmax, int64_t n, uint8_t* occluded,
                            hipStream_t stream) const
    {
        return check(rise_hip_intersect_any(scene_, rays_soa, tmax, n, occluded, stream), "IntersectShadowRay");
    }

    // IObjectManager::IntersectRay(ri, bHitFrontFaces, bHitBackFaces, ...)
    // (Interfaces/IObjectManager.h:37-43) for a batch: the face flags select
    // the triangles, hits (device, n records) receive id, mesh, world range,
    // barycentrics, hit point and geometric normal.
    bool IntersectRay(const float* rays_soa, int64_t n, bool hit_front_faces, bool hit_back_faces,
                      rise_hip_hit* hits, hipStream_t stream) const
    {
        const int32_t faces = (hit_front_faces ? RISE_HIP_FACES_FRONT : 0) | (hit_back_faces ? RISE_HIP_FACES_BACK : 0);
        return check(rise_hip_intersect_closest_ex(scene_, rays_soa, n, faces, hits, stream), "IntersectRay");
    }

    // IObjectManager::IntersectShadowRay(ray, dHowFar, bHitFrontFaces,
    // bHitBackFaces) (IObjectManager.h:45-50) for a batch.
    bool IntersectShadowRay(const float* rays_soa, const float* tmax, int64_t n, bool hit_front_faces,
                            bool hit_back_faces, uint8_t* occluded, hipStream_t stream) const
    {
        const int32_t faces = (hit_front_faces ? RISE_HIP_FACES_FRONT : 0) | (hit_back_faces ? RISE_HIP_FACES_BACK : 0);
        return check(rise_hip_intersect_any_ex(scene_, rays_soa, tmax, n, faces, occluded, stream),
                     "IntersectShadowRay");
    }

    // The built 8-wide tree saved to / reloaded from a scene cache file
    // (rise_hip_scene_cache_*): a reload replaces PrepareForRendering with no build.
    bool SaveCache(const char* path) const { return check(rise_hip_scene_cache_write(scene_, path), "SaveCache"); }
    bool LoadCache(const char* path)
    {
        rise_hip_scene_release(scene_);
        scene_ = nullptr;
        return check(rise_hip_scene_cache_read(path, &scene_), "LoadCache");
    }

    bool GetStats(rise_hip_scene_stats& out) const
    {
        return check(rise_hip_scene_get_stats(scene_, &out), "GetStats");
    }

private:
    std::vector<rise_hip_mesh_desc> meshes_;
    std::vector<rise_hip_object_instance> objects_;
    rise_hip_scene* scene_ = nullptr;
};

// The spectral path-tracing rasterizer for one resident scene.
class HipSpectralRasterizer {
public:
    HipSpectralRasterizer() = default;
    HipSpectralRasterizer(const HipSpectralRasterizer&) = delete;
    HipSpectralRasterizer& operator=(const HipSpectralRasterizer&) = delete;
    ~HipSpectralRasterizer() { rise_hip_pt_scene_release(scene_); }

    int AddSpectrum(float begin, float end, const std::vector<float>& amplitudes)
    {
        amps_.push_back(amplitudes);
        spds_.push_back({begin, end, (int32_t)amplitudes.size(), nullptr});
        return (int)spds_.size() - 1;
    }

    int AddLambertianMaterial(int reflectance_spd, int emission_spd = -1, float emission_scale = 1.0f)
    {
        rise_hip_material_desc m{};
        m.type = RISE_HIP_MATERIAL_LAMBERTIAN;
        m.reflectance_spd = reflectance_spd;
        m.emission_spd = emission_spd;
        m.emission_scale = emission_scale;
        materials_.push_back(m);
        return (int)materials_.size() - 1;
    }

    // GGXMaterial with conductor Fresnel (ggx_material: rd, rs, alphax = alphay, ior, extinction).
    int AddGGXMaterial(int diffuse_spd, int specular_spd, float alpha, int ior_spd, int extinction_spd)
    {
        rise_hip_material_desc m{};
        m.type = RISE_HIP_MATERIAL_GGX;
        m.reflectance_spd = diffuse_spd;
        m.emission_spd = -1;
        m.specular_spd = specular_spd;
        m.ior_spd = ior_spd;
        m.extinction_spd = extinction_spd;
        m.alpha = alpha;
        materials_.push_back(m);
        return (int)materials_.size() - 1;
    }

    // DielectricMaterial, smooth (dielectric_material: ior, tau).
    int AddDielectricMaterial(int ior_spd, int tau_spd = -1)
    {
        rise_hip_material_desc m{};
        m.type = RISE_HIP_MATERIAL_DIELECTRIC;
        m.reflectance_spd = -1;
        m.emission_spd = -1;
        m.specular_spd = -1;
        m.ior_spd = ior_spd;
        m.extinction_spd = tau_spd;
        materials_.push_back(m);
        return (int)materials_.size() - 1;
    }

    // Job::AddHosekWilkieSkylight (Job.cpp:2717-2750): the sky map and, when
    // sky.create_sun is set, its matched sun light of power
    // sky.sun_intensity_scale; applied at PrepareForRendering.
    void SetSky(const rise_hip_sky_desc& sky)
    {
        sky_ = sky;
        has_sky_ = true;
    }

    // Job::AddDirectionalLight (Job.cpp:5424): light arriving from `direction`
    // (toward the light) with RGB colour times power; at most 4 with the sky's
    // sun.  Applied at PrepareForRendering.
    void AddDirectionalLight(double power, const double color[3], const double direction[3])
    {
        rise_hip_directional_light_desc d{};
        for (int c = 0; c < 3; ++c) {
            d.color[c] = (float)color[c];
            d.direction[c] = (float)direction[c];
        }
        d.power = (float)power;
        dir_lights_.push_back(d);
    }

    // Job::AddHomogeneousMedium with `phase hg g`: RGB absorption and scattering per unit length.
    int AddHomogeneousMedium(const float sigma_a[3], const float sigma_s[3], float hg_g)
    {
        rise_hip_medium_desc m{};
        for (int c = 0; c < 3; ++c) {
            m.sigma_a[c] = sigma_a[c];
            m.sigma_s[c] = sigma_s[c];
        }
        m.hg_g = hg_g;
        media_.push_back(m);
        return (int)media_.size() - 1;
    }

    // standard_object: a triangle mesh with its material, interior_medium (-1: none)
    // and casts_shadows.
    void AddObject(const float* vertices, int64_t n_tris, int material, bool double_sided = false,
                   int interior_medium = -1, bool casts_shadows = true)
    {
        objects_.push_back({vertices, n_tris, material, double_sided ? 1 : 0});
        object_media_.push_back({interior_medium, casts_shadows ? 1 : 0});
        if (interior_medium >= 0 || !casts_shadows) use_media_ = true;
    }

    bool PrepareForRendering()
    {
        for (size_t i = 0; i < spds_.size(); ++i) spds_[i].amplitudes = amps_[i].data();
        rise_hip_pt_scene_release(scene_);
        scene_ = nullptr;
        if (!check(rise_hip_pt_scene_create_with_media(objects_.data(), objects_.size(), materials_.data(),
                                                       materials_.size(), spds_.data(), spds_.size(), media_.data(),
                                                       media_.size(), use_media_ ? object_media_.data() : nullptr,
                                                       &scene_),
                   "PrepareForRendering"))
            return false;
        if (!dir_lights_.empty() &&
            !check(rise_hip_pt_scene_set_directional_lights(scene_, dir_lights_.data(), dir_lights_.size()),
                   "PrepareForRendering (directional lights)"))
            return false;
        return !has_sky_ || check(rise_hip_pt_scene_set_sky(scene_, &sky_), "PrepareForRendering (sky)");
    }

    // Job::LoadAsciiSceneViaCst (Job.h:2843): reads a native scene file and
    // makes its scene resident; *desc receives the render its camera, film
    // and rasterizer chunks describe, `outputs` (optional) its
    // file_rasterizeroutput chunks as (description, file name) pairs.
    // Replaces anything added before.
    bool LoadAsciiScene(const char* path, rise_hip_render_desc* desc,
                        std::vector<std::pair<rise_hip_output_desc, std::string>>* outputs = nullptr)
    {
        rise_hip_scene_file* f = nullptr;
        if (!check(rise_hip_scene_file_parse(path, &f), "LoadAsciiScene")) return false;
        rise_hip_pt_scene_release(scene_);
        scene_ = nullptr;
        bool ok = check(rise_hip_scene_file_upload(f, &scene_), "LoadAsciiScene (upload)");
        if (ok && desc) ok = check(rise_hip_scene_file_get_render_desc(f, desc), "LoadAsciiScene (render)");
        rise_hip_scene_file_info info{};
        if (ok && outputs) ok = check(rise_hip_scene_file_get_info(f, &info), "LoadAsciiScene (info)");
        for (int32_t i = 0; ok && outputs && i < info.n_outputs; ++i) {
            rise_hip_output_desc od{};
            const char* pattern = nullptr;
            ok = check(rise_hip_scene_file_get_output(f, i, &od, &pattern), "LoadAsciiScene (output)");
            if (ok) outputs->push_back({od, std::string(pattern) + "." + rise_hip_output_extension(od.format)});
        }
        rise_hip_scene_file_release(f);
        return ok;
    }

    // FileRasterizerOutput::OutputImage: writes a rendered film (device,
    // W*H*3 mean XYZ) to `path` as `od` describes.
    static bool OutputImage(const float* film_xyz, int width, int height, const rise_hip_output_desc& od,
                            const char* path, hipStream_t stream)
    {
        return check(rise_hip_output_write(film_xyz, width, height, &od, path, stream), "OutputImage");
    }

    rise_hip_pt_scene* Scene() const { return scene_; }

    // RasterizeScene: renders desc into film_xyz (device, W*H*3 mean XYZ).
    bool RasterizeScene(const rise_hip_render_desc& desc, float* film_xyz, hipStream_t stream,
                        rise_hip_render_stats* stats = nullptr) const
    {
        return check(rise_hip_render(scene_, &desc, film_xyz, nullptr, stats, stream), "RasterizeScene");
    }

private:
    std::vector<std::vector<float>> amps_;
    std::vector<rise_hip_spd_desc> spds_;
    std::vector<rise_hip_material_desc> materials_;
    std::vector<rise_hip_object_desc> objects_;
    std::vector<rise_hip_medium_desc> media_;
    std::vector<rise_hip_object_media> object_media_;
    bool use_media_ = false;
    rise_hip_sky_desc sky_{};
    bool has_sky_ = false;
    std::vector<rise_hip_directional_light_desc> dir_lights_;
    rise_hip_pt_scene* scene_ = nullptr;
};

// A continuous-batching render server over resident scenes: a job runner
// submits each incoming render and polls for finished ones.
class HipRenderServer {
public:
    explicit HipRenderServer(hipStream_t stream, int64_t hbm_budget_bytes = 0)
    {
        rise_hip_server_config cfg{};
        cfg.hbm_budget_bytes = hbm_budget_bytes;
        ok_ = check(rise_hip_server_create(&cfg, stream, &server_), "HipRenderServer");
    }
    HipRenderServer(const HipRenderServer&) = delete;
    HipRenderServer& operator=(const HipRenderServer&) = delete;
    ~HipRenderServer() { rise_hip_server_release(server_); }

    bool Ok() const { return ok_; }

    bool Submit(const HipSpectralRasterizer& scene, const rise_hip_render_desc& desc, float* film_xyz,
                uint64_t& ticket)
    {
        rise_hip_render_request q{};
        q.scene = scene.Scene();
        q.desc = desc;
        q.film_xyz = film_xyz;
        return check(rise_hip_server_submit(server_, &q, &ticket), "Submit");
    }

    // Finished tickets since the last call (non-blocking, or all when wait is set).
    bool Finished(std::vector<uint64_t>& tickets, bool wait)
    {
        uint64_t t[64];
        int64_t lat[64];
        size_t n = 0;
        do {
            if (!check(wait ? rise_hip_server_drain(server_, t, lat, 64, &n) : rise_hip_server_poll(server_, t, lat, 64, &n),
                       "Finished"))
                return false;
            tickets.insert(tickets.end(), t, t + n);
        } while (n == 64);
        return true;
    }

private:
    rise_hip_server* server_ = nullptr;
    bool ok_ = false;
};

}  // namespace rise_hip_host
