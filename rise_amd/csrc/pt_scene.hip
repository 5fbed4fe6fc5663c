// pt_scene.hip -- spectral path-tracing scenes: BVH, material, spectrum and
// luminary tables resident in HBM (rise_hip_pt_scene_create / _release).
#include "rise_hip_render.h"

#include "api_common.h"
#include "bvh_build.h"
#include "env_sampler.h"
#include "pt_internal.h"
#include "rise_hip_rgb_uplift.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

using namespace rise_hip;

// ---------------------------------------------------------------------------
// Scene
// ---------------------------------------------------------------------------

namespace {

template <class T>
rise_hip_status upload(rise_hip_pt_scene* s, const std::vector<T>& v, const T** out)
{
    const size_t bytes = std::max<size_t>(1, v.size()) * sizeof(T);
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return fail_hip(e, "hipMalloc (scene)");
    s->allocs.push_back(p);
    if (!v.empty()) {
        e = hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
        if (e != hipSuccess) return fail_hip(e, "hipMemcpy (scene)");
    }
    *out = static_cast<const T*>(p);
    return RISE_HIP_OK;
}

void free_scene(rise_hip_pt_scene* s)
{
    release_solo_batch(s);
    for (void* p : s->allocs) (void)hipFree(p);
    delete s;
}

}  // namespace

extern "C" rise_hip_status rise_hip_pt_scene_create(const rise_hip_object_desc* objects, size_t n_objects,
                                                    const rise_hip_material_desc* materials, size_t n_materials,
                                                    const rise_hip_spd_desc* spds, size_t n_spds,
                                                    rise_hip_pt_scene** out_scene)
{
    return rise_hip_pt_scene_create_with_media(objects, n_objects, materials, n_materials, spds, n_spds, nullptr, 0,
                                               nullptr, out_scene);
}

extern "C" rise_hip_status rise_hip_pt_scene_create_with_media(
    const rise_hip_object_desc* objects, size_t n_objects, const rise_hip_material_desc* materials,
    size_t n_materials, const rise_hip_spd_desc* spds, size_t n_spds, const rise_hip_medium_desc* media,
    size_t n_media, const rise_hip_object_media* object_media, rise_hip_pt_scene** out_scene)
{
    return pt_scene_create(objects, n_objects, materials, n_materials, spds, n_spds, media, n_media, object_media,
                           nullptr, out_scene);
}

rise_hip_status rise_hip::pt_scene_create(const rise_hip_object_desc* objects, size_t n_objects,
                                          const rise_hip_material_desc* materials, size_t n_materials,
                                          const rise_hip_spd_desc* spds, size_t n_spds,
                                          const rise_hip_medium_desc* media, size_t n_media,
                                          const rise_hip_object_media* object_media,
                                          const Bvh2Tree* const* object_trees, rise_hip_pt_scene** out_scene)
{
    clear_error();
    if (!out_scene) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "out_scene is NULL");
    *out_scene = nullptr;
    if ((n_objects && !objects) || (n_materials && !materials) || (n_spds && !spds) || (n_media && !media))
        return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL array with a nonzero count");
    if (n_objects > 0x7fffffff || n_materials > 0x7fffffff || n_spds > 0x7fffffff || n_media > 0x7fffffff)
        return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "too many objects, materials, spectra or media");
    // Media: luminances of the RGB coefficients (HomogeneousMedium::GetCoefficientsNM).
    std::vector<DevMedium> dmed(n_media);
    for (size_t i = 0; i < n_media; ++i) {
        const rise_hip_medium_desc& m = media[i];
        const std::string who = "medium " + std::to_string(i) + ": ";
        for (int c = 0; c < 3; ++c)
            if (!(m.sigma_a[c] >= 0.0f && m.sigma_s[c] >= 0.0f && std::isfinite(m.sigma_a[c]) &&
                  std::isfinite(m.sigma_s[c])))
                return fail(RISE_HIP_ERR_INVALID_ARGUMENT, who + "coefficients must be finite and non-negative");
        if (!(m.hg_g > -1.0f && m.hg_g < 1.0f))
            return fail(RISE_HIP_ERR_INVALID_ARGUMENT, who + "Henyey-Greenstein g must lie in (-1, 1)");
        if (m.reserved != 0) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, who + "reserved field must be 0");
        double t[3];
        for (int c = 0; c < 3; ++c) t[c] = (double)m.sigma_a[c] + (double)m.sigma_s[c];
        dmed[i].sigma_t = (float)(0.2126 * t[0] + 0.7152 * t[1] + 0.0722 * t[2]);
        dmed[i].sigma_s = (float)(0.2126 * (double)m.sigma_s[0] + 0.7152 * (double)m.sigma_s[1] +
                                  0.0722 * (double)m.sigma_s[2]);
        dmed[i].g = m.hg_g;
        dmed[i].pad = 0;
    }
    if (object_media)
        for (size_t o = 0; o < n_objects; ++o)
            if (object_media[o].interior_medium < -1 || object_media[o].interior_medium >= (int64_t)n_media)
                return fail(RISE_HIP_ERR_INVALID_ARGUMENT,
                            "object " + std::to_string(o) + ": interior medium out of range");
    // Spectra.
    std::vector<DevSpd> dspd(n_spds);
    std::vector<float> amps;
    for (size_t i = 0; i < n_spds; ++i) {
        const rise_hip_spd_desc& d = spds[i];
        if (d.n == RISE_HIP_SPD_SIGMOID) {
            if (!d.amplitudes)
                return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "spectrum " + std::to_string(i) + ": NULL coefficients");
            for (int k = 0; k < 4; ++k)
                if (!std::isfinite(d.amplitudes[k]))
                    return fail(RISE_HIP_ERR_INVALID_ARGUMENT,
                                "spectrum " + std::to_string(i) + ": non-finite sigmoid coefficient or scale");
            dspd[i].begin = 380.0f;
            dspd[i].end = 780.0f;
            dspd[i].n = RISE_HIP_SPD_SIGMOID;
            dspd[i].inv_delta = 0.0f;
            dspd[i].offset = (int)amps.size();
            amps.insert(amps.end(), d.amplitudes, d.amplitudes + 4);
            continue;
        }
        if (d.n <= 0 || !d.amplitudes || !(d.lambda_end > d.lambda_begin))
            return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "spectrum " + std::to_string(i) + " is empty or inverted");
        dspd[i].begin = d.lambda_begin;
        dspd[i].end = d.lambda_end;
        dspd[i].n = d.n;
        dspd[i].inv_delta = (float)((double)d.n / ((double)d.lambda_end - (double)d.lambda_begin));
        dspd[i].offset = (int)amps.size();
        amps.insert(amps.end(), d.amplitudes, d.amplitudes + d.n);
    }
    // Materials.
    std::vector<DevMaterial> dmat(n_materials);
    bool has_glossy = false, has_transmission = false;
    unsigned material_kinds = 0;
    for (size_t i = 0; i < n_materials; ++i) {
        const rise_hip_material_desc& m = materials[i];
        const std::string who = "material " + std::to_string(i) + ": ";
        if (m.type != RISE_HIP_MATERIAL_LAMBERTIAN && m.type != RISE_HIP_MATERIAL_GGX &&
            m.type != RISE_HIP_MATERIAL_DIELECTRIC)
            return fail(RISE_HIP_ERR_NOT_SUPPORTED, who + "unsupported type");
        auto bad = [&](int32_t k, bool required) { return k < -1 || k >= (int)n_spds || (required && k < 0); };
        if (bad(m.reflectance_spd, m.type == RISE_HIP_MATERIAL_LAMBERTIAN) || bad(m.emission_spd, false) ||
            bad(m.specular_spd, m.type == RISE_HIP_MATERIAL_GGX) || bad(m.ior_spd, false) ||
            bad(m.extinction_spd, false))
            return fail(RISE_HIP_ERR_INVALID_ARGUMENT, who + "spectrum index out of range");
        if (m.type != RISE_HIP_MATERIAL_LAMBERTIAN && m.emission_spd >= 0)
            return fail(RISE_HIP_ERR_NOT_SUPPORTED, who + "only Lambertian materials emit");
        if (m.type == RISE_HIP_MATERIAL_GGX && !(m.alpha >= 0.0f))
            return fail(RISE_HIP_ERR_INVALID_ARGUMENT, who + "negative or NaN roughness");
        dmat[i] = DevMaterial{m.type,        m.reflectance_spd, m.emission_spd,    m.emission_scale,
                              m.specular_spd, m.ior_spd,         m.extinction_spd, m.alpha};
        material_kinds |= 1u << m.type;
        if (m.type == RISE_HIP_MATERIAL_GGX) has_glossy = true;
        if (m.type == RISE_HIP_MATERIAL_DIELECTRIC) has_glossy = has_transmission = true;
    }
    // Objects -> one triangle list.
    int64_t total = 0;
    for (size_t o = 0; o < n_objects; ++o) {
        if (objects[o].n_tris < 0 || (objects[o].n_tris > 0 && !objects[o].vertices))
            return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "object " + std::to_string(o) + ": bad triangle buffer");
        if (objects[o].material < 0 || (size_t)objects[o].material >= n_materials)
            return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "object " + std::to_string(o) + ": material out of range");
        total += objects[o].n_tris;
    }
    if (total > 0x7fffffff) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "more than 2^31-1 triangles");
    std::vector<float> verts((size_t)total * 9);
    std::vector<int32_t> tri_obj((size_t)total);
    {
        int64_t off = 0;
        for (size_t o = 0; o < n_objects; ++o) {
            const int64_t k = objects[o].n_tris;
            if (k) std::memcpy(verts.data() + 9 * off, objects[o].vertices, (size_t)k * 9 * sizeof(float));
            for (int64_t i = 0; i < k; ++i) tri_obj[off + i] = (int32_t)o;
            off += k;
        }
    }
    WideBVH bvh;
    std::string err;
    bool any_tree = false;
    for (size_t o = 0; object_trees && o < n_objects; ++o) any_tree = any_tree || object_trees[o];
    try {
        if (any_tree) {
            std::vector<Bvh2Part> parts;
            int64_t off = 0;
            for (size_t o = 0; o < n_objects; ++o) {
                parts.push_back({off, objects[o].n_tris, object_trees[o]});
                off += objects[o].n_tris;
            }
            if (!build_wide_bvh_parts(verts.data(), total, nullptr, parts, BuildParams(), bvh, err))
                return fail(RISE_HIP_ERR_BUILD, "BVH build failed: " + err);
        } else if (!build_wide_bvh(verts.data(), total, nullptr, BuildParams(), bvh, err)) {
            return fail(RISE_HIP_ERR_BUILD, "BVH build failed: " + err);
        }
    } catch (const std::bad_alloc&) {
        return fail(RISE_HIP_ERR_OUT_OF_MEMORY, "host allocation failed during BVH build");
    }
    // Object id in v1.w, shadow flag in v2.w; inverse map global id -> leaf index.
    std::vector<uint32_t> leaf_of((size_t)total);
    for (size_t l = 0; l < bvh.tris.size(); ++l) {
        int32_t id;
        std::memcpy(&id, &bvh.tris[l].v0[3], 4);
        const int32_t ob = tri_obj[(size_t)id];
        std::memcpy(&bvh.tris[l].v1[3], &ob, 4);
        const uint32_t flags = (object_media && !object_media[ob].casts_shadows) ? kTriNoShadow : 0u;
        std::memcpy(&bvh.tris[l].v2[3], &flags, 4);
        leaf_of[(size_t)id] = (uint32_t)l;
    }
    // The shells of the objects holding a medium, for the shadow-ray
    // transmittance walk (LightSampler.cpp:369-470).
    WideBVH mbvh;
    bool has_media = false;
    if (object_media) {
        std::vector<float> mverts;
        std::vector<int32_t> mobj;
        int64_t off = 0;
        for (size_t o = 0; o < n_objects; ++o) {
            const int64_t k = objects[o].n_tris;
            if (object_media[o].interior_medium >= 0 && k > 0) {
                mverts.insert(mverts.end(), verts.begin() + 9 * off, verts.begin() + 9 * (off + k));
                mobj.insert(mobj.end(), (size_t)k, (int32_t)o);
            }
            off += k;
        }
        if (!mobj.empty()) {
            has_media = true;
            try {
                if (!build_wide_bvh(mverts.data(), (int64_t)mobj.size(), nullptr, BuildParams(), mbvh, err))
                    return fail(RISE_HIP_ERR_BUILD, "medium-boundary BVH build failed: " + err);
            } catch (const std::bad_alloc&) {
                return fail(RISE_HIP_ERR_OUT_OF_MEMORY, "host allocation failed during BVH build");
            }
            for (Tri48& t : mbvh.tris) {
                int32_t id;
                std::memcpy(&id, &t.v0[3], 4);
                std::memcpy(&t.v1[3], &mobj[(size_t)id], 4);
            }
        }
    }
    // Luminaries (LuminaryManager + LightSampler::Prepare): emitting objects,
    // area CDF over their triangles, selection weight = mean exitance * scale * area.
    std::vector<DevObject> dobj(n_objects);
    std::vector<DevLuminary> lums;
    std::vector<uint32_t> lum_tris;
    std::vector<float> lum_cdf;
    std::vector<double> weights;
    {
        int64_t off = 0;
        for (size_t o = 0; o < n_objects; ++o) {
            const int64_t k = objects[o].n_tris;
            dobj[o] = DevObject{objects[o].material, objects[o].double_sided ? 1 : 0, -1,
                                object_media ? object_media[o].interior_medium : -1};
            const DevMaterial& m = dmat[objects[o].material];
            if (m.emit_spd >= 0 && k > 0) {
                std::vector<double> areas((size_t)k);
                double tot = 0;
                for (int64_t i = 0; i < k; ++i) {
                    const float* f = verts.data() + 9 * (off + i);
                    const double ax = (double)f[3] - f[0], ay = (double)f[4] - f[1], az = (double)f[5] - f[2];
                    const double bx = (double)f[6] - f[0], by = (double)f[7] - f[1], bz = (double)f[8] - f[2];
                    const double cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
                    areas[(size_t)i] = std::sqrt(cx * cx + cy * cy + cz * cz) * 0.5;
                    tot += areas[(size_t)i];
                }
                const DevSpd& e = dspd[(size_t)m.emit_spd];
                double mean = 0;
                if (e.n == RISE_HIP_SPD_SIGMOID) {
                    // Mean over the 81 CIE sample wavelengths, 380..780 nm.
                    const float* c = &amps[(size_t)e.offset];
                    for (int j = 0; j < 81; ++j) mean += rise_hip_sigmoid_spectrum_eval(c, c[3], 380.0 + 5.0 * j);
                    mean /= 81.0;
                } else {
                    for (int j = 0; j < e.n; ++j) mean += amps[(size_t)(e.offset + j)];
                    mean /= e.n;
                }
                const double wgt = mean * m.emit_scale * tot;
                if (tot > 0 && wgt > 0) {
                    DevLuminary L{};
                    L.object = (int)o;
                    L.tri_first = (int)lum_tris.size();
                    L.tri_count = (int)k;
                    L.area = (float)tot;
                    L.material = objects[o].material;
                    double acc = 0;
                    for (int64_t i = 0; i < k; ++i) {
                        acc += areas[(size_t)i] / tot;
                        lum_tris.push_back(leaf_of[(size_t)(off + i)]);
                        lum_cdf.push_back((float)acc);
                    }
                    dobj[o].luminary = (int)lums.size();
                    lums.push_back(L);
                    weights.push_back(wgt);
                }
            }
            off += k;
        }
        double wt = 0;
        for (double w : weights) wt += w;
        double acc = 0;
        for (size_t i = 0; i < lums.size(); ++i) {
            acc += weights[i];
            lums[i].pdf_select = (float)(weights[i] / wt);
            lums[i].cdf_select = (float)(acc / wt);
        }
        if (!lums.empty()) lums.back().cdf_select = 1.0f;
    }

    rise_hip_pt_scene* s = new (std::nothrow) rise_hip_pt_scene;
    if (!s) return fail(RISE_HIP_ERR_OUT_OF_MEMORY, "host allocation failed");
    hipError_t he = hipGetDevice(&s->device);
    if (he != hipSuccess) {
        delete s;
        return fail_hip(he, "hipGetDevice");
    }
    rise_hip_status st = RISE_HIP_OK;
    PtSceneView& v = s->view;
    if (st == RISE_HIP_OK) st = upload(s, bvh.nodes, &v.nodes);
    if (st == RISE_HIP_OK) st = upload(s, bvh.tris, &v.tris);
    if (st == RISE_HIP_OK) st = upload(s, dobj, &v.objects);
    if (st == RISE_HIP_OK) st = upload(s, dmat, &v.materials);
    if (st == RISE_HIP_OK) st = upload(s, dspd, &v.spds);
    if (st == RISE_HIP_OK) st = upload(s, amps, &v.spd_amp);
    if (st == RISE_HIP_OK) st = upload(s, lums, &v.lums);
    if (st == RISE_HIP_OK) st = upload(s, lum_tris, &v.lum_tris);
    if (st == RISE_HIP_OK) st = upload(s, lum_cdf, &v.lum_cdf);
    if (st == RISE_HIP_OK && has_media) st = upload(s, dmed, &v.media);
    if (st == RISE_HIP_OK && has_media) st = upload(s, mbvh.nodes, &v.mnodes);
    if (st == RISE_HIP_OK && has_media) st = upload(s, mbvh.tris, &v.mtris);
    if (st != RISE_HIP_OK) {
        free_scene(s);
        return st;
    }
    v.depth = bvh.stats.depth;
    v.n_lums = (int)lums.size();
    s->n_tris = total;
    s->has_glossy = has_glossy;
    s->has_transmission = has_transmission;
    s->material_kinds = material_kinds;
    v.has_media = has_media ? 1 : 0;
    s->has_media = has_media;
    *out_scene = s;
    return RISE_HIP_OK;
}

extern "C" rise_hip_status rise_hip_pt_scene_release(rise_hip_pt_scene* scene)
{
    clear_error();
    if (scene) free_scene(scene);
    return RISE_HIP_OK;
}

// Job::AddHosekWilkieSkylight without the sun light: HosekWilkieSkyModel's
// constructor (Utilities/HosekWilkieSkyModel.cpp:98-150) evaluated once in
// FP64; the per-direction part runs on the device (sky_radiance).
namespace {

// The scene's directional lights as the kernels see them: the list set by
// rise_hip_pt_scene_set_directional_lights, then the sky's matched sun.
void refresh_dir_lights(rise_hip_pt_scene* s)
{
    PtSceneView& v = s->view;
    v.n_dir = 0;
    for (const DevDirLight& l : s->dir_lights) v.dir[v.n_dir++] = l;
    if (s->has_sun) v.dir[v.n_dir++] = s->sun;
}

// DirectionalLight (Lights/DirectionalLight.cpp:25-30, 80-100): normalized
// direction toward the light; Rec.709 luminance of the colour times the energy.
bool make_dir_light(const float dir[3], const float color[3], float power, DevDirLight& out)
{
    const double x = dir[0], y = dir[1], z = dir[2];
    const double m = std::sqrt(x * x + y * y + z * z);
    if (!(m > 0.0) || !std::isfinite(m) || !std::isfinite(power)) return false;
    out.d[0] = (float)(x / m);
    out.d[1] = (float)(y / m);
    out.d[2] = (float)(z / m);
    const double lum = 0.2126 * color[0] + 0.7152 * color[1] + 0.0722 * color[2];
    out.le = (float)(lum * (double)power);
    return true;
}

}  // namespace

extern "C" rise_hip_status rise_hip_pt_scene_set_directional_lights(rise_hip_pt_scene* scene,
                                                                    const rise_hip_directional_light_desc* lights,
                                                                    size_t n)
{
    clear_error();
    if (!scene) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "scene is NULL");
    if (n && !lights) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "lights is NULL with a nonzero count");
    if (n + (scene->has_sun ? 1 : 0) > (size_t)kMaxDirLights)
        return fail(RISE_HIP_ERR_NOT_SUPPORTED, "more than " + std::to_string(kMaxDirLights) +
                                                    " directional lights (a sky's sun counts as one)");
    std::vector<DevDirLight> v(n);
    for (size_t i = 0; i < n; ++i)
        if (!make_dir_light(lights[i].direction, lights[i].color, lights[i].power, v[i]))
            return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "directional light " + std::to_string(i) +
                                                           ": zero or non-finite direction or power");
    release_solo_batch(scene);  // a cached plan holds a copy of the scene tables
    scene->dir_lights = v;
    refresh_dir_lights(scene);
    return RISE_HIP_OK;
}

extern "C" rise_hip_status rise_hip_pt_scene_set_sky(rise_hip_pt_scene* scene, const rise_hip_sky_desc* sky)
{
    clear_error();
    if (!scene) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "scene is NULL");
    if (sky && (!(sky->turbidity >= 1.0f && sky->turbidity <= 64.0f) || !(sky->sky_scale >= 0.0f)))
        return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "bad sky turbidity or scale");
    if (sky && sky->create_sun && scene->dir_lights.size() + 1 > (size_t)kMaxDirLights)
        return fail(RISE_HIP_ERR_NOT_SUPPORTED, "no room for the sky's sun among the directional lights");
    if (sky && sky->create_sun && !std::isfinite(sky->sun_intensity_scale))
        return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "bad sun_intensity_scale");
    release_solo_batch(scene);  // a cached plan holds a copy of the scene tables
    DevSky& d = scene->view.sky;
    d = DevSky{};
    scene->view.env = nullptr;
    scene->has_sun = false;
    refresh_dir_lights(scene);
    if (!sky) return RISE_HIP_OK;
    const double el = sky->solar_elevation_deg * (M_PI / 180.0), az = sky->solar_azimuth_deg * (M_PI / 180.0);
    const double T = sky->turbidity;
    // Preetham 1999 Table 2: each coefficient is t * T + c.
    static const double kT[3][5] = {{0.1787, -0.3554, -0.0227, 0.1206, -0.0670},
                                    {-0.0193, -0.0665, -0.0004, -0.0641, -0.0033},
                                    {-0.0167, -0.0950, -0.0079, -0.0441, -0.0109}};
    static const double kC[3][5] = {{-1.4630, 0.4275, 5.3251, -2.5771, 0.3703},
                                    {-0.2592, 0.0008, 0.2125, -0.8989, 0.0452},
                                    {-0.2608, 0.0092, 0.2102, -1.6537, 0.0529}};
    const double ts = std::max(0.0, std::min(M_PI / 2.0, M_PI / 2.0 - el));
    double p[3][5];
    for (int c = 0; c < 3; ++c)
        for (int k = 0; k < 5; ++k) p[c][k] = kT[c][k] * T + kC[c][k];
    // Zenith luminance (kcd/m^2) and chromaticity (Preetham 1999 section 4.4).
    const double chi = (4.0 / 9.0 - T / 120.0) * (M_PI - 2.0 * ts);
    double yz = (4.0453 * T - 4.9710) * std::tan(chi) - 0.2155 * T + 2.4192;
    if (yz < 0) yz = 0;
    const double t1 = ts, t2 = ts * ts, t3 = t2 * ts;
    const double xz = (0.00166 * t3 - 0.00375 * t2 + 0.00209 * t1) * T * T +
                      (-0.02903 * t3 + 0.06377 * t2 - 0.03202 * t1 + 0.00394) * T +
                      (0.11693 * t3 - 0.21196 * t2 + 0.06052 * t1 + 0.25886);
    const double yc = (0.00275 * t3 - 0.00610 * t2 + 0.00317 * t1) * T * T +
                      (-0.04214 * t3 + 0.08970 * t2 - 0.04153 * t1 + 0.00516) * T +
                      (0.15346 * t3 - 0.26756 * t2 + 0.06670 * t1 + 0.26688);
    // Perez F at the zenith (cos theta = 1, gamma = theta_s): the ratio's denominator.
    auto perez = [&](int c, double cos_theta, double gamma) {
        const double sc = std::max(1e-3, cos_theta);
        const double cg = std::cos(gamma);
        return (1.0 + p[c][0] * std::exp(p[c][1] / sc)) * (1.0 + p[c][2] * std::exp(p[c][3] * gamma) + p[c][4] * cg * cg);
    };
    const double zv[3] = {yz, xz, yc};
    d.on = 1;
    d.theta_s = (float)ts;
    d.sun[0] = (float)(std::cos(el) * std::sin(az));
    d.sun[1] = (float)std::sin(el);
    d.sun[2] = (float)(std::cos(el) * std::cos(az));
    for (int c = 0; c < 3; ++c) {
        for (int k = 0; k < 5; ++k) d.pz[c][k] = (float)p[c][k];
        d.zen[c] = (float)(zv[c] / perez(c, 1.0, ts));
    }
    d.scale = sky->sky_scale;

    // The environment light sample's table (built in double on the host) and
    // the matched sun (Job::AddHosekWilkieSkylight, Job.cpp:2734-2746).
    SkyModelD md;
    for (int c = 0; c < 3; ++c)
        for (int k = 0; k < 5; ++k) md.p[c][k] = p[c][k];
    for (int c = 0; c < 3; ++c) md.zen[c] = zv[c];
    md.theta_s = ts;
    md.sun[0] = std::cos(el) * std::sin(az);
    md.sun[1] = std::sin(el);
    md.sun[2] = std::cos(el) * std::cos(az);
    std::vector<EnvTable> tab(1);
    build_env_table(md, tab[0]);
    if (tab[0].total > 0.0) {  // a black sky disables the sample (RayCaster.cpp:352-360)
        const EnvTable* dev = nullptr;
        const rise_hip_status st = upload(scene, tab, &dev);  // kept until the scene is released
        if (st != RISE_HIP_OK) return st;
        scene->view.env = dev;
    }
    if (sky->create_sun) {
        const float dir[3] = {(float)md.sun[0], (float)md.sun[1], (float)md.sun[2]};
        const float white[3] = {1.0f, 1.0f, 1.0f};
        make_dir_light(dir, white, sky->sun_intensity_scale, scene->sun);
        scene->has_sun = true;
    }
    refresh_dir_lights(scene);
    return RISE_HIP_OK;
}
