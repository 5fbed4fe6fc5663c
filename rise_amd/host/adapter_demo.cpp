// adapter_demo.cpp -- C++ caller of the GPU path through the host adapters.
// Built by tests (hipcc host compile, linked against librise_hip.so).
//   adapter_demo raycast   closest hit for a grid of rays through a triangle
//   adapter_demo render    a tiny spectral render of a lit floor
//   adapter_demo scene F   a scene file loaded by LoadAsciiScene, rendered at 64x64, 4 spp, written
//                          through its file_rasterizeroutput chunk to /tmp/adapter_demo_output.img
// Prints "ok ..." and exits 0 on success.
#include "rise_hip_adapters.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

using rise_hip_host::HipObjectManager;
using rise_hip_host::HipSpectralRasterizer;

static int raycast()
{
    // One triangle (0,0,0)-(1,0,0)-(0,1,0); rays straight down on a 4x4 grid.
    const float tri[9] = {0, 0, 0, 1, 0, 0, 0, 1, 0};
    HipObjectManager om;
    om.AddMesh(tri, 1);
    if (!om.PrepareForRendering()) return 1;
    const int n = 16;
    std::vector<float> rays(6 * n);
    for (int i = 0; i < n; ++i) {
        rays[0 * n + i] = 0.1f + 0.25f * (i % 4);
        rays[1 * n + i] = 0.1f + 0.25f * (i / 4);
        rays[2 * n + i] = 2.0f;
        rays[3 * n + i] = 0.0f;
        rays[4 * n + i] = 0.0f;
        rays[5 * n + i] = -1.0f;
    }
    float* d_rays = nullptr;
    int32_t* d_id = nullptr;
    float* d_t = nullptr;
    if (hipMalloc((void**)&d_rays, rays.size() * 4) != hipSuccess || hipMalloc((void**)&d_id, n * 4) != hipSuccess ||
        hipMalloc((void**)&d_t, n * 4) != hipSuccess)
        return 2;
    hipMemcpy(d_rays, rays.data(), rays.size() * 4, hipMemcpyHostToDevice);
    if (!om.IntersectRay(d_rays, n, d_id, d_t, nullptr)) return 3;
    std::vector<int32_t> id(n);
    std::vector<float> t(n);
    hipMemcpy(id.data(), d_id, n * 4, hipMemcpyDeviceToHost);
    hipMemcpy(t.data(), d_t, n * 4, hipMemcpyDeviceToHost);
    int hits = 0;
    for (int i = 0; i < n; ++i) {
        const float x = rays[i], y = rays[n + i];
        const bool inside = x + y <= 1.0f;
        if (inside != (id[i] == 0)) return 4;
        if (inside && std::fabs(t[i] - 2.0f) > 1e-6f) return 5;
        hits += inside;
    }
    hipFree(d_rays);
    hipFree(d_id);
    hipFree(d_t);
    std::printf("ok raycast %d hits\n", hits);
    return 0;
}

static int render()
{
    // A grey floor under a downward-facing emitter quad.
    const float floor[18] = {-4, 0, -4, -4, 0, 4, 4, 0, 4, -4, 0, -4, 4, 0, 4, 4, 0, -4};
    const float light[18] = {-0.5f, 2, -0.5f, 0.5f, 2, -0.5f, 0.5f, 2, 0.5f, -0.5f, 2, -0.5f, 0.5f, 2, 0.5f,
                             -0.5f, 2, 0.5f};
    HipSpectralRasterizer r;
    const int grey = r.AddSpectrum(380, 720, {0.5f});
    const int emit = r.AddSpectrum(380, 720, {10.0f});
    const int mf = r.AddLambertianMaterial(grey);
    const int ml = r.AddLambertianMaterial(grey, emit, 3.14159265f);
    r.AddObject(floor, 2, mf);
    r.AddObject(light, 2, ml);
    if (!r.PrepareForRendering()) return 1;
    rise_hip_render_desc d{};
    d.width = d.height = 32;
    d.spp = d.spp_total = 8;
    d.lambda_begin = 380;
    d.lambda_end = 720;
    d.max_diffuse_bounce = 3;
    d.rr_min_depth = 3;
    d.rr_threshold = 0.05f;
    d.seed = 1;
    const float pos[3] = {0, 1, 5}, look[3] = {0, 0, 0}, up[3] = {0, 1, 0};
    std::memcpy(d.camera.position, pos, 12);
    std::memcpy(d.camera.look_at, look, 12);
    std::memcpy(d.camera.up, up, 12);
    d.camera.fov_deg = 45;
    float* film = nullptr;
    if (hipMalloc((void**)&film, 32 * 32 * 3 * 4) != hipSuccess) return 2;
    rise_hip_render_stats st{};
    if (!r.RasterizeScene(d, film, nullptr, &st)) return 3;
    std::vector<float> h(32 * 32 * 3);
    hipMemcpy(h.data(), film, h.size() * 4, hipMemcpyDeviceToHost);
    hipFree(film);
    double y = 0;
    for (int i = 0; i < 32 * 32; ++i) y += h[3 * i + 1];
    y /= 32 * 32;
    if (!(y > 0) || st.paths != 32 * 32 * 8) return 4;
    std::printf("ok render meanY %.4f paths %lld rays %lld\n", y, (long long)st.paths,
                (long long)(st.extension_rays + st.shadow_rays));
    return 0;
}

static int scene(const char* path)
{
    HipSpectralRasterizer r;
    rise_hip_render_desc d{};
    std::vector<std::pair<rise_hip_output_desc, std::string>> outs;
    if (!r.LoadAsciiScene(path, &d, &outs)) return 1;
    d.width = d.height = 64;
    d.spp = d.spp_total = 4;
    float* film = nullptr;
    if (hipMalloc((void**)&film, 64 * 64 * 3 * 4) != hipSuccess) return 2;
    rise_hip_render_stats st{};
    if (!r.RasterizeScene(d, film, nullptr, &st)) return 3;
    std::vector<float> h(64 * 64 * 3);
    hipMemcpy(h.data(), film, h.size() * 4, hipMemcpyDeviceToHost);
    // the scene's own output chunk, written next to the working directory's temp files
    const bool wrote = !outs.empty() && HipSpectralRasterizer::OutputImage(film, 64, 64, outs[0].first,
                                                                           "/tmp/adapter_demo_output.img", nullptr);
    hipFree(film);
    double y = 0;
    for (int i = 0; i < 64 * 64; ++i) y += h[3 * i + 1];
    y /= 64 * 64;
    if (!(y > 0) || st.paths != 64 * 64 * 4 || !wrote) return 4;
    std::printf("ok scene meanY %.4f paths %lld output %s\n", y, (long long)st.paths, outs[0].second.c_str());
    return 0;
}

int main(int argc, char** argv)
{
    if (argc > 1 && std::strcmp(argv[1], "raycast") == 0) return raycast();
    if (argc > 1 && std::strcmp(argv[1], "render") == 0) return render();
    if (argc > 2 && std::strcmp(argv[1], "scene") == 0) return scene(argv[2]);
    std::fprintf(stderr, "usage: adapter_demo raycast|render|scene <file>\n");
    return 64;
}
