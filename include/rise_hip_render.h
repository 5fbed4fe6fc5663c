/*
 * rise_hip_render.h -- C ABI of the spectral (hero-wavelength) path tracer.
 *
 * Replaces the pass loop and per-pixel integration of the reference
 * (paths relative to src/Library/ of RISE):
 *   - IRasterizer::RasterizeScene (Interfaces/IRasterizer.h:70) driven by
 *     PixelBasedRasterizerHelper::RasterizeScene / RasterizeScenePass
 *     (Rendering/PixelBasedRasterizerHelper.cpp:876-1300), and
 *   - PathTracingSpectralRasterizer::IntegratePixel / IntegratePixelSpectral
 *     (Rendering/PathTracingSpectralRasterizer.cpp:121-479) with
 *     PathTracingIntegrator::IntegrateRayHWSS / IntegrateFromHitHWSS
 *     (Shaders/PathTracingIntegrator.cpp:3939-5126),
 * configured as RISE_API_CreatePathTracingSpectralRasterizer does
 * (RISE_API.h:3112-3141): HWSS with 4 wavelengths, box pixel filter, no
 * denoiser, no adaptive sampling.
 *
 * Conventions as in rise_hip.h.  All *_desc inputs are host memory and are
 * copied; film outputs are device memory.
 */
#ifndef RISE_HIP_RENDER_H
#define RISE_HIP_RENDER_H

#include "rise_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* A step spectrum: n equal bins over [lambda_begin, lambda_end]
 * (SpectralPacket, Utilities/Color/SpectralPacket.h:109-124, 285-297).
 * With n = RISE_HIP_SPD_SIGMOID it is instead an RGB-uplifted sigmoid
 * spectrum (RGBSigmoidPolynomial, see rise_hip_rgb_uplift.h): amplitudes =
 * {c0, c1, c2, scale}, range fixed at [380, 780] nm. */
#define RISE_HIP_SPD_SIGMOID (-1)
typedef struct rise_hip_spd_desc {
    float lambda_begin;
    float lambda_end;
    int32_t n;
    const float* amplitudes; /* n values (sigmoid: 4) */
} rise_hip_spd_desc;

typedef enum rise_hip_material_type {
    RISE_HIP_MATERIAL_LAMBERTIAN = 0, /* LambertianMaterial (+ LambertianLuminaireMaterial emitter) */
    RISE_HIP_MATERIAL_GGX = 1,        /* GGXMaterial, conductor Fresnel, isotropic roughness
                                         (Materials/GGXBRDF.cpp, GGXSPF.cpp) */
    RISE_HIP_MATERIAL_DIELECTRIC = 2  /* DielectricMaterial, smooth (Materials/DielectricSPF.cpp) */
} rise_hip_material_type;

typedef struct rise_hip_material_desc {
    int32_t type;            /* rise_hip_material_type */
    int32_t reflectance_spd; /* Lambertian reflectance; GGX diffuse `rd` (-1: none) */
    int32_t emission_spd;    /* -1: not an emitter (LambertianEmitter otherwise; Lambertian only) */
    float emission_scale;    /* LambertianEmitter scale: L = spd / pi * scale */
    int32_t specular_spd;    /* GGX specular tint `rs` */
    int32_t ior_spd;         /* GGX conductor n; dielectric index of refraction */
    int32_t extinction_spd;  /* GGX conductor k; dielectric `tau`, transmittance per unit length (-1: none) */
    float alpha;             /* GGX roughness (alphax = alphay) */
} rise_hip_material_desc;

/* A world-space triangle mesh with one material (an Object with identity
 * transform over a TriangleMeshGeometryIndexed with face normals). */
typedef struct rise_hip_object_desc {
    const float* vertices; /* n_tris * 9 floats */
    int64_t n_tris;
    int32_t material;
    int32_t double_sided;  /* TriangleMeshGeometryIndexed bDoubleSided */
} rise_hip_object_desc;

typedef struct rise_hip_pt_scene rise_hip_pt_scene;

/* Builds the BVH, material, spectrum and luminary tables and uploads them.
 * Objects whose material emits become area luminaries (LuminaryManager,
 * LightSampler::Prepare, Lights/LightSampler.cpp:635-760).  Synchronous. */
rise_hip_status rise_hip_pt_scene_create(const rise_hip_object_desc* objects, size_t n_objects,
                                         const rise_hip_material_desc* materials, size_t n_materials,
                                         const rise_hip_spd_desc* spds, size_t n_spds,
                                         rise_hip_pt_scene** out_scene);

/* Participating media (row "media" of the scene format): the scene-file
 * chunk `homogeneous_medium` with `phase hg g` (Job::AddHomogeneousMedium;
 * Materials/HomogeneousMedium.cpp, Materials/HenyeyGreensteinPhaseFunction.cpp).
 * RGB coefficients per unit length; the spectral renderer uses their
 * luminances at every wavelength (HomogeneousMedium::GetCoefficientsNM). */
typedef struct rise_hip_medium_desc {
    float sigma_a[3];  /* absorption */
    float sigma_s[3];  /* scattering */
    float hg_g;        /* Henyey-Greenstein asymmetry, in (-1, 1) */
    int32_t reserved;  /* 0 */
} rise_hip_medium_desc;

/* Per object: `standard_object`'s `interior_medium` (index into the media,
 * -1: none) and `casts_shadows` (0: shadow rays pass through it, the usual
 * setting for a medium's shell).  A medium is entered through a dielectric
 * boundary (the index-of-refraction stack, Utilities/MediumTracking.h); its
 * shell must be a closed single-sided mesh with outward normals, since shadow
 * rays enter and leave media by face orientation (LightSampler.cpp:300-330). */
typedef struct rise_hip_object_media {
    int32_t interior_medium;
    int32_t casts_shadows;
} rise_hip_object_media;

/* rise_hip_pt_scene_create with participating media: `object_media` has one
 * entry per object (NULL: no media, every object casts shadows). */
rise_hip_status rise_hip_pt_scene_create_with_media(const rise_hip_object_desc* objects, size_t n_objects,
                                                    const rise_hip_material_desc* materials, size_t n_materials,
                                                    const rise_hip_spd_desc* spds, size_t n_spds,
                                                    const rise_hip_medium_desc* media, size_t n_media,
                                                    const rise_hip_object_media* object_media,
                                                    rise_hip_pt_scene** out_scene);

rise_hip_status rise_hip_pt_scene_release(rise_hip_pt_scene* scene);

/* The spectral sky of the scene-file chunk `hosek_wilkie_skylight`
 * (Job::AddHosekWilkieSkylight, Job.cpp:2717-2750; the reference's model is
 * Preetham 1999, Utilities/HosekWilkieSkyModel.cpp): a radiance map seen by
 * rays leaving the scene and importance-sampled by next-event estimation
 * (EnvironmentSampler, Rendering/EnvironmentSampler.cpp; LightSampler.cpp:
 * 2278-2358), with BSDF-sampled escapes weighted against that sample by the
 * power heuristic; and, when create_sun is set, its matched directional light
 * `__hw_sun__` toward the sun with colour (1, 1, 1) and power
 * sun_intensity_scale.  Renders and batches created afterwards use it (NULL
 * removes both). */
typedef struct rise_hip_sky_desc {
    float solar_elevation_deg;  /* solar_elevation (45) */
    float solar_azimuth_deg;    /* solar_azimuth (0) */
    float turbidity;            /* turbidity (3) */
    float sky_scale;            /* sky_intensity_scale (1) */
    int32_t create_sun;         /* create_sun (the chunk's default is TRUE) */
    float sun_intensity_scale;  /* sun_intensity_scale (3.14) */
} rise_hip_sky_desc;

rise_hip_status rise_hip_pt_scene_set_sky(rise_hip_pt_scene* scene, const rise_hip_sky_desc* sky);

/* Directional lights (the scene-file chunk `directional_light`,
 * Lights/DirectionalLight.cpp): light arriving from `direction` (toward the
 * light; normalized here) with RGB `color` times `power`.  Next-event
 * estimation evaluates each one deterministically at every surface vertex
 * and wavelength as ComputeDirectLightingNM does (LightSampler.cpp:1968-2000):
 * Rec.709 luminance of the colour x power x BSDF x cos, with a shadow ray to
 * infinity; no random numbers, no MIS.  Replaces the scene's earlier list; a
 * sky's matched sun comes on top.  At most 3 lights with a sun, 4 without. */
typedef struct rise_hip_directional_light_desc {
    float direction[3];
    float color[3];
    float power;
} rise_hip_directional_light_desc;

rise_hip_status rise_hip_pt_scene_set_directional_lights(rise_hip_pt_scene* scene,
                                                         const rise_hip_directional_light_desc* lights, size_t n);

/* Pinhole camera: vertical field of view, square pixels, pixel row 0 at the top. */
typedef struct rise_hip_camera_desc {
    float position[3];
    float look_at[3];
    float up[3];
    float fov_deg;
} rise_hip_camera_desc;

/* Random numbers of the bounce loops: every sampler.Get1D() of the
 * reference's integrator draws from an Owen-scrambled Sobol sequence
 * (Sampling/SobolSequence.h, Utilities/SobolSampler.h, ZSobolSampler.h) at
 * the dimension the reference's consumption order gives it, per pixel
 * (PathTracingSpectralRasterizer.cpp:229-345).
 *   RISE_HIP_SAMPLER_ZSOBOL (0, the reference's default, blue_noise_sampler
 *     TRUE): index Morton(x, y) * 2^ceil(log2 spp) + sample, and the sample's
 *     position inside the pixel's block shuffled per dimension pair.
 *   RISE_HIP_SAMPLER_SOBOL (1, blue_noise_sampler FALSE): index = sample,
 *     seed HashCombine(x, y), shuffled the same way.
 *   RISE_HIP_SAMPLER_ZSOBOL_UNSHUFFLED (2): the Z-order draws exactly as the
 *     reference makes them.  Every dimension pair then reuses one index, and
 *     the estimate is biased (about -1% on a spectral furnace); kept for
 *     comparison with the reference's noise.
 * The hero wavelength, the per-pixel seed of the pixel positions and
 * free-flight distances, which the reference draws from a Mersenne Twister,
 * come from a counter-based hash of (seed, path). */
typedef enum rise_hip_sampler {
    RISE_HIP_SAMPLER_ZSOBOL = 0,
    RISE_HIP_SAMPLER_SOBOL = 1,
    RISE_HIP_SAMPLER_ZSOBOL_UNSHUFFLED = 2
} rise_hip_sampler;

typedef struct rise_hip_render_desc {
    int32_t width;
    int32_t height;
    int32_t spp;          /* samples rendered by this call */
    int32_t spp_offset;   /* index of the first sample (progressive passes) */
    int32_t spp_total;    /* samples per pixel of the whole render (random-number layout) */
    float lambda_begin;   /* HWSS range, nm */
    float lambda_end;
    int32_t max_diffuse_bounce;      /* StabilityConfig bounce caps (Utilities/StabilityConfig.h:103-112) */
    int32_t max_glossy_bounce;
    int32_t max_transmission_bounce;
    int32_t rr_min_depth;            /* Russian roulette (PathTransportUtilities.h:86-110) */
    float rr_threshold;
    uint64_t seed;
    rise_hip_camera_desc camera;
    int32_t sky_background;          /* camera rays that miss see the sky (radiance_background) */
    int32_t sampler;                 /* rise_hip_sampler: the bounce loops' random numbers */
} rise_hip_render_desc;

typedef struct rise_hip_render_stats {
    int64_t extension_rays;  /* closest-hit queries */
    int64_t shadow_rays;     /* any-hit queries (next-event estimation) */
    int64_t paths;           /* camera samples */
    int64_t bounces;         /* extension passes launched */
    int64_t ior_overflows;   /* refractions into a 4th nested dielectric: the index-of-refraction stack holds
                                3 objects (the reference's is unbounded, Utilities/IORStack.h:95), so such a
                                path goes on with the enclosing index; nonzero means results differ there */
} rise_hip_render_stats;

/* Renders samples [spp_offset, spp_offset + spp) of every pixel.
 * film_xyz: device, width*height*3 floats: per-pixel mean XYZ of those
 *   samples (the ProgressiveFilm colorSum / alphaSum resolve,
 *   Rendering/ProgressiveFilm.h:36-60, before the XYZ -> RGB conversion).
 * film_xyz_sq: device, optional (NULL): per-pixel mean of squared XYZ.
 * stats: host, optional (NULL); when given the call waits for the stream.
 * The scene keeps the work buffers and tables of its last call and reuses
 * them while the request is unchanged; a changed request waits for the
 * previous call's stream before replacing them. */
rise_hip_status rise_hip_render(rise_hip_pt_scene* scene, const rise_hip_render_desc* desc, float* film_xyz,
                                float* film_xyz_sq, rise_hip_render_stats* stats, hipStream_t stream);

/* ---- Batches: several renders (scenes, image sizes, sample counts) in one
 * launch sequence ----
 *
 * Replaces a loop of Job::Rasterize calls (Job.cpp:9268), one per job, each
 * running IRasterizer::RasterizeScene to completion: here the requests share
 * one pool of paths in flight, so small renders fill the GPU together.  Each
 * request's film is identical to what rise_hip_render writes for it alone. */
typedef struct rise_hip_render_request {
    rise_hip_pt_scene* scene;
    rise_hip_render_desc desc;
    float* film_xyz;     /* device, width*height*3 floats (as rise_hip_render) */
    float* film_xyz_sq;  /* device, optional */
} rise_hip_render_request;

typedef struct rise_hip_batch rise_hip_batch;

/* Plans the batch (pass layout, work buffers, device tables).  Synchronous.
 * The scenes must outlive the batch; the films are written by each launch. */
rise_hip_status rise_hip_batch_create(const rise_hip_render_request* requests, size_t n_requests,
                                      rise_hip_batch** out_batch);

/* Launch-shape choices of a batch (0 everywhere: defaults).  They change
 * speed only: films are the same under every tuning. */
typedef struct rise_hip_batch_tuning {
    int32_t max_paths;    /* paths in flight per pass (0: 4194304); at least one chunk (<= 1048576 paths) */
    int32_t trace_block;  /* workgroup size of the traversal kernels: 64, 128 or 256 (0: 128) */
    int32_t stack;        /* traversal stack: 0 auto (LDS, sized by tree depth), 8 or 16 LDS entries per
                             lane when the trees fit, -1 private memory */
    int32_t persistent;   /* traversal schedule: 0 auto (persistent waves for scenes of >= 4096
                             triangles), 1 persistent waves with dynamic fetch, 2 one ray per thread */
} rise_hip_batch_tuning;

/* rise_hip_batch_create with explicit tuning (NULL: defaults).  No
 * counterpart in the reference: its CPU scheduler has no launch shapes. */
rise_hip_status rise_hip_batch_create_tuned(const rise_hip_render_request* requests, size_t n_requests,
                                            const rise_hip_batch_tuning* tuning, rise_hip_batch** out_batch);

/* The tuning a batch resolved to (auto choices made concrete). */
rise_hip_status rise_hip_batch_get_tuning(const rise_hip_batch* batch, rise_hip_batch_tuning* out_tuning);

/* Enqueues the whole batch on `stream`: kernel launches only (no copies, no
 * allocation, no host synchronization), so the call may be captured in a
 * hipGraph and replayed.  Launches of one batch must not overlap. */
rise_hip_status rise_hip_batch_launch(rise_hip_batch* batch, hipStream_t stream);

/* Ray and path counts of the last launch (waits for `stream`). */
rise_hip_status rise_hip_batch_get_stats(rise_hip_batch* batch, hipStream_t stream, rise_hip_render_stats* out_stats);

rise_hip_status rise_hip_batch_release(rise_hip_batch* batch);

/* ---- Serving: continuous batching of render requests ----
 *
 * Replaces the reference's one-job-at-a-time rendering (Job::Rasterize,
 * Job.cpp:9268, called per request by its front ends) for a stream of
 * requests: scenes stay resident on the device (rise_hip_pt_scene), requests
 * are submitted at any time, admitted while their film accumulators (48 B
 * per pixel) fit the HBM budget, and share every pass with the requests
 * already rendering.  Each film is bit-identical to rise_hip_render of that
 * request alone.  All work runs on the stream given at creation. */
typedef struct rise_hip_server_config {
    int64_t hbm_budget_bytes;     /* film accumulators of admitted requests (0: 8 GiB); a request larger
                                     than the budget is still admitted when nothing else is */
    int32_t max_requests;         /* admitted at once (0: 256, at most 256) */
    int32_t reserved;
    rise_hip_batch_tuning tuning; /* launch shapes (max_paths: the path pool) */
} rise_hip_server_config;

typedef struct rise_hip_server_stats {
    int64_t submitted, completed, admitted, waiting;
    int64_t passes, paths, extension_rays, shadow_rays;  /* over completed passes */
    int64_t hbm_in_use;                                  /* film accumulators now allocated */
    int64_t cancelled, timed_out, failed;                /* finished requests by their status */
    int64_t alloc_retries;                               /* film allocations that failed and were retried */
    int64_t ior_overflows;                               /* as rise_hip_render_stats::ior_overflows */
} rise_hip_server_stats;

/* How a served request finished. */
typedef enum rise_hip_request_status {
    RISE_HIP_REQUEST_DONE = 0,       /* rendered; the film holds the result */
    RISE_HIP_REQUEST_CANCELLED = 1,  /* rise_hip_server_cancel; the film is not written */
    RISE_HIP_REQUEST_TIMED_OUT = 2,  /* its timeout passed before its last pass was launched */
    RISE_HIP_REQUEST_FAILED = 3      /* its film accumulators could not be allocated */
} rise_hip_request_status;

typedef struct rise_hip_server_result {
    uint64_t ticket;
    int64_t latency_ns;  /* submit to finish */
    int32_t status;      /* rise_hip_request_status */
    int32_t reserved;
} rise_hip_server_result;

typedef struct rise_hip_server rise_hip_server;

rise_hip_status rise_hip_server_create(const rise_hip_server_config* config, hipStream_t stream,
                                       rise_hip_server** out_server);

/* Queues a request (host call; nothing waits).  The scene and the films must
 * stay valid until the request is reported done. */
rise_hip_status rise_hip_server_submit(rise_hip_server* server, const rise_hip_render_request* request,
                                       uint64_t* out_ticket);

/* Makes progress without blocking: retires finished passes, admits waiting
 * requests, launches passes (up to 4 in flight).  Reports up to `capacity`
 * finished requests: their tickets and submit-to-finish latency in ns. */
rise_hip_status rise_hip_server_poll(rise_hip_server* server, uint64_t* tickets, int64_t* latency_ns,
                                     size_t capacity, size_t* n_done);

/* As poll, but blocks until every submitted request has finished. */
rise_hip_status rise_hip_server_drain(rise_hip_server* server, uint64_t* tickets, int64_t* latency_ns,
                                      size_t capacity, size_t* n_done);

/* poll / drain reporting each finished request's status as well. */
rise_hip_status rise_hip_server_poll_results(rise_hip_server* server, rise_hip_server_result* results,
                                             size_t capacity, size_t* n_done);
rise_hip_status rise_hip_server_drain_results(rise_hip_server* server, rise_hip_server_result* results,
                                              size_t capacity, size_t* n_done);

/* Cancels a request (the reference's IsCancelled check between passes,
 * Rendering/PixelBasedRasterizerHelper.cpp:964-967).  A waiting request
 * finishes at once; an admitted one gets no further chunks, its film is left
 * unwritten and it finishes once the passes already holding its chunks have
 * run.  A request whose last pass is already launched completes normally.
 * Fails for an unknown or already finished ticket. */
rise_hip_status rise_hip_server_cancel(rise_hip_server* server, uint64_t ticket);

/* Ends the request as RISE_HIP_REQUEST_TIMED_OUT, as rise_hip_server_cancel
 * does, if its last pass has not been launched `timeout_ns` after its
 * submission (0: no timeout).  Checked at every poll and drain. */
rise_hip_status rise_hip_server_set_timeout(rise_hip_server* server, uint64_t ticket, int64_t timeout_ns);

/* Admission and its failures: a request whose film accumulators cannot be
 * allocated stays first in the queue and is tried again at later calls, as
 * memory frees; after 3 failures while nothing else is admitted it finishes
 * as RISE_HIP_REQUEST_FAILED.  Test hook with no counterpart in the
 * reference: the next n film allocations report out-of-memory without
 * allocating, so tests can force that path from the host. */
rise_hip_status rise_hip_server_fail_next_allocations(rise_hip_server* server, int32_t n);

rise_hip_status rise_hip_server_get_stats(const rise_hip_server* server, rise_hip_server_stats* out_stats);

/* Waits for the stream, then frees everything (unfinished requests are dropped). */
rise_hip_status rise_hip_server_release(rise_hip_server* server);

#ifdef __cplusplus
}
#endif

#endif /* RISE_HIP_RENDER_H */
