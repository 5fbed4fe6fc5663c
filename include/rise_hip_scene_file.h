/*
 * rise_hip_scene_file.h -- RISE scene files (native "RISE ASCII SCENE 7"
 * text) read into the spectral path tracer's scene tables.
 *
 * Replaces, for the chunks the path-tracing workloads use, the reference's
 * scene loading: Cst::ParseToCst + Cst::DeriveToJob (src/Library/Cst/Cst.cpp)
 * dispatching to the chunk parsers of Parsers/ChunkParserRegistry.cpp, which
 * call the IJob builders (Job.cpp).  Supported chunks and the reference
 * parsers they follow:
 *   spectral_painter (:1086-1170)          -> a spectrum (SpectralPacket of the
 *                                            control points, Job.cpp:1557-1582)
 *   uniformcolor_painter (:1005-1040)      -> the RGB-uplifted sigmoid spectrum
 *                                            (colorspace Rec709RGB_Linear, RISERGB
 *                                            or sRGB; see rise_hip_rgb_uplift.h)
 *   lambertian_material (:3090)            lambertian_luminaire_material (:3373)
 *   ggx_material (:3826; conductor Fresnel, isotropic, not emissive)
 *   dielectric_material (:3200; smooth: scattering >= 1e6)
 *   homogeneous_medium (:6142; RGB coefficients, phase "hg g" or "isotropic")
 *   hosek_wilkie_skylight (:6931; create_sun FALSE)
 *   rawmesh2_geometry (:5104; Geometry/TriangleMeshLoaderRAW2.cpp format)
 *   risemesh_geometry (:5137; .risemesh v1-v5, its BVH2 cache used for the
 *                      object's subtree; see rise_hip_risemesh.h)
 *   clippedplane_geometry (:5004; as two triangles)
 *   standard_object (:6411; position / orientation / scale, interior_medium,
 *                    casts_shadows)
 *   pinhole_camera (:4282)   film (:4814)
 *   pathtracing_spectral_rasterizer (:8894)
 *   file_rasterizeroutput (:9207; see rise_hip_output.h)
 * Chunks with no effect on the spectral path tracer's image (shaders and
 * shader ops, the other rasterizers' settings) are skipped;
 * any other chunk, and any parameter value of a supported chunk that the
 * GPU path cannot represent, is an error naming the chunk and line.
 * Parsing needs no GPU.
 */
#ifndef RISE_HIP_SCENE_FILE_H
#define RISE_HIP_SCENE_FILE_H

#include "rise_hip_output.h"
#include "rise_hip_render.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rise_hip_scene_file rise_hip_scene_file;

typedef struct rise_hip_scene_file_info {
    int64_t n_objects;
    int64_t n_materials;
    int64_t n_spds;
    int64_t n_media;
    int64_t n_tris;
    int32_t has_sky;
    int32_t n_skipped_chunks;  /* chunks skipped as irrelevant to the image */
    int32_t n_outputs;         /* file_rasterizeroutput chunks */
    int32_t n_directional_lights;  /* directional_light chunks (a sky's sun not counted) */
} rise_hip_scene_file_info;

/* Reads and checks a scene file (relative mesh and spectrum files resolve
 * against the scene file's directory). */
rise_hip_status rise_hip_scene_file_parse(const char* path, rise_hip_scene_file** out_file);

rise_hip_status rise_hip_scene_file_get_info(const rise_hip_scene_file* file, rise_hip_scene_file_info* out_info);

/* The tables rise_hip_pt_scene_create_with_media takes (owned by `file`);
 * *sky is NULL when the scene has no sky. */
rise_hip_status rise_hip_scene_file_get_tables(const rise_hip_scene_file* file,
                                               const rise_hip_object_desc** objects,
                                               const rise_hip_object_media** object_media,
                                               const rise_hip_material_desc** materials,
                                               const rise_hip_spd_desc** spds, const rise_hip_medium_desc** media,
                                               const rise_hip_sky_desc** sky);

/* The file's directional_light chunks (n_directional_lights of them, owned by
 * `file`), as rise_hip_pt_scene_set_directional_lights takes them. */
rise_hip_status rise_hip_scene_file_get_directional_lights(const rise_hip_scene_file* file,
                                                           const rise_hip_directional_light_desc** lights);

/* The render the file's camera, film and rasterizer chunks describe. */
rise_hip_status rise_hip_scene_file_get_render_desc(const rise_hip_scene_file* file, rise_hip_render_desc* out_desc);

/* Output `index` (0 .. n_outputs-1) of the file's file_rasterizeroutput
 * chunks (FileRasterizerOutput, Parsers/ChunkParserRegistry.cpp:9207-9325):
 * its description for rise_hip_output_write and its path pattern (owned by
 * `file`; the file name is pattern + "." + rise_hip_output_extension). */
rise_hip_status rise_hip_scene_file_get_output(const rise_hip_scene_file* file, int32_t index,
                                               rise_hip_output_desc* desc, const char** pattern);

/* Builds the device scene (rise_hip_pt_scene_create_with_media, then
 * rise_hip_pt_scene_set_directional_lights and rise_hip_pt_scene_set_sky when
 * the file has them). */
rise_hip_status rise_hip_scene_file_upload(const rise_hip_scene_file* file, rise_hip_pt_scene** out_scene);

rise_hip_status rise_hip_scene_file_release(rise_hip_scene_file* file);

#ifdef __cplusplus
}
#endif

#endif /* RISE_HIP_SCENE_FILE_H */
