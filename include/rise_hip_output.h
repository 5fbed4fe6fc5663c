/*
 * rise_hip_output.h -- the output stage: a rendered XYZ film resolved to an
 * image file.
 *
 * Replaces the reference's FileRasterizerOutput path for the formats it
 * writes without external libraries (Rendering/FileRasterizerOutput.cpp,
 * Rendering/FrameEncoders.cpp:47-160, Rendering/DisplayTransformWriter.cpp,
 * Rendering/DisplayTransform.h, RasterImages/{PPM,TGA,HDR}Writer.cpp), plus
 * PNG and EXR written here directly (the reference uses libpng / OpenEXR):
 *   film XYZ -> linear Rec.709 (ColorUtils::XYZtoRec709RGB, with its gamut
 *   clip) -> for LDR formats, exposure 2^EV and a display transform (none,
 *   Reinhard, ACES, AgX, Hable) -> output colour space (linear Rec.709 or
 *   sRGB) -> clamp and integerize (8 or 16 bits) or, for HDR formats, RGBE
 *   (Radiance .hdr) or half / float (OpenEXR, uncompressed scanlines).
 * The per-pixel work runs in one GPU kernel over the film in device memory
 * (double precision, the reference's operation order, so PPM / TGA / HDR
 * bytes equal the reference writers'); the host adds headers and writes.
 */
#ifndef RISE_HIP_OUTPUT_H
#define RISE_HIP_OUTPUT_H

#include "rise_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum rise_hip_image_format {
    RISE_HIP_IMAGE_TGA = 0, /* 32-bit BGRA, bottom-left origin flag as the reference writes it */
    RISE_HIP_IMAGE_PPM = 1, /* binary P6, 8 bits */
    RISE_HIP_IMAGE_PNG = 2, /* RGBA, 8 or 16 bits (bpp) */
    RISE_HIP_IMAGE_HDR = 3, /* Radiance RGBE, flat scanlines */
    RISE_HIP_IMAGE_EXR = 6  /* OpenEXR scanline, no compression; float channels if bpp >= 32, else half */
} rise_hip_image_format;

typedef enum rise_hip_color_space {
    RISE_HIP_COLOR_REC709_LINEAR = 0,
    RISE_HIP_COLOR_SRGB = 1
} rise_hip_color_space;

typedef enum rise_hip_display_transform {
    RISE_HIP_DISPLAY_DEFAULT = -1, /* ACES for LDR formats, none for HDR formats */
    RISE_HIP_DISPLAY_NONE = 0,
    RISE_HIP_DISPLAY_REINHARD = 1,
    RISE_HIP_DISPLAY_ACES = 2,
    RISE_HIP_DISPLAY_AGX = 3,
    RISE_HIP_DISPLAY_HABLE = 4
} rise_hip_display_transform;

typedef struct rise_hip_output_desc {
    int32_t format;            /* rise_hip_image_format */
    int32_t bpp;               /* PNG: 16, else 8; EXR: >= 32 float, else half; others: 8 */
    int32_t color_space;       /* rise_hip_color_space */
    int32_t display_transform; /* rise_hip_display_transform; ignored (none) for HDR formats */
    double exposure_ev;        /* LDR formats only, as the reference */
    int32_t with_alpha;        /* EXR: write an A channel (1.0) */
    int32_t reserved;
} rise_hip_output_desc;

/* Resolves a device film (height x width x 3 floats, mean XYZ as
 * rise_hip_render writes it) on `stream` and writes the file at `path`
 * (waits for the stream). */
rise_hip_status rise_hip_output_write(const float* film_xyz, int32_t width, int32_t height,
                                      const rise_hip_output_desc* desc, const char* path, void* stream);

/* The per-pixel stage alone, device to device on `stream` (no copies, waits
 * or allocations, so it can be captured in a graph): writes width x height
 * pixels of 4 bytes (TGA / PPM / PNG 8-bit: R, G, B, A; HDR: R, G, B, E),
 * 8 bytes (PNG 16-bit: four native-endian uint16) or 12 bytes (EXR: three
 * floats) to `out`, row-major, film row order. */
rise_hip_status rise_hip_output_resolve(const float* film_xyz, int32_t width, int32_t height,
                                        const rise_hip_output_desc* desc, void* out, void* stream);

/* The same from a film in host memory, resolved on host threads with the
 * same per-pixel code (no GPU). */
rise_hip_status rise_hip_output_write_host(const float* film_xyz, int32_t width, int32_t height,
                                           const rise_hip_output_desc* desc, const char* path);

/* The file name the reference forms for a file_rasterizeroutput pattern:
 * pattern + "." + the format's extension (tga, ppm, png, hdr, exr). */
const char* rise_hip_output_extension(int32_t format);

#ifdef __cplusplus
}
#endif

#endif /* RISE_HIP_OUTPUT_H */
