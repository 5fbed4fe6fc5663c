// output.hip -- the output stage: XYZ film -> image file.
//
// One kernel converts every pixel (colour pipeline of output_color.h, FP64)
// into the bytes or floats the format stores; the host then frames them
// (headers, row order, PNG's zlib stream, EXR's offset table) and writes the
// file.  The kernel reads 12 bytes and writes 4-16 bytes per pixel: a
// bandwidth-bound pass over a film already in HBM.
#include "rise_hip_output.h"

#include "api_common.h"
#include "output_color.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <thread>
#include <vector>

using namespace rise_hip;

namespace {

enum Mode : int { kRgba8 = 0, kRgba16 = 1, kRgbe = 2, kFloat3 = 3 };

struct Plan {
    out::Params p;
    int mode;
    size_t bytes_per_px;
};

bool is_hdr(int format) { return format == RISE_HIP_IMAGE_HDR || format == RISE_HIP_IMAGE_EXR; }

rise_hip_status make_plan(const rise_hip_output_desc* d, int32_t w, int32_t h, Plan& pl)
{
    if (!d) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL output description");
    if (w <= 0 || h <= 0 || (int64_t)w * h > (int64_t)1 << 31)
        return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "bad image size");
    if (d->color_space != RISE_HIP_COLOR_REC709_LINEAR && d->color_space != RISE_HIP_COLOR_SRGB)
        return fail(RISE_HIP_ERR_NOT_SUPPORTED, "colour space must be linear Rec.709 or sRGB");
    if (d->display_transform < -1 || d->display_transform > 4)
        return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "unknown display transform");
    const bool hdr = is_hdr(d->format);
    switch (d->format) {
    case RISE_HIP_IMAGE_TGA:
    case RISE_HIP_IMAGE_PPM:
        pl.mode = kRgba8;
        break;
    case RISE_HIP_IMAGE_PNG:  // PNGWriter: 16 bits on request, else 8
        pl.mode = d->bpp == 16 ? kRgba16 : kRgba8;
        break;
    case RISE_HIP_IMAGE_HDR:
        pl.mode = kRgbe;
        break;
    case RISE_HIP_IMAGE_EXR:  // EXRFrameEncoder: float channels from 32 bits up, else half
        pl.mode = kFloat3;
        break;
    default:
        return fail(RISE_HIP_ERR_NOT_SUPPORTED, "unsupported image format");
    }
    // FileRasterizerOutput / FrameEncoderBase::Encode: exposure and display
    // transforms are for LDR formats only, and wrap the writer only when one
    // of them is not the identity.
    const int dt = d->display_transform < 0 ? (hdr ? 0 : 2) : d->display_transform;
    pl.p.color_space = d->color_space;
    pl.p.display = hdr ? 0 : dt;
    pl.p.exposure = hdr ? 1.0 : std::pow(2.0, d->exposure_ev);
    pl.p.tone = !hdr && (dt != 0 || d->exposure_ev != 0.0);
    pl.bytes_per_px = pl.mode == kRgba16 ? 8 : pl.mode == kFloat3 ? 12 : 4;
    return RISE_HIP_OK;
}

RISE_HD inline void encode_pixel(const out::Params& p, int mode, const float* px, void* out, int64_t i)
{
    double c[3];
    out::pixel_colour(p, px[0], px[1], px[2], c);
    if (mode == kRgba8) {
        uint8_t* o = (uint8_t*)out + 4 * i;
        for (int k = 0; k < 3; ++k) o[k] = out::integerize<uint8_t>(c[k], 255.0);
        o[3] = out::integerize<uint8_t>(1.0, 255.0);
    } else if (mode == kRgba16) {
        uint16_t* o = (uint16_t*)out + 4 * i;
        for (int k = 0; k < 3; ++k) o[k] = out::integerize<uint16_t>(c[k], 65535.0);
        o[3] = out::integerize<uint16_t>(1.0, 65535.0);
    } else if (mode == kRgbe) {
        out::rgbe(c, (uint8_t*)out + 4 * i);
    } else {
        float* o = (float*)out + 3 * i;
        for (int k = 0; k < 3; ++k) o[k] = (float)c[k];
    }
}

__global__ void k_output(const float* __restrict__ film, int64_t n, out::Params p, int mode, void* out)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        encode_pixel(p, mode, film + 3 * i, out, i);
}

// ---- framing --------------------------------------------------------------

struct Crc32Table {
    uint32_t t[256];
    Crc32Table()
    {
        for (uint32_t k = 0; k < 256; ++k) {
            uint32_t v = k;
            for (int j = 0; j < 8; ++j) v = (v & 1) ? 0xEDB88320u ^ (v >> 1) : v >> 1;
            t[k] = v;
        }
    }
};

// CRC-32 (IEEE 802.3, reflected), as PNG chunks carry it.
uint32_t crc32_update(uint32_t c, const uint8_t* p, size_t n)
{
    static const Crc32Table table;  // initialized once, thread-safely
    for (size_t i = 0; i < n; ++i) c = table.t[(c ^ p[i]) & 0xff] ^ (c >> 8);
    return c;
}

void put_be32(std::vector<uint8_t>& b, uint32_t v)
{
    for (int s = 24; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s));
}

void png_chunk(std::vector<uint8_t>& f, const char* type, const std::vector<uint8_t>& data)
{
    put_be32(f, (uint32_t)data.size());
    const size_t at = f.size();
    f.insert(f.end(), type, type + 4);
    f.insert(f.end(), data.begin(), data.end());
    const uint32_t crc = crc32_update(0xffffffffu, f.data() + at, 4 + data.size()) ^ 0xffffffffu;
    put_be32(f, crc);
}

// RGBA rows (filter 0) in a zlib stream of stored deflate blocks.
std::vector<uint8_t> png_file(const uint8_t* px, int w, int h, int bytes_per_sample)
{
    const size_t row = (size_t)w * 4 * bytes_per_sample;
    std::vector<uint8_t> raw;
    raw.reserve((row + 1) * h);
    for (int y = 0; y < h; ++y) {
        raw.push_back(0);
        const uint8_t* r = px + (size_t)y * row;
        if (bytes_per_sample == 1) {
            raw.insert(raw.end(), r, r + row);
        } else {
            for (size_t k = 0; k < row; k += 2) {  // 16-bit samples are big-endian
                raw.push_back(r[k + 1]);
                raw.push_back(r[k]);
            }
        }
    }
    std::vector<uint8_t> z{0x78, 0x01};
    uint32_t a = 1, b = 0;
    for (uint8_t v : raw) {
        a = (a + v) % 65521u;
        b = (b + a) % 65521u;
    }
    size_t off = 0;
    do {
        const size_t len = std::min<size_t>(65535, raw.size() - off);
        z.push_back(off + len == raw.size() ? 1 : 0);
        z.push_back((uint8_t)(len & 0xff));
        z.push_back((uint8_t)(len >> 8));
        z.push_back((uint8_t)(~len & 0xff));
        z.push_back((uint8_t)((~len >> 8) & 0xff));
        z.insert(z.end(), raw.begin() + (std::ptrdiff_t)off, raw.begin() + (std::ptrdiff_t)(off + len));
        off += len;
    } while (off < raw.size());
    put_be32(z, (b << 16) | a);
    std::vector<uint8_t> f{0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    std::vector<uint8_t> ihdr;
    put_be32(ihdr, (uint32_t)w);
    put_be32(ihdr, (uint32_t)h);
    ihdr.insert(ihdr.end(), {(uint8_t)(8 * bytes_per_sample), 6, 0, 0, 0});
    png_chunk(f, "IHDR", ihdr);
    png_chunk(f, "IDAT", z);
    png_chunk(f, "IEND", {});
    return f;
}

// IEEE half, round to nearest even (what OpenEXR's half(float) does).
uint16_t to_half(float v)
{
    uint32_t x;
    std::memcpy(&x, &v, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0u));  // inf / nan
    if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);  // rounds past the largest half
    if (ax < 0x38800000u) {                                     // subnormal half or zero
        if (ax < 0x33000000u) return (uint16_t)sign;
        const uint32_t m = (ax & 0x7fffffu) | 0x800000u;
        const int shift = 126 - (int)(ax >> 23);  // 14 .. 24
        const uint32_t half_m = m >> shift;
        const uint32_t rem = m & ((1u << shift) - 1), mid = 1u << (shift - 1);
        return (uint16_t)(sign | (half_m + (rem > mid || (rem == mid && (half_m & 1)))));
    }
    const uint32_t e = (ax >> 23) - 112;
    const uint32_t m = ax & 0x7fffffu;
    uint32_t h = (e << 10) | (m >> 13);
    const uint32_t rem = m & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1))) ++h;
    return (uint16_t)(sign | h);
}

void put_le(std::vector<uint8_t>& b, const void* p, size_t n)
{
    b.insert(b.end(), (const uint8_t*)p, (const uint8_t*)p + n);
}

void exr_attr(std::vector<uint8_t>& b, const char* name, const char* type, const std::vector<uint8_t>& v)
{
    b.insert(b.end(), name, name + std::strlen(name) + 1);
    b.insert(b.end(), type, type + std::strlen(type) + 1);
    const int32_t n = (int32_t)v.size();
    put_le(b, &n, 4);
    b.insert(b.end(), v.begin(), v.end());
}

std::vector<uint8_t> exr_file(const float* rgb, int w, int h, bool half, bool alpha)
{
    std::vector<uint8_t> f{0x76, 0x2f, 0x31, 0x01, 2, 0, 0, 0};
    const char* names[4] = {"A", "B", "G", "R"};  // channels in sorted order
    const int first = alpha ? 0 : 1;
    std::vector<uint8_t> ch;
    for (int c = first; c < 4; ++c) {
        ch.insert(ch.end(), names[c], names[c] + 2);
        const int32_t type = half ? 1 : 2, one = 1;
        put_le(ch, &type, 4);
        ch.insert(ch.end(), {0, 0, 0, 0});
        put_le(ch, &one, 4);
        put_le(ch, &one, 4);
    }
    ch.push_back(0);
    exr_attr(f, "channels", "chlist", ch);
    exr_attr(f, "compression", "compression", {0});
    std::vector<uint8_t> box;
    const int32_t bx[4] = {0, 0, w - 1, h - 1};
    put_le(box, bx, 16);
    exr_attr(f, "dataWindow", "box2i", box);
    exr_attr(f, "displayWindow", "box2i", box);
    exr_attr(f, "lineOrder", "lineOrder", {0});
    std::vector<uint8_t> v;
    const float one = 1.0f, zero[2] = {0.0f, 0.0f};
    put_le(v, &one, 4);
    exr_attr(f, "pixelAspectRatio", "float", v);
    v.clear();
    put_le(v, zero, 8);
    exr_attr(f, "screenWindowCenter", "v2f", v);
    v.clear();
    put_le(v, &one, 4);
    exr_attr(f, "screenWindowWidth", "float", v);
    f.push_back(0);
    const int nch = 4 - first;
    const size_t line = (size_t)w * nch * (half ? 2 : 4);
    uint64_t at = f.size() + 8 * (uint64_t)h;
    for (int y = 0; y < h; ++y, at += 8 + line) put_le(f, &at, 8);
    for (int y = 0; y < h; ++y) {
        const int32_t yy = y, sz = (int32_t)line;
        put_le(f, &yy, 4);
        put_le(f, &sz, 4);
        for (int c = first; c < 4; ++c)
            for (int x = 0; x < w; ++x) {
                const float val = c == 0 ? 1.0f : rgb[3 * ((size_t)y * w + x) + (3 - c)];  // B=2, G=1, R=0
                if (half) {
                    const uint16_t hv = to_half(val);
                    put_le(f, &hv, 2);
                } else {
                    put_le(f, &val, 4);
                }
            }
    }
    return f;
}

// The writers' framing of the per-pixel data (row y of the film is the
// reference's raster row y).
std::vector<uint8_t> frame(const rise_hip_output_desc& d, const Plan& pl, const std::vector<uint8_t>& px, int w,
                           int h)
{
    std::vector<uint8_t> f;
    const size_t n = (size_t)w * h;
    switch (d.format) {
    case RISE_HIP_IMAGE_PPM: {  // PPMWriter: "P6\n<w> <h>\n255\n", RGB
        char hdr[64];
        std::snprintf(hdr, sizeof(hdr), "P6\n%d %d\n255\n", w, h);
        f.insert(f.end(), hdr, hdr + std::strlen(hdr));
        for (size_t i = 0; i < n; ++i) f.insert(f.end(), &px[4 * i], &px[4 * i] + 3);
        break;
    }
    case RISE_HIP_IMAGE_TGA: {  // TGAWriter: 18-byte header, 32 bpp, descriptor 8 (bottom-left origin), BGRA
        const uint8_t hdr[18] = {0, 0, 2, 0, 0, 0, 0, 0, 0, 0, 0, 0, (uint8_t)(w & 0xff), (uint8_t)(w >> 8),
                                 (uint8_t)(h & 0xff), (uint8_t)(h >> 8), 32, 8};
        f.insert(f.end(), hdr, hdr + 18);
        for (int y = h - 1; y >= 0; --y)  // bottom row first
            for (int x = 0; x < w; ++x) {
                const uint8_t* p = &px[4 * ((size_t)y * w + x)];
                f.insert(f.end(), {p[2], p[1], p[0], p[3]});
            }
        break;
    }
    case RISE_HIP_IMAGE_PNG:
        f = png_file(px.data(), w, h, pl.mode == kRgba16 ? 2 : 1);
        break;
    case RISE_HIP_IMAGE_HDR: {  // HDRWriter: Radiance header, flat RGBE scanlines
        char hdr[256];
        std::snprintf(hdr, sizeof(hdr),
                      "#?RADIANCE\n# Generated by rise_hip\nFORMAT=32-bit_rle_rgbe\n\n-Y %d +X %d\n", h, w);
        f.insert(f.end(), hdr, hdr + std::strlen(hdr));
        f.insert(f.end(), px.begin(), px.end());
        break;
    }
    default:  // EXR
        f = exr_file((const float*)px.data(), w, h, d.bpp < 32, d.with_alpha != 0);
        break;
    }
    return f;
}

rise_hip_status write_file(const char* path, const std::vector<uint8_t>& f)
{
    std::FILE* fp = std::fopen(path, "wb");
    if (!fp) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, std::string("cannot create ") + path);
    const bool ok = std::fwrite(f.data(), 1, f.size(), fp) == f.size();
    if (std::fclose(fp) != 0 || !ok) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, std::string("cannot write ") + path);
    return RISE_HIP_OK;
}

}  // namespace

extern "C" rise_hip_status rise_hip_output_write(const float* film, int32_t w, int32_t h,
                                                 const rise_hip_output_desc* desc, const char* path, void* stream)
{
    clear_error();
    if (!film || !path) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL film or path");
    Plan pl{};
    rise_hip_status st = make_plan(desc, w, h, pl);
    if (st != RISE_HIP_OK) return st;
    const int64_t n = (int64_t)w * h;
    hipStream_t s = (hipStream_t)stream;
    void* d_out = nullptr;
    hipError_t e = hipMalloc(&d_out, (size_t)n * pl.bytes_per_px);
    if (e != hipSuccess) return fail_hip(e, "hipMalloc (output)");
    std::vector<uint8_t> px;
    try {
        px.resize((size_t)n * pl.bytes_per_px);
    } catch (const std::bad_alloc&) {
        (void)hipFree(d_out);
        return fail(RISE_HIP_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
    const int block = 256;
    const unsigned grid = (unsigned)std::min<int64_t>((n + block - 1) / block, 8192);
    hipLaunchKernelGGL(k_output, dim3(grid), dim3(block), 0, s, film, n, pl.p, pl.mode, d_out);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(px.data(), d_out, px.size(), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d_out);
    if (e != hipSuccess) return fail_hip(e, "output resolve");
    try {
        return write_file(path, frame(*desc, pl, px, w, h));
    } catch (const std::bad_alloc&) {
        return fail(RISE_HIP_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
}

extern "C" rise_hip_status rise_hip_output_resolve(const float* film, int32_t w, int32_t h,
                                                   const rise_hip_output_desc* desc, void* out, void* stream)
{
    clear_error();
    if (!film || !out) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL film or output");
    Plan pl{};
    rise_hip_status st = make_plan(desc, w, h, pl);
    if (st != RISE_HIP_OK) return st;
    const int64_t n = (int64_t)w * h;
    const int block = 256;
    const unsigned grid = (unsigned)std::min<int64_t>((n + block - 1) / block, 8192);
    hipLaunchKernelGGL(k_output, dim3(grid), dim3(block), 0, (hipStream_t)stream, film, n, pl.p, pl.mode, out);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? RISE_HIP_OK : fail_hip(e, "output resolve launch");
}

extern "C" rise_hip_status rise_hip_output_write_host(const float* film, int32_t w, int32_t h,
                                                      const rise_hip_output_desc* desc, const char* path)
{
    clear_error();
    if (!film || !path) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL film or path");
    Plan pl{};
    rise_hip_status st = make_plan(desc, w, h, pl);
    if (st != RISE_HIP_OK) return st;
    const int64_t n = (int64_t)w * h;
    try {
        std::vector<uint8_t> px((size_t)n * pl.bytes_per_px);
        const int nt = host_threads();
        std::vector<std::thread> th;
        for (int t = 0; t < nt; ++t)
            th.emplace_back([&, t] {
                for (int64_t i = t; i < n; i += nt) encode_pixel(pl.p, pl.mode, film + 3 * i, px.data(), i);
            });
        for (auto& x : th) x.join();
        return write_file(path, frame(*desc, pl, px, w, h));
    } catch (const std::bad_alloc&) {
        return fail(RISE_HIP_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
}

extern "C" const char* rise_hip_output_extension(int32_t format)
{
    switch (format) {
    case RISE_HIP_IMAGE_TGA: return "tga";
    case RISE_HIP_IMAGE_PPM: return "ppm";
    case RISE_HIP_IMAGE_PNG: return "png";
    case RISE_HIP_IMAGE_HDR: return "hdr";
    case RISE_HIP_IMAGE_EXR: return "exr";
    default: return "out";
    }
}
