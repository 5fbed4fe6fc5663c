// risemesh.cpp -- RISE's binary mesh files (.risemesh v1-v5) and their BVH2
// cache block, read and written on the host.
//
// Layout (TriangleMeshGeometryIndexed.cpp:586-1000; little-endian scalars as
// MemoryBuffer stores them, Utilities/MemoryBuffer.cpp:398-470, 605-660):
//   "RISETMGI" | uint32 version
//   v1-v3: uint32 (legacy max per octant node) | char (legacy recursion level)
//   char face_normals
//   uint32 n_points | n_points x 3 doubles
//   uint32 n_normals | n_normals x 3 doubles
//   uint32 n_coords | n_coords x 2 doubles
//   v5: uint32 n_colors | n_colors x 3 doubles
//   uint32 n_tris | n_tris x 3 x (uint32 point, uint32 normal, uint32 coord)
//   char double_sided
//   v1-v3: char use_bsp | char have_tree | [serialized BSP/octree]
//   v3+ (n_tris > 0): char have_bvh | [BVH2 block]
// BVH2 block (Acceleration/BVH.h:338-423): "BVH2" | uint32 1 | uint32 n_nodes |
// n_nodes x 32-byte nodes | uint32 n_prims | n_prims x uint32 | 6 doubles (box).
#include "rise_hip_risemesh.h"

#include "api_common.h"
#include "bvh_build.h"
#include "risemesh_internal.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <unordered_map>
#include <new>
#include <string>
#include <tuple>
#include <vector>

using namespace rise_hip;

struct rise_hip_risemesh {
    int32_t version = 0;
    bool face_normals = false;
    bool double_sided = false;
    std::vector<double> points;   // 3 per point
    std::vector<double> normals;  // 3 per normal
    std::vector<double> coords;   // 2 per coordinate
    std::vector<double> colors;   // 3 per color
    std::vector<uint32_t> tris;   // 9 per triangle: (point, normal, coord) per corner
    bool has_bvh = false;
    bool bvh_valid = false;
    Bvh2Tree bvh;
    std::string bvh_error;
};

namespace {

constexpr char kSignature[8] = {'R', 'I', 'S', 'E', 'T', 'M', 'G', 'I'};
constexpr uint32_t kWriteVersion = 5;

struct Reader {
    const std::vector<char>& buf;
    size_t pos = 0;
    std::string err;

    bool need(size_t n)
    {
        if (pos > buf.size() || n > buf.size() - pos) {
            if (err.empty()) err = "file truncated at byte " + std::to_string(pos);
            return false;
        }
        return true;
    }
    bool bytes(void* dst, size_t n)
    {
        if (!need(n)) return false;
        std::memcpy(dst, buf.data() + pos, n);
        pos += n;
        return true;
    }
    bool u32(uint32_t& v) { return bytes(&v, 4); }
    bool u8(uint8_t& v) { return bytes(&v, 1); }
    bool f64(double& v) { return bytes(&v, 8); }
    // n items of `per` doubles, with the count checked against the bytes left.
    bool doubles(uint32_t n, int per, std::vector<double>& out)
    {
        if (!need((size_t)n * per * 8)) return false;
        out.resize((size_t)n * per);
        return bytes(out.data(), out.size() * 8);
    }
};

struct Writer {
    std::vector<char> buf;
    void bytes(const void* p, size_t n) { buf.insert(buf.end(), (const char*)p, (const char*)p + n); }
    void u32(uint32_t v) { bytes(&v, 4); }
    void u8(uint8_t v) { bytes(&v, 1); }
    void f64(double v) { bytes(&v, 8); }
};

bool read_file(const std::string& path, std::vector<char>& out, std::string& err)
{
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    const std::streamoff size = f.tellg();
    if (size < 0) {
        err = "cannot read " + path;
        return false;
    }
    out.resize((size_t)size);
    f.seekg(0);
    if (size > 0 && !f.read(out.data(), size)) {
        err = "cannot read " + path;
        return false;
    }
    return true;
}

bool parse(const std::vector<char>& data, rise_hip_risemesh& m, std::string& err)
{
    Reader r{data};
    char sig[8];
    uint32_t version = 0;
    if (!r.bytes(sig, 8) || std::memcmp(sig, kSignature, 8) != 0) {
        err = "not a .risemesh file (signature RISETMGI not found)";
        return false;
    }
    if (!r.u32(version) || version < 1 || version > 5) {
        err = "unsupported .risemesh version " + std::to_string(version) + " (1..5 are known)";
        return false;
    }
    m.version = (int32_t)version;
    uint8_t c = 0;
    uint32_t legacy = 0;
    if (version < 4 && (!r.u32(legacy) || !r.u8(c))) {
        err = r.err;
        return false;
    }
    if (!r.u8(c)) {
        err = r.err;
        return false;
    }
    m.face_normals = c != 0;
    uint32_t n = 0;
    if (!r.u32(n) || !r.doubles(n, 3, m.points) || !r.u32(n) || !r.doubles(n, 3, m.normals) || !r.u32(n) ||
        !r.doubles(n, 2, m.coords)) {
        err = r.err;
        return false;
    }
    if (version >= 5 && (!r.u32(n) || !r.doubles(n, 3, m.colors))) {
        err = r.err;
        return false;
    }
    uint32_t nt = 0;
    if (!r.u32(nt) || !r.need((size_t)nt * 36)) {
        err = r.err;
        return false;
    }
    m.tris.resize((size_t)nt * 9);
    if (!r.bytes(m.tris.data(), m.tris.size() * 4)) {
        err = r.err;
        return false;
    }
    const size_t np = m.points.size() / 3, nn = m.normals.size() / 3, nc = m.coords.size() / 2;
    for (size_t t = 0; t < nt; ++t)
        for (int k = 0; k < 3; ++k) {
            const uint32_t* e = &m.tris[9 * t + 3 * k];
            // The reference takes &pPoints[i] / &pCoords[i] unchecked and the
            // normal only with per-vertex normals; out-of-range indices are
            // refused here.
            if (e[0] >= np || (nc > 0 && e[2] >= nc) || (!m.face_normals && nn > 0 && e[1] >= nn)) {
                err = "triangle " + std::to_string(t) + " indexes past the point, normal or coordinate lists";
                return false;
            }
        }
    if (!r.u8(c)) {
        err = r.err;
        return false;
    }
    m.double_sided = c != 0;
    if (version < 4) {
        uint8_t use_bsp = 0, have_tree = 0;
        if (!r.u8(use_bsp) || !r.u8(have_tree)) {
            err = r.err;
            return false;
        }
        if (have_tree && version >= 2) {
            err = "legacy BSP/octree block (version " + std::to_string(version) +
                  ") is not supported; re-save the mesh as version 4 or 5";
            return false;
        }
    }
    if (version >= 3 && nt > 0) {
        uint8_t have = 0;
        if (!r.u8(have)) {
            err = r.err;
            return false;
        }
        if (have) {
            m.has_bvh = true;
            // A bad cache is not fatal: the reference logs and rebuilds
            // (TriangleMeshGeometryIndexed.cpp:935-944).
            char magic[4];
            uint32_t ver = 0, nodes = 0, prims = 0;
            std::string why;
            if (!r.bytes(magic, 4) || std::memcmp(magic, "BVH2", 4) != 0 || !r.u32(ver) || ver != 1) {
                why = "bad BVH2 magic or version";
            } else if (!r.u32(nodes) || !r.need((size_t)nodes * 32)) {
                why = "BVH2 node block truncated";
            } else {
                m.bvh.nodes.resize(nodes);
                r.bytes(m.bvh.nodes.data(), (size_t)nodes * 32);
                if (!r.u32(prims) || !r.need((size_t)prims * 4)) {
                    why = "BVH2 primitive block truncated";
                } else {
                    m.bvh.prims.resize(prims);
                    r.bytes(m.bvh.prims.data(), (size_t)prims * 4);
                    double box[6];
                    for (double& v : box)
                        if (!r.f64(v)) why = "BVH2 box truncated";
                }
            }
            if (why.empty()) {
                std::string e2;
                if (check_bvh2(m.bvh, nt, e2)) m.bvh_valid = true;
                else why = e2;
            }
            if (!m.bvh_valid) {
                m.bvh_error = why;
                m.bvh = Bvh2Tree();
            }
        }
    }
    return true;
}

std::vector<float> triangles_of(const rise_hip_risemesh& m)
{
    const size_t nt = m.tris.size() / 9;
    std::vector<float> out(nt * 9);
    for (size_t t = 0; t < nt; ++t)
        for (int k = 0; k < 3; ++k) {
            const double* p = &m.points[3 * (size_t)m.tris[9 * t + 3 * k]];
            for (int j = 0; j < 3; ++j) out[9 * t + 3 * k + j] = (float)p[j];
        }
    return out;
}

}  // namespace

namespace rise_hip {

const Bvh2Tree* risemesh_tree(const rise_hip_risemesh* m) { return m->bvh_valid ? &m->bvh : nullptr; }

std::vector<double> risemesh_triangles_f64(const rise_hip_risemesh* m)
{
    const size_t nt = m->tris.size() / 9;
    std::vector<double> out(nt * 9);
    for (size_t t = 0; t < nt; ++t)
        for (int k = 0; k < 3; ++k)
            for (int j = 0; j < 3; ++j) out[9 * t + 3 * k + j] = m->points[3 * (size_t)m->tris[9 * t + 3 * k] + j];
    return out;
}

bool risemesh_double_sided(const rise_hip_risemesh* m) { return m->double_sided; }

}  // namespace rise_hip

extern "C" rise_hip_status rise_hip_risemesh_read(const char* path, rise_hip_risemesh** out_mesh)
{
    clear_error();
    if (!path || !out_mesh) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    *out_mesh = nullptr;
    rise_hip_risemesh* m = new (std::nothrow) rise_hip_risemesh;
    if (!m) return fail(RISE_HIP_ERR_OUT_OF_MEMORY, "host allocation failed");
    std::string err;
    try {
        std::vector<char> data;
        if (!read_file(path, data, err) || !parse(data, *m, err)) {
            delete m;
            return fail(RISE_HIP_ERR_INVALID_ARGUMENT, std::string(path) + ": " + err);
        }
    } catch (const std::bad_alloc&) {
        delete m;
        return fail(RISE_HIP_ERR_OUT_OF_MEMORY, "host allocation failed while reading a .risemesh file");
    }
    *out_mesh = m;
    return RISE_HIP_OK;
}

extern "C" rise_hip_status rise_hip_risemesh_get_info(const rise_hip_risemesh* m, rise_hip_risemesh_info* out)
{
    clear_error();
    if (!m || !out) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = rise_hip_risemesh_info{};
    out->version = m->version;
    out->face_normals = m->face_normals ? 1 : 0;
    out->double_sided = m->double_sided ? 1 : 0;
    out->has_bvh_cache = m->has_bvh ? 1 : 0;
    out->bvh_cache_valid = m->bvh_valid ? 1 : 0;
    out->n_points = (int64_t)(m->points.size() / 3);
    out->n_normals = (int64_t)(m->normals.size() / 3);
    out->n_coords = (int64_t)(m->coords.size() / 2);
    out->n_colors = (int64_t)(m->colors.size() / 3);
    out->n_tris = (int64_t)(m->tris.size() / 9);
    out->n_bvh_nodes = (int64_t)m->bvh.nodes.size();
    return RISE_HIP_OK;
}

extern "C" rise_hip_status rise_hip_risemesh_get_triangles(const rise_hip_risemesh* m, float* out_tris)
{
    clear_error();
    if (!m || (!out_tris && !m->tris.empty())) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    const std::vector<float> t = triangles_of(*m);
    if (!t.empty()) std::memcpy(out_tris, t.data(), t.size() * sizeof(float));
    return RISE_HIP_OK;
}

extern "C" rise_hip_status rise_hip_risemesh_write(const char* path, const float* tris, int64_t n_tris,
                                                   const float* uvs, int32_t double_sided, int32_t with_bvh_cache)
{
    clear_error();
    if (!path || n_tris < 0 || (n_tris > 0 && !tris)) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "bad argument");
    if (n_tris > 0x7fffffff) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "more than 2^31-1 triangles");
    try {
        for (int64_t i = 0; i < 9 * n_tris; ++i)
            if (!std::isfinite(tris[i])) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "non-finite vertex coordinate");
        // Exact-match vertex and coordinate deduplication (by bit pattern).
        struct Key3Hash {
            size_t operator()(const std::tuple<uint32_t, uint32_t, uint32_t>& k) const
            {
                uint64_t h = (uint64_t)std::get<0>(k) * 0x9E3779B97F4A7C15ull;
                h ^= (uint64_t)std::get<1>(k) * 0xC2B2AE3D27D4EB4Full + (h << 6) + (h >> 2);
                h ^= (uint64_t)std::get<2>(k) * 0x165667B19E3779F9ull + (h << 6) + (h >> 2);
                return (size_t)h;
            }
        };
        struct Key2Hash {
            size_t operator()(const std::pair<uint32_t, uint32_t>& k) const
            {
                return (size_t)(((uint64_t)k.first << 32 | k.second) * 0x9E3779B97F4A7C15ull);
            }
        };
        auto bits = [](float f) {
            uint32_t u;
            std::memcpy(&u, &f, 4);
            return f == 0.0f ? 0u : u;  // +0 and -0 are one vertex
        };
        std::unordered_map<std::tuple<uint32_t, uint32_t, uint32_t>, uint32_t, Key3Hash> pmap;
        std::unordered_map<std::pair<uint32_t, uint32_t>, uint32_t, Key2Hash> cmap;
        pmap.reserve((size_t)n_tris * 3);
        std::vector<double> points, coords;
        std::vector<uint32_t> idx((size_t)n_tris * 9, 0);
        for (int64_t t = 0; t < n_tris; ++t)
            for (int k = 0; k < 3; ++k) {
                const float* v = tris + 9 * t + 3 * k;
                auto key = std::make_tuple(bits(v[0]), bits(v[1]), bits(v[2]));
                auto it = pmap.find(key);
                if (it == pmap.end()) {
                    it = pmap.emplace(key, (uint32_t)(points.size() / 3)).first;
                    points.insert(points.end(), {(double)v[0], (double)v[1], (double)v[2]});
                }
                idx[9 * t + 3 * k] = it->second;
                if (uvs) {
                    const float* c = uvs + 6 * t + 2 * k;
                    auto ck = std::make_pair(bits(c[0]), bits(c[1]));
                    auto ci = cmap.find(ck);
                    if (ci == cmap.end()) {
                        ci = cmap.emplace(ck, (uint32_t)(coords.size() / 2)).first;
                        coords.insert(coords.end(), {(double)c[0], (double)c[1]});
                    }
                    idx[9 * t + 3 * k + 2] = ci->second;
                }
            }
        if (!uvs && n_tris > 0) coords = {0.0, 0.0};  // one shared (0, 0): every corner indexes a coordinate
        Writer w;
        w.bytes(kSignature, 8);
        w.u32(kWriteVersion);
        w.u8(1);  // face normals (no per-vertex normals are stored)
        w.u32((uint32_t)(points.size() / 3));
        w.bytes(points.data(), points.size() * 8);
        w.u32(0);  // normals
        w.u32((uint32_t)(coords.size() / 2));
        w.bytes(coords.data(), coords.size() * 8);
        w.u32(0);  // colors
        w.u32((uint32_t)n_tris);
        w.bytes(idx.data(), idx.size() * 4);
        w.u8(double_sided ? 1 : 0);
        if (n_tris > 0) {
            if (with_bvh_cache) {
                Bvh2Tree tree;
                float lo[3], hi[3];
                std::string err;
                if (!build_bvh2(tris, n_tris, BuildParams(), tree, lo, hi, err))
                    return fail(RISE_HIP_ERR_BUILD, "BVH2 build failed: " + err);
                w.u8(1);
                w.bytes("BVH2", 4);
                w.u32(1);
                w.u32((uint32_t)tree.nodes.size());
                w.bytes(tree.nodes.data(), tree.nodes.size() * sizeof(Bvh2Node));
                w.u32((uint32_t)tree.prims.size());
                w.bytes(tree.prims.data(), tree.prims.size() * 4);
                for (int k = 0; k < 3; ++k) w.f64(lo[k]);
                for (int k = 0; k < 3; ++k) w.f64(hi[k]);
            } else {
                w.u8(0);
            }
        }
        std::FILE* f = std::fopen(path, "wb");
        if (!f) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, std::string("cannot create ") + path);
        const bool ok = std::fwrite(w.buf.data(), 1, w.buf.size(), f) == w.buf.size();
        if (std::fclose(f) != 0 || !ok) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, std::string("cannot write ") + path);
    } catch (const std::bad_alloc&) {
        return fail(RISE_HIP_ERR_OUT_OF_MEMORY, "host allocation failed while writing a .risemesh file");
    }
    return RISE_HIP_OK;
}

extern "C" rise_hip_status rise_hip_risemesh_upload(const rise_hip_risemesh* m, rise_hip_scene** out_scene)
{
    clear_error();
    if (!m || !out_scene) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    *out_scene = nullptr;
    WideBVH bvh;
    std::string err;
    const int64_t n = (int64_t)(m->tris.size() / 9);
    try {
        const std::vector<float> t = triangles_of(*m);
        const std::vector<Bvh2Part> parts{{0, n, m->bvh_valid ? &m->bvh : nullptr}};
        if (!build_wide_bvh_parts(t.data(), n, nullptr, parts, BuildParams(), bvh, err))
            return fail(RISE_HIP_ERR_BUILD, "BVH build failed: " + err);
    } catch (const std::bad_alloc&) {
        return fail(RISE_HIP_ERR_OUT_OF_MEMORY, "host allocation failed during BVH build");
    }
    return upload_scene(bvh, n, bvh.stats.imported_tris > 0 ? 1 : 0, out_scene);
}

extern "C" rise_hip_status rise_hip_risemesh_release(rise_hip_risemesh* m)
{
    clear_error();
    delete m;
    return RISE_HIP_OK;
}
