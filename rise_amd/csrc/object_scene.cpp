// object_scene.cpp -- builds the trees of a scene of objects with transforms.
#include "object_scene.h"

#include "api_common.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>

namespace rise_hip {

namespace {

bool is_identity(const double* m)
{
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            if (m[4 * i + j] != (i == j ? 1.0 : 0.0)) return false;
    return true;
}

bool all_finite(const double* m)
{
    for (int k = 0; k < 16; ++k)
        if (!std::isfinite(m[k])) return false;
    return true;
}

// Point3Ops::Transform in double (row-vector convention, translation in row 3).
void xform(const double* m, const double p[3], double out[3])
{
    for (int j = 0; j < 3; ++j) out[j] = m[j] * p[0] + m[4 + j] * p[1] + m[8 + j] * p[2] + m[12 + j];
}

// Rounding bound of a world position taken to object space and back:
// 2^-40 times 4 ||L||_inf ||L^-1||_inf of the linear parts (the two matrices'
// upper 3x3), far above the few ulps of the two transforms.
double transform_slack(const double* m, const double* mi)
{
    double n = 0.0, ni = 0.0;
    for (int i = 0; i < 3; ++i) {
        n = std::max(n, std::fabs(m[4 * i]) + std::fabs(m[4 * i + 1]) + std::fabs(m[4 * i + 2]));
        ni = std::max(ni, std::fabs(mi[4 * i]) + std::fabs(mi[4 * i + 1]) + std::fabs(mi[4 * i + 2]));
    }
    return 4.0 * n * ni * 0x1p-40;
}

// World box of an object: the mesh's vertex bounds in object space, widened by
// 1e-7 of their size and coordinates plus 1e-10 (the hit point the reference
// transforms back lies within the certifier's 1e-12 tolerance and the 1e-12
// back-off of the triangle), its eight corners taken to world space, and the
// result widened again by the larger of 1e-7 and the transform's slack.  It
// contains every world point the reference can report for the object, and
// every point of the exact world ray that passes within rounding of one.
void world_box(const std::vector<double>& v64, const double* to_world, double slack, double lo[3], double hi[3])
{
    double olo[3] = {INFINITY, INFINITY, INFINITY}, ohi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = 0; i < v64.size() / 3; ++i)
        for (int a = 0; a < 3; ++a) {
            const double v = v64[3 * i + a];
            olo[a] = std::min(olo[a], v);
            ohi[a] = std::max(ohi[a], v);
        }
    double span = 0.0;
    for (int a = 0; a < 3; ++a) span = std::max({span, ohi[a] - olo[a], std::fabs(olo[a]), std::fabs(ohi[a])});
    const double pad_o = span * 1e-7 + 1e-10;
    for (int a = 0; a < 3; ++a) {
        olo[a] -= pad_o;
        ohi[a] += pad_o;
        lo[a] = INFINITY;
        hi[a] = -INFINITY;
    }
    for (int c = 0; c < 8; ++c) {
        const double p[3] = {(c & 1) ? ohi[0] : olo[0], (c & 2) ? ohi[1] : olo[1], (c & 4) ? ohi[2] : olo[2]};
        double w[3];
        xform(to_world, p, w);
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], w[a]);
            hi[a] = std::max(hi[a], w[a]);
        }
    }
    double wspan = 0.0;
    for (int a = 0; a < 3; ++a) wspan = std::max({wspan, hi[a] - lo[a], std::fabs(lo[a]), std::fabs(hi[a])});
    const double pad_w = wspan * std::max(1e-7, slack) + 1e-10;
    for (int a = 0; a < 3; ++a) {
        lo[a] -= pad_w;
        hi[a] += pad_w;
    }
}

}  // namespace

void mesh_vertices(const rise_hip_mesh_desc& m, std::vector<float>& f, std::vector<double>& d)
{
    const size_t n = (size_t)m.n_tris * 9;
    const size_t f0 = f.size(), d0 = d.size();
    f.resize(f0 + n);
    d.resize(d0 + n);
    for (size_t i = 0; i < n; ++i) {
        if (m.vertices_f64) {
            d[d0 + i] = m.vertices_f64[i];
            f[f0 + i] = (float)m.vertices_f64[i];
        } else {
            f[f0 + i] = m.vertices[i];
            d[d0 + i] = m.vertices[i];
        }
    }
}

std::vector<double> leaf_rows(const WideBVH& bvh, const std::vector<double>& rows, int stride)
{
    std::vector<double> out(bvh.tris.size() * (size_t)stride);
    for (size_t i = 0; i < bvh.tris.size(); ++i) {
        int32_t p;
        std::memcpy(&p, &bvh.tris[i].v0[3], 4);
        for (int k = 0; k < stride; ++k) out[(size_t)stride * i + k] = rows[(size_t)stride * (size_t)p + k];
    }
    return out;
}

std::vector<double> leaf_vertices(const WideBVH& bvh, const std::vector<double>& v64)
{
    return leaf_rows(bvh, v64, 9);
}

void mesh_attrs(const rise_hip_mesh_desc& m, std::vector<double>& a)
{
    const size_t a0 = a.size();
    a.resize(a0 + (size_t)m.n_tris * kAttrDoubles, 0.0);
    for (int64_t i = 0; i < m.n_tris; ++i) {
        double* row = a.data() + a0 + (size_t)i * kAttrDoubles;
        if (m.normals)
            for (int k = 0; k < 9; ++k) row[k] = m.normals[9 * i + k];
        if (m.coords)
            for (int k = 0; k < 6; ++k) row[9 + k] = m.coords[6 * i + k];
    }
}

uint32_t mesh_tri_flags(const rise_hip_mesh_desc& m)
{
    return ((m.flags & RISE_HIP_MESH_DOUBLE_SIDED) ? kTriDoubleSided : 0u) | (m.normals ? kTriNormals : 0u) |
           (m.coords ? kTriCoords : 0u);
}

rise_hip_status build_object_scene(const rise_hip_mesh_desc* meshes, size_t n_meshes,
                                   const rise_hip_object_instance* objects, size_t n_objects, ObjectSceneHost& out)
{
    if (n_meshes > 0 && !meshes) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "meshes is NULL");
    if (n_objects > 0 && !objects) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "objects is NULL");
    if (n_objects > 0x7fffffff) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "more than 2^31-1 objects");
    for (size_t m = 0; m < n_meshes; ++m) {
        if (meshes[m].n_tris < 0) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "negative triangle count");
        if (meshes[m].n_tris > 0x7fffffff) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "mesh over 2^31-1 triangles");
        if (meshes[m].n_tris > 0 && !meshes[m].vertices && !meshes[m].vertices_f64)
            return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "mesh vertices is NULL");
        if (meshes[m].flags & ~RISE_HIP_MESH_DOUBLE_SIDED)
            return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "unknown mesh flags");
    }
    int64_t world_total = 0;
    for (size_t k = 0; k < n_objects; ++k) {
        const rise_hip_object_instance& ob = objects[k];
        if (ob.mesh < 0 || (size_t)ob.mesh >= n_meshes)
            return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "object " + std::to_string(k) + ": mesh index out of range");
        if (ob.reserved != 0)
            return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "object " + std::to_string(k) + ": reserved must be 0");
        if (!all_finite(ob.to_world) || !all_finite(ob.to_object))
            return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "object " + std::to_string(k) + ": non-finite matrix entry");
        if (is_identity(ob.to_world) && is_identity(ob.to_object)) world_total += meshes[ob.mesh].n_tris;
    }
    if (world_total > 0x7fffffff) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "more than 2^31-1 world triangles");

    std::string err;
    try {
        // Identity objects: one world-space tree built over the running index,
        // then each leaf triangle takes its caller id (v0.w) and its object (v1.w).
        std::vector<float> verts;
        std::vector<double> verts64, attrs;
        bool world_has_attr = false;
        std::vector<int32_t> ids((size_t)world_total), obj_of((size_t)world_total);
        int64_t off = 0;
        for (size_t k = 0; k < n_objects; ++k) {
            const rise_hip_object_instance& ob = objects[k];
            if (!(is_identity(ob.to_world) && is_identity(ob.to_object))) continue;
            const rise_hip_mesh_desc& m = meshes[ob.mesh];
            if (m.n_tris == 0) continue;
            mesh_vertices(m, verts, verts64);
            mesh_attrs(m, attrs);
            world_has_attr = world_has_attr || m.normals || m.coords;
            for (int64_t i = 0; i < m.n_tris; ++i) {
                ids[(size_t)(off + i)] = m.tri_ids ? m.tri_ids[i] : (int32_t)i;
                obj_of[(size_t)(off + i)] = (int32_t)k;
            }
            off += m.n_tris;
        }
        out.world_tris = world_total;
        if (!build_wide_bvh(verts.data(), world_total, nullptr, BuildParams(), out.world, err))
            return fail(RISE_HIP_ERR_BUILD, "BVH build failed: " + err);
        out.world_tri64 = leaf_vertices(out.world, verts64);
        if (world_has_attr) out.world_attr = leaf_rows(out.world, attrs, kAttrDoubles);
        for (Tri48& tr : out.world.tris) {
            int32_t p;
            std::memcpy(&p, &tr.v0[3], 4);
            std::memcpy(&tr.v0[3], &ids[(size_t)p], 4);
            const int32_t k = obj_of[(size_t)p];
            std::memcpy(&tr.v1[3], &k, 4);
            const uint32_t fl = mesh_tri_flags(meshes[objects[k].mesh]);
            std::memcpy(&tr.v2[3], &fl, 4);
        }
        // Transformed objects: one object-space tree per mesh they use.
        std::vector<int32_t> blas_of(n_meshes, -1);
        out.total_tris = world_total;
        for (size_t k = 0; k < n_objects; ++k) {
            const rise_hip_object_instance& ob = objects[k];
            if (is_identity(ob.to_world) && is_identity(ob.to_object)) continue;
            const rise_hip_mesh_desc& m = meshes[ob.mesh];
            if (m.n_tris == 0) continue;  // nothing to hit
            if (blas_of[ob.mesh] < 0) {
                WideBVH b;
                std::vector<float> mv;
                std::vector<double> mv64, ma;
                mesh_vertices(m, mv, mv64);
                if (!build_wide_bvh(mv.data(), m.n_tris, nullptr, BuildParams(), b, err))
                    return fail(RISE_HIP_ERR_BUILD, "BVH build failed: " + err);
                out.blas_tri64.push_back(leaf_vertices(b, mv64));
                if (m.normals || m.coords) {
                    mesh_attrs(m, ma);
                    out.blas_attr.push_back(leaf_rows(b, ma, kAttrDoubles));
                } else {
                    out.blas_attr.emplace_back();
                }
                const uint32_t fl = mesh_tri_flags(m);
                for (Tri48& tr : b.tris) {
                    int32_t p;
                    std::memcpy(&p, &tr.v0[3], 4);
                    const int32_t id = m.tri_ids ? m.tri_ids[p] : p;
                    std::memcpy(&tr.v0[3], &id, 4);
                    std::memcpy(&tr.v1[3], &ob.mesh, 4);
                    std::memcpy(&tr.v2[3], &fl, 4);
                }
                blas_of[ob.mesh] = (int32_t)out.blas.size();
                out.total_tris += m.n_tris;
                out.blas.push_back(std::move(b));
            }
            DevObject d{};
            std::memcpy(d.to_world, ob.to_world, sizeof(d.to_world));
            std::memcpy(d.to_object, ob.to_object, sizeof(d.to_object));
            d.slack = transform_slack(ob.to_world, ob.to_object);
            world_box(out.blas_tri64[(size_t)blas_of[ob.mesh]], ob.to_world, d.slack, d.lo, d.hi);
            d.blas = blas_of[ob.mesh];
            d.object = (int32_t)k;
            out.objects.push_back(d);
        }
        // Top-level tree: each object's world box as the triangle (lo, hi,
        // (lo.x, lo.y, hi.z)), whose bounds are the box, rounded outward to
        // float; the leaf record's id is the object's position in `objects`.
        const size_t no = out.objects.size();
        std::vector<float> boxes(no * 9);
        for (size_t k = 0; k < no; ++k) {
            float lo[3], hi[3];
            for (int a = 0; a < 3; ++a) {
                lo[a] = (float)out.objects[k].lo[a];
                if ((double)lo[a] > out.objects[k].lo[a]) lo[a] = std::nextafter(lo[a], -INFINITY);
                hi[a] = (float)out.objects[k].hi[a];
                if ((double)hi[a] < out.objects[k].hi[a]) hi[a] = std::nextafter(hi[a], INFINITY);
            }
            const float v[9] = {lo[0], lo[1], lo[2], hi[0], hi[1], hi[2], lo[0], lo[1], hi[2]};
            std::memcpy(&boxes[k * 9], v, sizeof(v));
        }
        if (!build_wide_bvh(boxes.data(), (int64_t)no, nullptr, BuildParams(), out.top, err))
            return fail(RISE_HIP_ERR_BUILD, "top-level BVH build failed: " + err);
        int blas_depth = 0;
        for (const WideBVH& b : out.blas) blas_depth = std::max(blas_depth, b.stats.depth);
        out.stack_need = std::max(out.world.stats.depth, no > 0 ? out.top.stats.depth + blas_depth : 0);
        if (out.stack_need > kObjStack)
            return fail(RISE_HIP_ERR_BUILD, "top-level and object trees too deep for the traversal stack");
    } catch (const std::bad_alloc&) {
        return fail(RISE_HIP_ERR_OUT_OF_MEMORY, "host allocation failed during BVH build");
    }
    return RISE_HIP_OK;
}

}  // namespace rise_hip
