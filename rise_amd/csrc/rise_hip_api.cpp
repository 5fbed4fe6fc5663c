// rise_hip_api.cpp -- C ABI: scene upload and batch ray queries.
#include "rise_hip.h"
#include "rise_hip_risemesh.h"

#include "api_common.h"
#include "bvh_build.h"
#include "object_kernels.h"
#include "object_scene.h"
#include "trace_kernels.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <thread>
#include <chrono>
#include <cstdio>
#include <fstream>
#include <cstring>
#include <new>
#include <string>
#include <vector>

using namespace rise_hip;

struct rise_hip_scene {
    NodeSlot* d_nodes = nullptr;
    Tri48* d_tris = nullptr;
    double* d_tri64 = nullptr;  // leaf records' double vertices (9 each), or NULL: the float vertices are exact
    double* d_attr = nullptr;   // leaf records' normals and texture coordinates (kAttrDoubles each), or NULL
    // Dynamic-fetch counters: each query takes the next slot, so queries on
    // different streams do not share a counter (up to kFetchSlots in flight).
    static constexpr int kFetchSlots = 64;
    unsigned long long* d_fetch = nullptr;
    std::atomic<unsigned> next_slot{0};
    std::atomic<int> arith{RISE_HIP_ARITH_IEEE};
    int device = 0;
    rise_hip_scene_stats stats;
    // Transformed objects (rise_hip_scene_upload_objects); none otherwise.
    DevObject* d_objects = nullptr;
    BlasView* d_blas = nullptr;
    NodeSlot* d_top_nodes = nullptr;
    Tri48* d_top_leaves = nullptr;
    std::vector<void*> blas_allocs;
    int32_t n_objects = 0;
    int32_t stack_need = 0;  // node-group stack entries an object query needs
    bool world_empty = false;
    ObjectSceneView objects_view() const
    {
        return ObjectSceneView{world_empty ? nullptr : d_nodes, d_tris, d_tri64, d_attr, d_objects, n_objects,
                               d_blas, d_top_nodes, d_top_leaves};
    }
    unsigned long long* fetch_slot()
    {
        return d_fetch + next_slot.fetch_add(1u) % kFetchSlots;
    }
    TraceScene view() { return view_on(fetch_slot()); }
    TraceScene view_on(unsigned long long* fetch_counter) const
    {
        return TraceScene{d_nodes, d_tris, stats.depth, fetch_counter, arith.load(), d_tri64, d_attr};
    }
};

namespace rise_hip {

static thread_local std::string g_last_error;

rise_hip_status fail(rise_hip_status s, const std::string& msg)
{
    g_last_error = msg;
    return s;
}

rise_hip_status fail_hip(hipError_t e, const char* what)
{
    return fail(e == hipErrorOutOfMemory ? RISE_HIP_ERR_OUT_OF_MEMORY : RISE_HIP_ERR_DEVICE,
                std::string(what) + ": " + hipGetErrorString(e));
}

void clear_error() { g_last_error.clear(); }

int host_threads()
{
    int n = (int)std::thread::hardware_concurrency();
    if (const char* e = std::getenv("OMP_NUM_THREADS")) {
        const int v = std::atoi(e);
        if (v > 0) n = v;
    }
    return std::max(1, std::min(n, 16));
}

}  // namespace rise_hip

extern "C" {

const char* rise_hip_last_error(void) { return g_last_error.c_str(); }

const char* rise_hip_version(void) { return "0.1.0"; }

rise_hip_status rise_hip_scene_upload(const rise_hip_mesh_desc* meshes, size_t n_meshes, rise_hip_scene** out_scene)
{
    clear_error();
    if (!out_scene) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "out_scene is NULL");
    *out_scene = nullptr;
    if (n_meshes > 0 && !meshes) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "meshes is NULL");
    int64_t total = 0;
    bool any64 = false, any_attr = false;
    for (size_t m = 0; m < n_meshes; ++m) {
        if (meshes[m].n_tris < 0) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "negative triangle count");
        if (meshes[m].n_tris > 0 && !meshes[m].vertices && !meshes[m].vertices_f64)
            return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "mesh vertices is NULL");
        if (meshes[m].flags & ~RISE_HIP_MESH_DOUBLE_SIDED)
            return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "unknown mesh flags");
        total += meshes[m].n_tris;
        any64 = any64 || (meshes[m].n_tris > 0 && meshes[m].vertices_f64);
        any_attr = any_attr || (meshes[m].n_tris > 0 && (meshes[m].normals || meshes[m].coords));
    }
    if (total > 0x7fffffff) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "more than 2^31-1 triangles");

    // Concatenate meshes (identity transforms: world space == object space).
    std::vector<float> verts;
    std::vector<double> verts64, attrs;
    std::vector<int32_t> ids;
    try {
        verts.reserve((size_t)total * 9);
        verts64.reserve((size_t)total * 9);
        ids.resize((size_t)total);
    } catch (const std::bad_alloc&) {
        return fail(RISE_HIP_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
    int64_t off = 0;
    for (size_t m = 0; m < n_meshes; ++m) {
        const int64_t k = meshes[m].n_tris;
        if (k == 0) continue;
        mesh_vertices(meshes[m], verts, verts64);
        if (any_attr) mesh_attrs(meshes[m], attrs);
        for (int64_t i = 0; i < k; ++i) ids[off + i] = meshes[m].tri_ids ? meshes[m].tri_ids[i] : (int32_t)(off + i);
        off += k;
    }

    WideBVH bvh;
    std::vector<double> tri64, leaf_attr;
    std::string err;
    BuildParams params;
    try {
        // Built over the running index, then each leaf triangle takes its
        // caller id (v0.w) and its mesh index (v1.w, reported in hit records).
        if (!build_wide_bvh(verts.data(), total, nullptr, params, bvh, err))
            return fail(RISE_HIP_ERR_BUILD, "BVH build failed: " + err);
        if (any64) tri64 = leaf_vertices(bvh, verts64);
        if (any_attr) leaf_attr = leaf_rows(bvh, attrs, kAttrDoubles);
        std::vector<int32_t> mesh_of((size_t)total);
        int64_t o = 0;
        for (size_t m = 0; m < n_meshes; ++m)
            for (int64_t i = 0; i < meshes[m].n_tris; ++i) mesh_of[(size_t)(o++)] = (int32_t)m;
        for (Tri48& tr : bvh.tris) {
            int32_t p;
            std::memcpy(&p, &tr.v0[3], 4);
            std::memcpy(&tr.v0[3], &ids[(size_t)p], 4);
            const int32_t m = mesh_of[(size_t)p];
            std::memcpy(&tr.v1[3], &m, 4);
            const uint32_t fl = mesh_tri_flags(meshes[m]);
            std::memcpy(&tr.v2[3], &fl, 4);
        }
    } catch (const std::bad_alloc&) {
        return fail(RISE_HIP_ERR_OUT_OF_MEMORY, "host allocation failed during BVH build");
    }

    return upload_scene(bvh, total, 0, out_scene, any64 ? &tri64 : nullptr, any_attr ? &leaf_attr : nullptr);
}

}  // extern "C"

namespace rise_hip {

rise_hip_status upload_scene(const WideBVH& bvh, int64_t total, int32_t bvh_source, rise_hip_scene** out_scene,
                             const std::vector<double>* tri64, const std::vector<double>* attr)
{
    rise_hip_scene* s = new (std::nothrow) rise_hip_scene;
    if (!s) return fail(RISE_HIP_ERR_OUT_OF_MEMORY, "host allocation failed");
    hipError_t e = hipGetDevice(&s->device);
    if (e != hipSuccess) {
        delete s;
        return fail_hip(e, "hipGetDevice");
    }
    const size_t nb = bvh.nodes.size() * sizeof(NodeSlot);
    const size_t tb = std::max<size_t>(bvh.tris.size(), 1) * sizeof(Tri48);
    e = hipMalloc((void**)&s->d_nodes, nb);
    if (e == hipSuccess) e = hipMalloc((void**)&s->d_tris, tb);
    if (e == hipSuccess) e = hipMalloc((void**)&s->d_fetch, sizeof(unsigned long long) * rise_hip_scene::kFetchSlots);
    if (e == hipSuccess) e = hipMemcpy(s->d_nodes, bvh.nodes.data(), nb, hipMemcpyHostToDevice);
    if (e == hipSuccess && !bvh.tris.empty())
        e = hipMemcpy(s->d_tris, bvh.tris.data(), bvh.tris.size() * sizeof(Tri48), hipMemcpyHostToDevice);
    const size_t vb = tri64 ? std::max<size_t>(tri64->size(), 1) * sizeof(double) : 0;
    if (e == hipSuccess && tri64) e = hipMalloc((void**)&s->d_tri64, vb);
    if (e == hipSuccess && tri64 && !tri64->empty())
        e = hipMemcpy(s->d_tri64, tri64->data(), tri64->size() * sizeof(double), hipMemcpyHostToDevice);
    const size_t ab = (attr && !attr->empty()) ? attr->size() * sizeof(double) : 0;
    if (e == hipSuccess && ab) e = hipMalloc((void**)&s->d_attr, ab);
    if (e == hipSuccess && ab) e = hipMemcpy(s->d_attr, attr->data(), ab, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(s->d_nodes);
        (void)hipFree(s->d_tris);
        (void)hipFree(s->d_tri64);
        (void)hipFree(s->d_attr);
        (void)hipFree(s->d_fetch);
        delete s;
        return fail_hip(e, "scene upload");
    }
    std::memset(&s->stats, 0, sizeof(s->stats));
    s->stats.n_tris = total;
    s->stats.n_nodes = (int64_t)bvh.nodes.size();
    s->stats.n_bvh2_nodes = bvh.stats.bvh2_nodes;
    s->stats.depth = bvh.stats.depth;
    s->stats.node_bytes = (int32_t)sizeof(NodeSlot);
    s->stats.tri_bytes = (int32_t)sizeof(Tri48);
    s->stats.build_ms = bvh.stats.build_ms;
    s->stats.sah_cost = bvh.stats.sah_cost;
    s->stats.device_bytes = (int64_t)(nb + tb + vb + ab);
    s->stats.bvh_source = bvh_source;
    *out_scene = s;
    return RISE_HIP_OK;
}

}  // namespace rise_hip

extern "C" {

rise_hip_status rise_hip_scene_upload_objects(const rise_hip_mesh_desc* meshes, size_t n_meshes,
                                              const rise_hip_object_instance* objects, size_t n_objects,
                                              rise_hip_scene** out_scene)
{
    clear_error();
    if (!out_scene) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "out_scene is NULL");
    *out_scene = nullptr;
    ObjectSceneHost host;
    rise_hip_status st = build_object_scene(meshes, n_meshes, objects, n_objects, host);
    if (st != RISE_HIP_OK) return st;
    rise_hip_scene* s = nullptr;
    st = upload_scene(host.world, host.world_tris, 0, &s, &host.world_tri64,
                      host.world_attr.empty() ? nullptr : &host.world_attr);
    if (st != RISE_HIP_OK) return st;
    s->world_empty = host.world_tris == 0;
    s->stats.n_tris = host.total_tris;
    hipError_t e = hipSuccess;
    std::vector<BlasView> views;
    for (size_t k = 0; k < host.blas.size(); ++k) {
        if (e != hipSuccess) break;
        const WideBVH& b = host.blas[k];
        const std::vector<double>& v64 = host.blas_tri64[k];
        const std::vector<double>& at = host.blas_attr[k];
        const size_t nb = b.nodes.size() * sizeof(NodeSlot), tb = std::max<size_t>(b.tris.size(), 1) * sizeof(Tri48);
        const size_t vb = std::max<size_t>(v64.size(), 1) * sizeof(double);
        void *dn = nullptr, *dt = nullptr, *dv = nullptr, *da = nullptr;
        e = hipMalloc(&dn, nb);
        if (e == hipSuccess) s->blas_allocs.push_back(dn);
        if (e == hipSuccess) e = hipMalloc(&dt, tb);
        if (e == hipSuccess) s->blas_allocs.push_back(dt);
        if (e == hipSuccess) e = hipMalloc(&dv, vb);
        if (e == hipSuccess) s->blas_allocs.push_back(dv);
        if (e == hipSuccess) e = hipMemcpy(dn, b.nodes.data(), nb, hipMemcpyHostToDevice);
        if (e == hipSuccess && !b.tris.empty())
            e = hipMemcpy(dt, b.tris.data(), b.tris.size() * sizeof(Tri48), hipMemcpyHostToDevice);
        if (e == hipSuccess && !v64.empty())
            e = hipMemcpy(dv, v64.data(), v64.size() * sizeof(double), hipMemcpyHostToDevice);
        if (e == hipSuccess && !at.empty()) e = hipMalloc(&da, at.size() * sizeof(double));
        if (e == hipSuccess && da) s->blas_allocs.push_back(da);
        if (e == hipSuccess && da) e = hipMemcpy(da, at.data(), at.size() * sizeof(double), hipMemcpyHostToDevice);
        views.push_back(BlasView{(const NodeSlot*)dn, (const Tri48*)dt, (const double*)dv, (const double*)da});
        s->stats.device_bytes += (int64_t)(nb + tb + vb + at.size() * sizeof(double));
    }
    if (e == hipSuccess && !host.objects.empty()) {
        const size_t ob = host.objects.size() * sizeof(DevObject), vb = views.size() * sizeof(BlasView);
        const size_t tn = host.top.nodes.size() * sizeof(NodeSlot), tl = host.top.tris.size() * sizeof(Tri48);
        e = hipMalloc((void**)&s->d_objects, ob);
        if (e == hipSuccess) e = hipMalloc((void**)&s->d_blas, vb);
        if (e == hipSuccess) e = hipMalloc((void**)&s->d_top_nodes, tn);
        if (e == hipSuccess) e = hipMalloc((void**)&s->d_top_leaves, tl);
        if (e == hipSuccess) e = hipMemcpy(s->d_objects, host.objects.data(), ob, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(s->d_blas, views.data(), vb, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(s->d_top_nodes, host.top.nodes.data(), tn, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(s->d_top_leaves, host.top.tris.data(), tl, hipMemcpyHostToDevice);
        s->stats.device_bytes += (int64_t)(ob + vb + tn + tl);
    }
    if (e != hipSuccess) {
        (void)rise_hip_scene_release(s);
        return fail_hip(e, "object scene upload");
    }
    s->n_objects = (int32_t)host.objects.size();
    s->stack_need = host.stack_need;
    *out_scene = s;
    return RISE_HIP_OK;
}

rise_hip_status rise_hip_scene_release(rise_hip_scene* scene)
{
    clear_error();
    if (!scene) return RISE_HIP_OK;
    for (void* p : scene->blas_allocs) (void)hipFree(p);
    (void)hipFree(scene->d_objects);
    (void)hipFree(scene->d_blas);
    (void)hipFree(scene->d_top_nodes);
    (void)hipFree(scene->d_top_leaves);
    (void)hipFree(scene->d_nodes);
    (void)hipFree(scene->d_tris);
    (void)hipFree(scene->d_tri64);
    (void)hipFree(scene->d_attr);
    (void)hipFree(scene->d_fetch);
    delete scene;
    return RISE_HIP_OK;
}

rise_hip_status rise_hip_scene_get_stats(const rise_hip_scene* scene, rise_hip_scene_stats* out_stats)
{
    clear_error();
    if (!scene || !out_stats) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    *out_stats = scene->stats;
    return RISE_HIP_OK;
}

rise_hip_status rise_hip_scene_set_arithmetic(rise_hip_scene* scene, rise_hip_arith arith)
{
    clear_error();
    if (!scene) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "scene is NULL");
    if (arith != RISE_HIP_ARITH_IEEE && arith != RISE_HIP_ARITH_X86_FMA)
        return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "unknown arithmetic model");
    scene->arith.store((int)arith);
    return RISE_HIP_OK;
}

static rise_hip_status closest_impl(rise_hip_scene* scene, const float* rays, int64_t n, int32_t* prim_id, float* t,
                                    uint64_t* counters, hipStream_t stream)
{
    clear_error();
    if (!scene) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "scene is NULL");
    if (n < 0) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "negative ray count");
    if (n == 0) return RISE_HIP_OK;
    if (!rays || !prim_id || !t) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL buffer");
    if (scene->n_objects > 0) {
        if (counters)
            return fail(RISE_HIP_ERR_NOT_SUPPORTED, "counted queries on a scene with transformed objects");
        const hipError_t e = launch_closest_objects(scene->objects_view(), scene->arith.load(), scene->stack_need, rays, n,
                                                    3u, prim_id, t, nullptr, scene->fetch_slot(), stream);
        return e == hipSuccess ? RISE_HIP_OK : fail_hip(e, "closest-hit launch");
    }
    const hipError_t e = launch_closest(scene->view(), rays, n, prim_id, t,
                                        reinterpret_cast<unsigned long long*>(counters), stream);
    if (e != hipSuccess) return fail_hip(e, "closest-hit launch");
    return RISE_HIP_OK;
}

static rise_hip_status any_impl(rise_hip_scene* scene, const float* rays, const float* tmax, int64_t n,
                                uint8_t* occluded, uint64_t* counters, hipStream_t stream)
{
    clear_error();
    if (!scene) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "scene is NULL");
    if (n < 0) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "negative ray count");
    if (n == 0) return RISE_HIP_OK;
    if (!rays || !occluded) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL buffer");
    if (scene->n_objects > 0) {
        if (counters)
            return fail(RISE_HIP_ERR_NOT_SUPPORTED, "counted queries on a scene with transformed objects");
        const hipError_t e = launch_any_objects(scene->objects_view(), scene->arith.load(), scene->stack_need, rays, tmax,
                                                n, 3u, occluded, scene->fetch_slot(), stream);
        return e == hipSuccess ? RISE_HIP_OK : fail_hip(e, "any-hit launch");
    }
    const hipError_t e = launch_any(scene->view(), rays, tmax, n, occluded,
                                    reinterpret_cast<unsigned long long*>(counters), stream);
    if (e != hipSuccess) return fail_hip(e, "any-hit launch");
    return RISE_HIP_OK;
}

rise_hip_status rise_hip_intersect_closest(rise_hip_scene* scene, const float* rays_soa, int64_t n, int32_t* prim_id,
                                           float* t, hipStream_t stream)
{
    return closest_impl(scene, rays_soa, n, prim_id, t, nullptr, stream);
}

rise_hip_status rise_hip_intersect_any(rise_hip_scene* scene, const float* rays_soa, const float* tmax, int64_t n,
                                       uint8_t* occluded, hipStream_t stream)
{
    return any_impl(scene, rays_soa, tmax, n, occluded, nullptr, stream);
}

rise_hip_status rise_hip_intersect_closest_ex(rise_hip_scene* scene, const float* rays_soa, int64_t n, int32_t faces,
                                              rise_hip_hit* hits, hipStream_t stream)
{
    clear_error();
    if (!scene) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "scene is NULL");
    if (n < 0) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "negative ray count");
    if (faces < 0 || faces > 3) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "faces must be a combination of RISE_HIP_FACES_*");
    if (n == 0) return RISE_HIP_OK;
    if (!rays_soa || !hits) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL buffer");
    const hipError_t e = scene->n_objects > 0
                             ? launch_closest_objects(scene->objects_view(), scene->arith.load(), scene->stack_need, rays_soa, n,
                                                      (unsigned)faces, nullptr, nullptr, hits, scene->fetch_slot(), stream)
                             : launch_closest_ex(scene->view(), rays_soa, n, (unsigned)faces, hits, stream);
    if (e != hipSuccess) return fail_hip(e, "closest-hit launch");
    return RISE_HIP_OK;
}

rise_hip_status rise_hip_intersect_any_ex(rise_hip_scene* scene, const float* rays_soa, const float* tmax, int64_t n,
                                          int32_t faces, uint8_t* occluded, hipStream_t stream)
{
    clear_error();
    if (!scene) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "scene is NULL");
    if (n < 0) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "negative ray count");
    if (faces < 0 || faces > 3) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "faces must be a combination of RISE_HIP_FACES_*");
    if (n == 0) return RISE_HIP_OK;
    if (!rays_soa || !occluded) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL buffer");
    const hipError_t e = scene->n_objects > 0
                             ? launch_any_objects(scene->objects_view(), scene->arith.load(), scene->stack_need, rays_soa, tmax, n,
                                                  (unsigned)faces, occluded, scene->fetch_slot(), stream)
                             : launch_any_ex(scene->view(), rays_soa, tmax, n, (unsigned)faces, occluded, stream);
    if (e != hipSuccess) return fail_hip(e, "any-hit launch");
    return RISE_HIP_OK;
}

rise_hip_status rise_hip_intersect_closest_noalloc(rise_hip_scene* scene, const float* rays_soa, int64_t n,
                                                   int32_t* prim_id, float* t, hipStream_t stream)
{
    clear_error();
    if (!scene) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "scene is NULL");
    if (n < 0) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "negative ray count");
    if (n == 0) return RISE_HIP_OK;
    if (!rays_soa || !prim_id || !t) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL buffer");
    unsigned long long* counter = nullptr;
    hipError_t e = noalloc_scratch(&counter);
    if (e != hipSuccess) return fail_hip(e, "scratch address (hipGetSymbolAddress)");
    if (scene->n_objects > 0)
        e = launch_closest_objects(scene->objects_view(), scene->arith.load(), scene->stack_need, rays_soa, n, 3u,
                                   prim_id, t, nullptr, counter, stream);
    else
        e = launch_closest(scene->view_on(counter), rays_soa, n, prim_id, t, nullptr, stream);
    if (e != hipSuccess) return fail_hip(e, "closest-hit launch");
    return RISE_HIP_OK;
}

rise_hip_status rise_hip_intersect_closest_counted(rise_hip_scene* scene, const float* rays_soa, int64_t n,
                                                   int32_t* prim_id, float* t, uint64_t* counters, hipStream_t stream)
{
    if (!counters) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "counters is NULL");
    return closest_impl(scene, rays_soa, n, prim_id, t, counters, stream);
}

rise_hip_status rise_hip_intersect_any_counted(rise_hip_scene* scene, const float* rays_soa, const float* tmax,
                                               int64_t n, uint8_t* occluded, uint64_t* counters, hipStream_t stream)
{
    if (!counters) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "counters is NULL");
    return any_impl(scene, rays_soa, tmax, n, occluded, counters, stream);
}

}  // extern "C"

namespace {

struct CacheHeader {
    char magic[8];
    uint32_t version, node_bytes, depth, reserved;
    uint64_t n_nodes, n_tris;
};
static_assert(sizeof(CacheHeader) == 40, "cache header layout");
constexpr char kCacheMagic[8] = {'R', 'H', 'B', 'V', 'H', '8', 0, 0};

// Bounds and shape of a wide tree read from a file: every internal child and
// leaf triangle range inside the arrays, children after their parent (so the
// traversal cannot loop), every node reached from one parent only (a tree,
// not a DAG, so the level computed here is the node's only depth), depth
// within the traversal stack.
bool check_wide(const std::vector<NodeSlot>& nodes, uint64_t n_tris, int& depth, std::string& err)
{
    const uint64_t nn = nodes.size();
    std::vector<uint8_t> level(nn, 0);
    level[0] = 1;
    depth = 1;
    for (uint64_t i = 0; i < nn; ++i) {
        const Node8& n = nodes[i];
        if (level[i] == 0) continue;  // unreachable from the root: never traversed
        const unsigned internal = (unsigned)__builtin_popcount(n.imask);
        if (internal > 0 && (n.child_base <= i || (uint64_t)n.child_base + internal > nn)) {
            err = "node " + std::to_string(i) + ": child range out of bounds";
            return false;
        }
        for (unsigned k = 0; k < internal; ++k) {
            if (level[n.child_base + k] != 0) {
                err = "node " + std::to_string(n.child_base + k) + " is reached from two parents";
                return false;
            }
            level[n.child_base + k] = (uint8_t)std::min(255, level[i] + 1);
            depth = std::max(depth, level[i] + 1);
        }
        for (int s = 0; s < 8; ++s) {
            if (n.meta[s] == 0 || ((n.imask >> s) & 1u)) continue;
            const uint64_t cnt = n.meta[s] >> 5, off = n.meta[s] & 31u;
            if (cnt == 0 || (uint64_t)n.tri_base + off + cnt > n_tris) {
                err = "node " + std::to_string(i) + ": leaf triangles out of bounds";
                return false;
            }
        }
    }
    if (depth > kMaxStack) {
        err = "tree deeper than the traversal stack";
        return false;
    }
    return true;
}

}  // namespace

extern "C" rise_hip_status rise_hip_scene_cache_write(const rise_hip_scene* scene, const char* path)
{
    clear_error();
    if (!scene || !path) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if (scene->n_objects > 0 || scene->d_tri64 || scene->d_attr)
        return fail(RISE_HIP_ERR_NOT_SUPPORTED,
                    "scene cache of a scene with transformed objects, double-precision vertices or vertex "
                    "attributes");
    const uint64_t nn = (uint64_t)scene->stats.n_nodes, nt = (uint64_t)scene->stats.n_tris;
    std::vector<NodeSlot> nodes;
    std::vector<Tri48> tris;
    try {
        nodes.resize(nn);
        tris.resize(nt);
    } catch (const std::bad_alloc&) {
        return fail(RISE_HIP_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
    hipError_t e = hipSetDevice(scene->device);
    if (e == hipSuccess) e = hipMemcpy(nodes.data(), scene->d_nodes, nn * sizeof(NodeSlot), hipMemcpyDeviceToHost);
    if (e == hipSuccess && nt) e = hipMemcpy(tris.data(), scene->d_tris, nt * sizeof(Tri48), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail_hip(e, "scene cache copy");
    CacheHeader h{};
    std::memcpy(h.magic, kCacheMagic, 8);
    h.version = 1;
    h.node_bytes = (uint32_t)sizeof(NodeSlot);
    h.depth = (uint32_t)scene->stats.depth;
    h.n_nodes = nn;
    h.n_tris = nt;
    std::FILE* f = std::fopen(path, "wb");
    if (!f) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, std::string("cannot create ") + path);
    bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1 && std::fwrite(nodes.data(), sizeof(NodeSlot), nn, f) == nn &&
              std::fwrite(tris.data(), sizeof(Tri48), nt, f) == nt;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? RISE_HIP_OK : fail(RISE_HIP_ERR_INVALID_ARGUMENT, std::string("cannot write ") + path);
}

extern "C" rise_hip_status rise_hip_scene_cache_read(const char* path, rise_hip_scene** out_scene)
{
    clear_error();
    if (!path || !out_scene) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    *out_scene = nullptr;
    const auto t0 = std::chrono::steady_clock::now();
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, std::string("cannot open ") + path);
    const uint64_t size = (uint64_t)f.tellg();
    f.seekg(0);
    CacheHeader h{};
    if (size < sizeof(h) || !f.read((char*)&h, sizeof(h)) || std::memcmp(h.magic, kCacheMagic, 8) != 0 ||
        h.version != 1)
        return fail(RISE_HIP_ERR_INVALID_ARGUMENT, std::string(path) + ": not a scene cache (version 1)");
    if (h.node_bytes != sizeof(NodeSlot) || h.n_nodes == 0 || h.n_tris > 0x7fffffffull ||
        h.n_nodes > (size - sizeof(h)) / sizeof(NodeSlot) ||
        size != sizeof(h) + h.n_nodes * sizeof(NodeSlot) + h.n_tris * sizeof(Tri48))
        return fail(RISE_HIP_ERR_INVALID_ARGUMENT, std::string(path) + ": sizes do not match the file");
    WideBVH bvh;
    std::string err;
    int depth = 0;
    try {
        bvh.nodes.resize(h.n_nodes);
        bvh.tris.resize(h.n_tris);
        if (!f.read((char*)bvh.nodes.data(), (std::streamsize)(h.n_nodes * sizeof(NodeSlot))) ||
            !f.read((char*)bvh.tris.data(), (std::streamsize)(h.n_tris * sizeof(Tri48))))
            return fail(RISE_HIP_ERR_INVALID_ARGUMENT, std::string(path) + ": file truncated");
        if (!check_wide(bvh.nodes, h.n_tris, depth, err))
            return fail(RISE_HIP_ERR_INVALID_ARGUMENT, std::string(path) + ": " + err);
    } catch (const std::bad_alloc&) {
        return fail(RISE_HIP_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
    bvh.stats.nodes = (int64_t)h.n_nodes;
    bvh.stats.tris = (int64_t)h.n_tris;
    bvh.stats.depth = depth;
    bvh.stats.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return upload_scene(bvh, (int64_t)h.n_tris, 2, out_scene);
}