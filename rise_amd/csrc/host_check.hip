// host_check.hip -- runs the kernels' traversal code (trace_core.h) on host
// threads, for validation on machines without a GPU.
#include "rise_hip_debug.h"

#include "api_common.h"
#include "bvh_build.h"
#include "object_scene.h"
#include "object_trace.h"
#include "risemesh_internal.h"
#include "trace_core.h"

#include <atomic>
#include <cstring>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

using namespace rise_hip;

namespace {

struct HostStack {
    uint2 e[kMaxStack];
    int sp = 0;
    void push(uint2 v)
    {
        if (sp >= kMaxStack) throw std::runtime_error("host traversal stack overflow");
        e[sp++] = v;
    }
    uint2 pop() { return e[--sp]; }
    bool empty() const { return sp == 0; }
};

}  // namespace

namespace {

rise_hip_status check_trace_args(const float* rays, int64_t n, int32_t mode, int32_t* prim_id, float* t,
                                 uint8_t* occluded)
{
    if (n < 0) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "negative count");
    if (n > 0 && !rays) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL buffer");
    if ((mode & 1) == 0 && n > 0 && (!prim_id || !t)) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL output");
    if ((mode & 1) == 1 && n > 0 && !occluded) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL output");
    if (mode < 0 || mode > 3) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "mode must be 0..3");
    return RISE_HIP_OK;
}

rise_hip_status trace_host(const WideBVH& bvh, int64_t n_tris, int32_t bvh_source, const float* rays,
                           const float* tmax, int64_t n, int32_t mode, int32_t* prim_id, float* t,
                           uint8_t* occluded, uint64_t* counters, rise_hip_scene_stats* stats);

}  // namespace

extern "C" rise_hip_status rise_hip_debug_node_accept_host(const void* node, const float ray[6], double best,
                                                           uint32_t* accept)
{
    clear_error();
    if (!node || !ray || !accept) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    uint4 w[5];
    std::memcpy(w, node, sizeof(w));
    RayState r;
    init_ray(r, ray[0], ray[1], ray[2], ray[3], ray[4], ray[5]);
    *accept = node_accept_mask(r, w[0], w[1], w[2], w[3], w[4], box_limit(r, best));
    return RISE_HIP_OK;
}

extern "C" rise_hip_status rise_hip_debug_trace_ex_host(const float* tris, int64_t n_tris, const float* rays,
                                                        const float* tmax, int64_t n, int32_t mode, int32_t faces,
                                                        rise_hip_hit* hits, uint8_t* occluded)
{
    clear_error();
    if (n_tris < 0 || n < 0) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "negative count");
    if ((n_tris > 0 && !tris) || (n > 0 && !rays)) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL buffer");
    if (mode < 0 || mode > 3) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "mode must be 0..3");
    if (faces < 0 || faces > 3) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "faces must be 0..3");
    if (n > 0 && (mode & 1) == 0 && !hits) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL output");
    if (n > 0 && (mode & 1) == 1 && !occluded) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL output");
    WideBVH bvh;
    std::string err;
    try {
        if (!build_wide_bvh(tris, n_tris, nullptr, BuildParams(), bvh, err))
            return fail(RISE_HIP_ERR_BUILD, "BVH build failed: " + err);
    } catch (const std::bad_alloc&) {
        return fail(RISE_HIP_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
    const NodeSlot* nodes = bvh.nodes.data();
    const Tri48* trs = bvh.tris.data();
    std::atomic<int64_t> next(0);
    std::atomic<bool> overflow(false);
    auto one = [&](auto r, int64_t i) {
        init_ray(r, rays[i], rays[n + i], rays[2 * n + i], rays[3 * n + i], rays[4 * n + i], rays[5 * n + i]);
        HostStack st;
        unsigned nn = 0, nt = 0;
        if ((mode & 1) == 0) {
            ClosestState cs;
            closest_begin(r, cs);
            cs.faces = (unsigned)faces;
            while (closest_step<false, true>(r, nodes, trs, st, cs, nn, nt)) {
            }
            rise_hip_hit h{};
            h.prim_id = cs.best_id;
            h.mesh = -1;
            h.range = INFINITY;
            if (cs.best_id >= 0) {
                double ab[2], nrm[3], pt[3], sn[3], uv[2];
                hit_attributes(r, trs + cs.best_leaf, cs.best_t, ab, nrm, pt, nullptr, nullptr, sn, uv);
                h.mesh = float_to_bits(trs[cs.best_leaf].v1[3]);
                h.range = world_range(r, cs.best_t);
                h.alpha = ab[0];
                h.beta = ab[1];
                for (int k = 0; k < 3; ++k) {
                    h.point[k] = (float)pt[k];
                    h.geom_normal[k] = (float)nrm[k];
                    h.shading_normal[k] = (float)sn[k];
                }
                h.uv[0] = (float)uv[0];
                h.uv[1] = (float)uv[1];
            }
            hits[i] = h;
        } else {
            AnyState as;
            const double hf = tmax ? (double)tmax[i] : kRiseInfinity;
            any_begin(r, scale_how_far(r, hf), as);
            as.faces = (unsigned)faces;
            while (any_step<false, true>(r, nodes, trs, st, as, nn, nt)) {
            }
            occluded[i] = as.hit ? 1 : 0;
        }
    };
    auto worker = [&]() {
        for (;;) {
            const int64_t s0 = next.fetch_add(64);
            if (s0 >= n) break;
            for (int64_t i = s0; i < std::min<int64_t>(s0 + 64, n); ++i) {
                try {
                    if (mode & 2) one(RayStateT<kArithX86Fma>{}, i);
                    else one(RayStateT<kArithIeee>{}, i);
                } catch (const std::exception&) {
                    overflow = true;
                }
            }
        }
    };
    std::vector<std::thread> th;
    for (int k = 1; k < host_threads(); ++k) th.emplace_back(worker);
    worker();
    for (auto& x : th) x.join();
    if (overflow) return fail(RISE_HIP_ERR_BUILD, "traversal stack overflow");
    return RISE_HIP_OK;
}

extern "C" rise_hip_status rise_hip_debug_objects_host(const rise_hip_mesh_desc* meshes, size_t n_meshes,
                                                       const rise_hip_object_instance* objects, size_t n_objects,
                                                       const float* rays, const float* tmax, int64_t n, int32_t mode,
                                                       int32_t faces, rise_hip_hit* hits, uint8_t* occluded)
{
    clear_error();
    if (n < 0) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "negative count");
    if (n > 0 && !rays) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL buffer");
    if (mode < 0 || mode > 3) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "mode must be 0..3");
    if (faces < 0 || faces > 3) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "faces must be 0..3");
    if (n > 0 && (mode & 1) == 0 && !hits) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL output");
    if (n > 0 && (mode & 1) == 1 && !occluded) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL output");
    ObjectSceneHost host;
    const rise_hip_status st = build_object_scene(meshes, n_meshes, objects, n_objects, host);
    if (st != RISE_HIP_OK) return st;
    std::vector<BlasView> views;
    for (size_t k = 0; k < host.blas.size(); ++k)
        views.push_back(BlasView{host.blas[k].nodes.data(), host.blas[k].tris.data(), host.blas_tri64[k].data(),
                                 host.blas_attr[k].empty() ? nullptr : host.blas_attr[k].data()});
    const ObjectSceneView sc{host.world_tris > 0 ? host.world.nodes.data() : nullptr, host.world.tris.data(),
                             host.world_tri64.data(), host.world_attr.empty() ? nullptr : host.world_attr.data(),
                             host.objects.data(), (int32_t)host.objects.size(), views.data(),
                             host.top.nodes.data(), host.top.tris.data()};
    std::atomic<int64_t> next(0);
    std::atomic<bool> overflow(false);
    auto one = [&](auto arith, int64_t i) {
        constexpr int A = decltype(arith)::value;
        float ray[6];
        for (int k = 0; k < 6; ++k) ray[k] = rays[(int64_t)k * n + i];
        ObjStack st;
        if ((mode & 1) == 0) {
            ObjectHit h;
            objects_closest<A>(sc, ray, (unsigned)faces, st, h);
            rise_hip_hit rec{};
            objects_record<A>(sc, ray, h, rec);
            hits[i] = rec;
        } else {
            const double hf = tmax ? (double)tmax[i] : kRiseInfinity;
            occluded[i] = objects_any<A>(sc, ray, hf, (unsigned)faces, st) ? 1 : 0;
        }
        if (st.overflow) overflow = true;
    };
    auto worker = [&]() {
        for (;;) {
            const int64_t s0 = next.fetch_add(64);
            if (s0 >= n) break;
            for (int64_t i = s0; i < std::min<int64_t>(s0 + 64, n); ++i) {
                try {
                    if (mode & 2) one(std::integral_constant<int, kArithX86Fma>{}, i);
                    else one(std::integral_constant<int, kArithIeee>{}, i);
                } catch (const std::exception&) {
                    overflow = true;
                }
            }
        }
    };
    std::vector<std::thread> th;
    for (int k = 1; k < host_threads(); ++k) th.emplace_back(worker);
    worker();
    for (auto& x : th) x.join();
    if (overflow) return fail(RISE_HIP_ERR_BUILD, "traversal stack overflow");
    return RISE_HIP_OK;
}

// The kernels' double slab test of one 8-wide node against an object-space
// ray made from a world ray and an object's inverse matrix.
extern "C" rise_hip_status rise_hip_debug_node_accept_obj_host(const void* node, const float ray[6],
                                                               const double to_object[16], double best,
                                                               uint32_t* accept)
{
    clear_error();
    if (!node || !ray || !to_object || !accept) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    uint4 w[5];
    std::memcpy(w, node, sizeof(w));
    const double ow[3] = {ray[0], ray[1], ray[2]}, dw[3] = {ray[3], ray[4], ray[5]};
    ObjRayT<kArithIeee> q;
    init_obj_ray(q, ow, dw, to_object);
    *accept = node_accept_mask(q, w[0], w[1], w[2], w[3], w[4], box_limit(q, best));
    return RISE_HIP_OK;
}

extern "C" rise_hip_status rise_hip_debug_trace_host(const float* tris, int64_t n_tris, const float* rays,
                                                     const float* tmax, int64_t n, int32_t mode, int32_t* prim_id,
                                                     float* t, uint8_t* occluded, uint64_t* counters,
                                                     rise_hip_scene_stats* stats)
{
    clear_error();
    if (n_tris < 0) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "negative count");
    if (n_tris > 0 && !tris) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL buffer");
    rise_hip_status st = check_trace_args(rays, n, mode, prim_id, t, occluded);
    if (st != RISE_HIP_OK) return st;
    WideBVH bvh;
    std::string err;
    try {
        if (!build_wide_bvh(tris, n_tris, nullptr, BuildParams(), bvh, err))
            return fail(RISE_HIP_ERR_BUILD, "BVH build failed: " + err);
    } catch (const std::bad_alloc&) {
        return fail(RISE_HIP_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
    return trace_host(bvh, n_tris, 0, rays, tmax, n, mode, prim_id, t, occluded, counters, stats);
}

extern "C" rise_hip_status rise_hip_debug_trace_host_risemesh(const rise_hip_risemesh* mesh, const float* rays,
                                                              const float* tmax, int64_t n, int32_t mode,
                                                              int32_t* prim_id, float* t, uint8_t* occluded,
                                                              uint64_t* counters, rise_hip_scene_stats* stats)
{
    clear_error();
    if (!mesh) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL mesh");
    rise_hip_status st = check_trace_args(rays, n, mode, prim_id, t, occluded);
    if (st != RISE_HIP_OK) return st;
    WideBVH bvh;
    std::string err;
    int64_t n_tris = 0;
    try {
        const std::vector<double> d = risemesh_triangles_f64(mesh);
        std::vector<float> tris(d.size());
        for (size_t i = 0; i < d.size(); ++i) tris[i] = (float)d[i];
        n_tris = (int64_t)(tris.size() / 9);
        const std::vector<Bvh2Part> parts{{0, n_tris, risemesh_tree(mesh)}};
        if (!build_wide_bvh_parts(tris.data(), n_tris, nullptr, parts, BuildParams(), bvh, err))
            return fail(RISE_HIP_ERR_BUILD, "BVH build failed: " + err);
    } catch (const std::bad_alloc&) {
        return fail(RISE_HIP_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
    return trace_host(bvh, n_tris, bvh.stats.imported_tris > 0 ? 1 : 0, rays, tmax, n, mode, prim_id, t, occluded,
                      counters, stats);
}

namespace {

rise_hip_status trace_host(const WideBVH& bvh, int64_t n_tris, int32_t bvh_source, const float* rays,
                           const float* tmax, int64_t n, int32_t mode, int32_t* prim_id, float* t,
                           uint8_t* occluded, uint64_t* counters, rise_hip_scene_stats* stats)
{
    if (stats) {
        *stats = rise_hip_scene_stats{};
        stats->n_tris = n_tris;
        stats->n_nodes = (int64_t)bvh.nodes.size();
        stats->n_bvh2_nodes = bvh.stats.bvh2_nodes;
        stats->depth = bvh.stats.depth;
        stats->node_bytes = (int32_t)sizeof(NodeSlot);
        stats->tri_bytes = (int32_t)sizeof(Tri48);
        stats->build_ms = bvh.stats.build_ms;
        stats->sah_cost = bvh.stats.sah_cost;
        stats->bvh_source = bvh_source;
    }
    const NodeSlot* nodes = bvh.nodes.data();
    const Tri48* trs = bvh.tris.data();
    std::atomic<int64_t> next(0);
    std::atomic<unsigned long long> cn(0), ct(0);
    std::atomic<bool> overflow(false);
    auto trace_one = [&](auto r, int64_t i, unsigned& nn, unsigned& nt) {
        init_ray(r, rays[i], rays[n + i], rays[2 * n + i], rays[3 * n + i], rays[4 * n + i], rays[5 * n + i]);
        HostStack st;
        if ((mode & 1) == 0) {
            double best;
            unsigned leaf;
            const int id = trace_closest<true>(r, nodes, trs, st, best, leaf, nn, nt);
            prim_id[i] = id;
            t[i] = id >= 0 ? (float)world_range(r, best) : INFINITY;
        } else {
            const double hf = tmax ? (double)tmax[i] : kRiseInfinity;
            occluded[i] = trace_any<true>(r, nodes, trs, st, scale_how_far(r, hf), nn, nt) ? 1 : 0;
        }
    };
    auto worker = [&]() {
        unsigned long long ln = 0, lt = 0;
        for (;;) {
            const int64_t s = next.fetch_add(64);
            if (s >= n) break;
            const int64_t e = std::min<int64_t>(s + 64, n);
            for (int64_t i = s; i < e; ++i) {
                unsigned nn = 0, nt = 0;
                try {
                    if (mode & 2) trace_one(RayStateT<kArithX86Fma>{}, i, nn, nt);
                    else trace_one(RayStateT<kArithIeee>{}, i, nn, nt);
                } catch (const std::exception&) {
                    overflow = true;
                }
                ln += nn;
                lt += nt;
            }
        }
        cn += ln;
        ct += lt;
    };
    const int nth = host_threads();
    std::vector<std::thread> th;
    for (int k = 1; k < nth; ++k) th.emplace_back(worker);
    worker();
    for (auto& x : th) x.join();
    if (overflow) return fail(RISE_HIP_ERR_BUILD, "traversal stack overflow");
    if (counters) {
        counters[0] = cn.load();
        counters[1] = ct.load();
    }
    return RISE_HIP_OK;
}

}  // namespace
