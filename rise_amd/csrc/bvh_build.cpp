// bvh_build.cpp -- binned-SAH BVH2 build, collapse to 8-wide nodes, quantization.
// See bvh_build.h for the reference lines this follows.
#include "bvh_build.h"

#include "api_common.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <future>
#include <limits>
#include <thread>

namespace rise_hip {

namespace {

struct Box {
    float lo[3], hi[3];
    void reset()
    {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::numeric_limits<float>::infinity();
            hi[k] = -std::numeric_limits<float>::infinity();
        }
    }
    void grow(const Box& b)
    {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], b.lo[k]);
            hi[k] = std::max(hi[k], b.hi[k]);
        }
    }
    void grow(const float p[3])
    {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], p[k]);
            hi[k] = std::max(hi[k], p[k]);
        }
    }
    bool empty() const { return lo[0] > hi[0]; }
    // Surface area in double (BVH.h:1083-1088 SurfaceArea: 0 for an inverted box).
    double area() const
    {
        const double ex = (double)hi[0] - lo[0], ey = (double)hi[1] - lo[1], ez = (double)hi[2] - lo[2];
        if (ex < 0 || ey < 0 || ez < 0) return 0.0;
        return 2.0 * (ex * ey + ey * ez + ez * ex);
    }
};

struct Node2 {
    Box box;
    int32_t left = -1;   // internal: children left, left + 1
    uint32_t first = 0;  // first index into idx of the subtree's triangles
    uint32_t count = 0;  // leaf: > 0
    uint32_t span = 0;   // triangles under the node: idx[first, first + span)
    uint8_t axis = 0;    // internal: split axis
};

struct Bvh2Builder {
    const BuildParams& P;
    std::vector<Box> pbox;
    std::vector<float> cent;  // 3 per prim
    std::vector<uint32_t> idx;
    std::vector<Node2> nodes;
    std::atomic<int64_t> next_node{1};
    std::atomic<int> spawn_budget{0};

    explicit Bvh2Builder(const BuildParams& p) : P(p) {}

    int64_t alloc_pair() { return next_node.fetch_add(2); }

    void make_leaf(int64_t ni, uint32_t first, uint32_t count)
    {
        nodes[ni].left = -1;
        nodes[ni].first = first;
        nodes[ni].count = count;
        nodes[ni].span = count;
    }

    void build(int64_t ni, uint32_t first, uint32_t last)
    {
        const uint32_t count = last - first;
        Box box, cbox;
        box.reset();
        cbox.reset();
        for (uint32_t i = first; i < last; ++i) {
            box.grow(pbox[idx[i]]);
            cbox.grow(&cent[3 * (size_t)idx[i]]);
        }
        nodes[ni].box = box;
        if (count <= 1) {
            make_leaf(ni, first, count);
            return;
        }

        // Binned SAH over all three axes.
        const int B = P.bins;
        struct Bin {
            Box b;
            uint32_t n;
        };
        int best_axis = -1, best_split = -1;
        double best_cost = std::numeric_limits<double>::infinity();
        const double parent_area = box.area();
        std::vector<Bin> bins((size_t)B * 3);
        std::vector<Box> rbox(B);
        std::vector<uint32_t> rcnt(B);
        float cmin[3], inv[3];
        bool ok_axis[3];
        for (int a = 0; a < 3; ++a) {
            cmin[a] = cbox.lo[a];
            const float ext = cbox.hi[a] - cbox.lo[a];
            ok_axis[a] = ext > 0.0f;
            inv[a] = ok_axis[a] ? (float)B / ext : 0.0f;
            for (int b = 0; b < B; ++b) {
                bins[(size_t)a * B + b].b.reset();
                bins[(size_t)a * B + b].n = 0;
            }
        }
        for (uint32_t i = first; i < last; ++i) {
            const uint32_t p = idx[i];
            for (int a = 0; a < 3; ++a) {
                if (!ok_axis[a]) continue;
                int b = (int)((cent[3 * (size_t)p + a] - cmin[a]) * inv[a]);
                b = std::min(std::max(b, 0), B - 1);
                Bin& bin = bins[(size_t)a * B + b];
                bin.b.grow(pbox[p]);
                bin.n++;
            }
        }
        for (int a = 0; a < 3; ++a) {
            if (!ok_axis[a]) continue;
            const Bin* ab = &bins[(size_t)a * B];
            Box acc;
            acc.reset();
            uint32_t n = 0;
            for (int b = B - 1; b >= 1; --b) {
                if (ab[b].n) acc.grow(ab[b].b);
                n += ab[b].n;
                rbox[b] = acc;
                rcnt[b] = n;
            }
            acc.reset();
            n = 0;
            for (int b = 0; b + 1 < B; ++b) {
                if (ab[b].n) acc.grow(ab[b].b);
                n += ab[b].n;
                const uint32_t nr = rcnt[b + 1];
                if (n == 0 || nr == 0) continue;
                const double cost = P.sah_traversal +
                                    P.sah_intersect * (acc.area() * n + rbox[b + 1].area() * nr) /
                                        (parent_area > 0 ? parent_area : 1.0);
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = a;
                    best_split = b;
                }
            }
        }

        const double leaf_cost = P.sah_intersect * (double)count;
        if (count <= (uint32_t)P.max_leaf && (best_axis < 0 || best_cost >= leaf_cost)) {
            make_leaf(ni, first, count);
            return;
        }

        uint32_t mid = first;
        if (best_axis >= 0) {
            const int a = best_axis;
            for (uint32_t i = first; i < last; ++i) {
                const uint32_t p = idx[i];
                int b = (int)((cent[3 * (size_t)p + a] - cmin[a]) * inv[a]);
                b = std::min(std::max(b, 0), B - 1);
                if (b <= best_split) std::swap(idx[i], idx[mid++]);
            }
        }
        if (mid == first || mid == last) {
            // Median fallback (BVH.h:1002-1012): all centroids coincide or the
            // bins collapsed; split by centroid order on the widest axis, ties
            // by primitive index so the result is deterministic.
            int a = 0;
            const float e0 = cbox.hi[0] - cbox.lo[0], e1 = cbox.hi[1] - cbox.lo[1], e2 = cbox.hi[2] - cbox.lo[2];
            if (e1 > e0) a = 1;
            if (e2 > std::max(e0, e1)) a = 2;
            best_axis = a;
            mid = first + count / 2;
            std::nth_element(idx.begin() + first, idx.begin() + mid, idx.begin() + last,
                             [&](uint32_t x, uint32_t y) {
                                 const float cx = cent[3 * (size_t)x + a], cy = cent[3 * (size_t)y + a];
                                 return cx < cy || (cx == cy && x < y);
                             });
        }

        const int64_t l = alloc_pair();
        nodes[ni].left = (int32_t)l;
        nodes[ni].count = 0;
        nodes[ni].first = first;
        nodes[ni].span = count;
        nodes[ni].axis = (uint8_t)(best_axis >= 0 ? best_axis : 0);
        const bool spawn = count > 65536 && spawn_budget.fetch_sub(1) > 0;
        if (spawn) {
            auto fut = std::async(std::launch::async, [this, l, first, mid] { build(l, first, mid); });
            build(l + 1, mid, last);
            fut.get();
        } else {
            build(l, first, mid);
            build(l + 1, mid, last);
        }
    }
};

float exp2f_int(int e) { return std::ldexp(1.0f, e); }

// Quantize one axis of all children of a node; returns false if the chosen
// exponent cannot represent the range (caller retries with e + 1).
bool quantize_axis(float p, int e, const float* clo, const float* chi, int nchild, uint8_t* qlo, uint8_t* qhi)
{
    const float s = exp2f_int(e);
    for (int c = 0; c < nchild; ++c) {
        double fl = std::floor(((double)clo[c] - p) / s);
        double fh = std::ceil(((double)chi[c] - p) / s);
        int lo = (int)std::max(0.0, std::min(255.0, fl));
        int hi = (int)std::max(0.0, std::min(256.0, fh));
        while (lo > 0 && std::fmaf((float)lo, s, p) > clo[c]) --lo;
        if (std::fmaf((float)lo, s, p) > clo[c]) return false;
        while (hi <= 255 && std::fmaf((float)hi, s, p) < chi[c]) ++hi;
        if (hi > 255) return false;
        if (std::fmaf((float)hi, s, p) < chi[c]) return false;
        qlo[c] = (uint8_t)lo;
        qhi[c] = (uint8_t)hi;
    }
    return true;
}

}  // namespace

void decode_child_box(const Node8& n, int s, float lo[3], float hi[3])
{
    const float sx = exp2f_int(n.ex), sy = exp2f_int(n.ey), sz = exp2f_int(n.ez);
    lo[0] = std::fmaf((float)n.qlox[s], sx, n.px);
    lo[1] = std::fmaf((float)n.qloy[s], sy, n.py);
    lo[2] = std::fmaf((float)n.qloz[s], sz, n.pz);
    hi[0] = std::fmaf((float)n.qhix[s], sx, n.px);
    hi[1] = std::fmaf((float)n.qhiy[s], sy, n.py);
    hi[2] = std::fmaf((float)n.qhiz[s], sz, n.pz);
}

namespace {

bool check_params(const BuildParams& params, int64_t n, std::string& err)
{
    if (n < 0 || n > (int64_t)0x7fffffff) {
        err = "triangle count out of range";
        return false;
    }
    if (params.max_leaf < 1 || params.max_leaf > kMaxLeafTris || params.bins < 2 || params.collapse_leaf < 0 ||
        params.collapse_leaf > kMaxLeafTris) {
        err = "invalid build parameters";
        return false;
    }
    return true;
}

void empty_scene(WideBVH& out)
{
    // Empty scene: a single node with no children; every ray misses.
    NodeSlot root;
    std::memset(&root, 0, sizeof(root));
    out.nodes.push_back(root);
    out.stats.nodes = 1;
    out.stats.depth = 1;
}

// Primitive boxes, padded outward, and centroids; the scene bounds.
bool prepare_prims(const float* tris, int64_t n, Bvh2Builder& b, WideBVH& out, std::string& err)
{
    b.pbox.resize((size_t)n);
    b.cent.resize((size_t)n * 3);
    b.idx.resize((size_t)n);
    Box scene;
    scene.reset();
    for (int64_t i = 0; i < n; ++i) {
        const float* f = tris + 9 * i;
        for (int k = 0; k < 9; ++k) {
            if (!std::isfinite(f[k])) {
                err = "non-finite vertex coordinate in triangle " + std::to_string(i);
                return false;
            }
        }
        Box bb;
        bb.reset();
        bb.grow(f);
        bb.grow(f + 3);
        bb.grow(f + 6);
        scene.grow(bb);
        b.pbox[i] = bb;
        b.idx[i] = (uint32_t)i;
    }
    // Pad every primitive box outward.  The double-precision certifier accepts
    // barycentrics up to 1e-12 outside [0,1] (RayTriangleIntersection.cpp:46,54),
    // i.e. hit points a hair outside the triangle; the pad (1e-9 of the scene
    // extent plus one float step) keeps those inside the box.
    const float ext = std::max(std::max(scene.hi[0] - scene.lo[0], scene.hi[1] - scene.lo[1]),
                               std::max(scene.hi[2] - scene.lo[2], 1e-30f));
    const float pad = ext * 1e-9f;
    for (int64_t i = 0; i < n; ++i) {
        Box& bb = b.pbox[i];
        for (int k = 0; k < 3; ++k) {
            bb.lo[k] = std::nextafter(bb.lo[k] - pad, -std::numeric_limits<float>::infinity());
            bb.hi[k] = std::nextafter(bb.hi[k] + pad, std::numeric_limits<float>::infinity());
            b.cent[3 * i + k] = 0.5f * bb.lo[k] + 0.5f * bb.hi[k];
        }
    }
    for (int k = 0; k < 3; ++k) {
        out.scene_lo[k] = scene.lo[k];
        out.scene_hi[k] = scene.hi[k];
    }
    return true;
}

int threads_of(const BuildParams& params)
{
    return params.threads > 0 ? params.threads : host_threads();
}

}  // namespace

// Checks a binary tree over `count` primitives (from a file: anything may be
// wrong): node 0 the root, children in range and each node reached once,
// prims a permutation, every prim position under exactly one leaf.
bool check_bvh2(const Bvh2Tree& t, uint32_t count, std::string& err)
{
    const size_t nn = t.nodes.size();
    if (nn == 0 || t.prims.size() != count) {
        err = "node or primitive count does not match the mesh";
        return false;
    }
    std::vector<uint8_t> seen(count, 0);
    for (uint32_t p : t.prims) {
        if (p >= count || seen[p]) {
            err = "primitive order is not a permutation";
            return false;
        }
        seen[p] = 1;
    }
    std::fill(seen.begin(), seen.end(), 0);
    std::vector<uint8_t> reached(nn, 0);
    std::vector<uint32_t> stack{0};
    reached[0] = 1;
    while (!stack.empty()) {
        const Bvh2Node& nd = t.nodes[stack.back()];
        stack.pop_back();
        if (nd.count == 0) {
            const uint64_t l = nd.first_or_left;
            if (l == 0 || l + 1 >= nn || reached[l] || reached[l + 1]) {
                err = "bad child index";
                return false;
            }
            reached[l] = reached[l + 1] = 1;
            stack.push_back((uint32_t)l);
            stack.push_back((uint32_t)l + 1);
        } else {
            if ((uint64_t)nd.first_or_left + nd.count > count) {
                err = "leaf out of range";
                return false;
            }
            for (uint32_t i = nd.first_or_left; i < nd.first_or_left + nd.count; ++i) {
                if (seen[i]) {
                    err = "a primitive lies under two leaves";
                    return false;
                }
                seen[i] = 1;
            }
        }
    }
    for (uint8_t v : seen)
        if (!v) {
            err = "a primitive lies under no leaf";
            return false;
        }
    return true;
}

namespace {

// Copies tree `t` over idx[first, first + count) into b.nodes from `base`
// and refits its boxes to the (padded) primitive boxes bottom-up, so a tree
// built in another space (an object's space before its transform) stays
// conservative.  Returns the root's index.
int64_t import_bvh2(Bvh2Builder& b, const Bvh2Tree& t, uint32_t first, int64_t base)
{
    const size_t nn = t.nodes.size();
    for (size_t j = 0; j < t.prims.size(); ++j) b.idx[first + j] = first + t.prims[j];
    for (size_t i = 0; i < nn; ++i) {
        const Bvh2Node& src = t.nodes[i];
        Node2& d = b.nodes[(size_t)(base + (int64_t)i)];
        if (src.count == 0) {
            d.left = (int32_t)(base + src.first_or_left);
            d.count = 0;
            d.axis = src.axis;
        } else {
            d.left = -1;
            d.first = first + src.first_or_left;
            d.count = src.count;
            d.span = src.count;
        }
    }
    // Leaves above kMaxLeafTris triangles (the reference's builder makes a few
    // when its split fails) become subtrees of halves.
    for (size_t i = 0; i < nn; ++i) {
        const int64_t root = base + (int64_t)i;
        if (b.nodes[(size_t)root].left >= 0 || b.nodes[(size_t)root].count <= (uint32_t)kMaxLeafTris) continue;
        std::vector<int64_t> todo{root};
        while (!todo.empty()) {
            const int64_t ni = todo.back();
            todo.pop_back();
            const uint32_t f = b.nodes[(size_t)ni].first, c = b.nodes[(size_t)ni].count;
            if (c <= (uint32_t)kMaxLeafTris) continue;
            const int64_t l = b.alloc_pair();
            const uint32_t h = c / 2;
            b.make_leaf(l, f, h);
            b.make_leaf(l + 1, f + h, c - h);
            Node2& nd = b.nodes[(size_t)ni];
            nd.left = (int32_t)l;
            nd.count = 0;
            nd.axis = 0;
            todo.push_back(l);
            todo.push_back(l + 1);
        }
    }
    // Post-order refit (explicit stack: the tree may be deep).
    std::vector<std::pair<int64_t, bool>> st{{base, false}};
    while (!st.empty()) {
        auto [ni, expanded] = st.back();
        st.pop_back();
        Node2& nd = b.nodes[(size_t)ni];
        if (nd.left < 0) {
            nd.box.reset();
            for (uint32_t i = nd.first; i < nd.first + nd.count; ++i) nd.box.grow(b.pbox[b.idx[i]]);
            continue;
        }
        if (!expanded) {
            st.push_back({ni, true});
            st.push_back({nd.left, false});
            st.push_back({nd.left + 1, false});
            continue;
        }
        const Node2& l = b.nodes[(size_t)nd.left];
        const Node2& r = b.nodes[(size_t)nd.left + 1];
        nd.box = l.box;
        nd.box.grow(r.box);
        nd.first = std::min(l.first, r.first);
        // A subtree whose triangles are not one contiguous run is never
        // collapsed into a single leaf.
        const bool contiguous = (l.first + l.span == r.first) || (r.first + r.span == l.first);
        nd.span = contiguous ? l.span + r.span : 0xffffffffu;
    }
    return base;
}

}  // namespace

static bool collapse_wide(const float* tris, int64_t n, const int32_t* ids, const BuildParams& params,
                          Bvh2Builder& b, int64_t n2, WideBVH& out, std::string& err,
                          std::chrono::steady_clock::time_point t0);

bool build_wide_bvh(const float* tris, int64_t n, const int32_t* ids, const BuildParams& params,
                    WideBVH& out, std::string& err)
{
    const auto t0 = std::chrono::steady_clock::now();
    out = WideBVH();
    if (!check_params(params, n, err)) return false;
    if (n == 0) {
        empty_scene(out);
        return true;
    }
    Bvh2Builder b(params);
    if (!prepare_prims(tris, n, b, out, err)) return false;
    b.nodes.resize((size_t)(2 * n + 1));
    b.spawn_budget = std::max(0, 4 * threads_of(params));
    b.build(0, 0, (uint32_t)n);
    return collapse_wide(tris, n, ids, params, b, b.next_node.load(), out, err, t0);
}

bool build_bvh2(const float* tris, int64_t n, const BuildParams& params, Bvh2Tree& out, float scene_lo[3],
                float scene_hi[3], std::string& err)
{
    out = Bvh2Tree();
    if (!check_params(params, n, err)) return false;
    if (n == 0) return true;
    Bvh2Builder b(params);
    WideBVH tmp;
    if (!prepare_prims(tris, n, b, tmp, err)) return false;
    for (int k = 0; k < 3; ++k) {
        scene_lo[k] = tmp.scene_lo[k];
        scene_hi[k] = tmp.scene_hi[k];
    }
    b.nodes.resize((size_t)(2 * n + 1));
    b.spawn_budget = std::max(0, 4 * threads_of(params));
    b.build(0, 0, (uint32_t)n);
    const int64_t n2 = b.next_node.load();
    // Node 0 is the root and pairs are allocated from 1, so the layout is
    // already the reference's (children at left, left + 1).
    out.nodes.resize((size_t)n2);
    for (int64_t i = 0; i < n2; ++i) {
        const Node2& s = b.nodes[(size_t)i];
        Bvh2Node& d = out.nodes[(size_t)i];
        for (int k = 0; k < 3; ++k) {
            d.lo[k] = s.box.lo[k];
            d.hi[k] = s.box.hi[k];
        }
        d.first_or_left = s.left >= 0 ? (uint32_t)s.left : s.first;
        d.count = s.left >= 0 ? 0 : (uint16_t)s.count;
        d.axis = s.left >= 0 ? s.axis : 0;
        d.pad = 0;
    }
    out.prims = b.idx;
    return true;
}

bool build_wide_bvh_parts(const float* tris, int64_t n, const int32_t* ids, const std::vector<Bvh2Part>& parts,
                          const BuildParams& params, WideBVH& out, std::string& err)
{
    const auto t0 = std::chrono::steady_clock::now();
    out = WideBVH();
    if (!check_params(params, n, err)) return false;
    int64_t covered = 0;
    size_t extra = 0;
    for (const Bvh2Part& p : parts) {
        if (p.first != covered || p.count < 0) {
            err = "parts must cover the triangles in order";
            return false;
        }
        covered += p.count;
        if (p.tree) extra += p.tree->nodes.size();
    }
    if (covered != n) {
        err = "parts must cover the triangles in order";
        return false;
    }
    if (n == 0) {
        empty_scene(out);
        return true;
    }
    Bvh2Builder b(params);
    if (!prepare_prims(tris, n, b, out, err)) return false;
    // Room for SAH-built parts, imported trees, leaves split on import and the top level.
    b.nodes.resize((size_t)(4 * n + 1) + extra + 2 * parts.size());
    b.spawn_budget = std::max(0, 4 * threads_of(params));
    // Subtree of every non-empty part: imported or built.
    std::vector<int64_t> roots;
    for (const Bvh2Part& p : parts) {
        if (p.count == 0) continue;
        if (p.tree) {
            if (!check_bvh2(*p.tree, (uint32_t)p.count, err)) {
                err = "binary tree of triangles " + std::to_string(p.first) + "+" + std::to_string(p.count) + ": " + err;
                return false;
            }
            const int64_t base = b.next_node.fetch_add((int64_t)p.tree->nodes.size());
            roots.push_back(import_bvh2(b, *p.tree, (uint32_t)p.first, base));
            out.stats.imported_tris += p.count;
        } else {
            const int64_t r = b.next_node.fetch_add(1);
            b.build(r, (uint32_t)p.first, (uint32_t)(p.first + p.count));
            roots.push_back(r);
        }
    }
    // Top level over the part roots: median splits by box centre on the
    // widest axis (there are few parts); a single part becomes the root.
    struct Top {
        Bvh2Builder& b;
        void make(int64_t ni, std::vector<int64_t> list)
        {
            if (list.size() == 1) {
                b.nodes[(size_t)ni] = b.nodes[(size_t)list[0]];
                return;
            }
            Box cb, box;
            cb.reset();
            box.reset();
            for (int64_t r : list) {
                const Box& x = b.nodes[(size_t)r].box;
                float c[3] = {0.5f * x.lo[0] + 0.5f * x.hi[0], 0.5f * x.lo[1] + 0.5f * x.hi[1],
                              0.5f * x.lo[2] + 0.5f * x.hi[2]};
                cb.grow(c);
                box.grow(x);
            }
            int a = 0;
            const float e0 = cb.hi[0] - cb.lo[0], e1 = cb.hi[1] - cb.lo[1], e2 = cb.hi[2] - cb.lo[2];
            if (e1 > e0) a = 1;
            if (e2 > std::max(e0, e1)) a = 2;
            const size_t mid = list.size() / 2;
            std::nth_element(list.begin(), list.begin() + (std::ptrdiff_t)mid, list.end(), [&](int64_t x, int64_t y) {
                const Box& bx = b.nodes[(size_t)x].box;
                const Box& by = b.nodes[(size_t)y].box;
                const float cx = bx.lo[a] + bx.hi[a], cy = by.lo[a] + by.hi[a];
                return cx < cy || (cx == cy && x < y);
            });
            const int64_t l = b.alloc_pair();
            std::vector<int64_t> left(list.begin(), list.begin() + (std::ptrdiff_t)mid);
            std::vector<int64_t> right(list.begin() + (std::ptrdiff_t)mid, list.end());
            Node2& nd = b.nodes[(size_t)ni];
            nd.box = box;
            nd.left = (int32_t)l;
            nd.count = 0;
            nd.first = 0;
            nd.span = 0xffffffffu;
            nd.axis = (uint8_t)a;
            make(l, std::move(left));
            make(l + 1, std::move(right));
        }
    };
    Top{b}.make(0, roots);
    return collapse_wide(tris, n, ids, params, b, b.next_node.load(), out, err, t0);
}

static bool collapse_wide(const float* tris, int64_t n, const int32_t* ids, const BuildParams& params,
                          Bvh2Builder& b, int64_t n2, WideBVH& out, std::string& err,
                          std::chrono::steady_clock::time_point t0)
{
    b.nodes.resize((size_t)n2);
    {
        // Over the nodes reachable from the root (a tree assembled from parts
        // leaves the copied part roots behind).
        const double ra = b.nodes[0].box.area();
        double c = 0;
        int64_t reach = 0;
        std::vector<int32_t> st{0};
        while (!st.empty()) {
            const Node2& nd = b.nodes[(size_t)st.back()];
            st.pop_back();
            ++reach;
            c += (nd.left < 0 ? params.sah_intersect * nd.count : params.sah_traversal) * nd.box.area();
            if (nd.left >= 0) {
                st.push_back(nd.left);
                st.push_back(nd.left + 1);
            }
        }
        out.stats.sah_cost = ra > 0 ? c / ra : 0;
        out.stats.bvh2_nodes = reach;
    }

    // Collapse to 8-wide nodes in breadth-first order.
    struct Pending {
        int32_t n2;
        uint32_t n8;
        int level;
    };
    std::vector<Pending> queue;
    queue.reserve((size_t)n2 / 4 + 2);
    out.nodes.reserve((size_t)n2 / 4 + 2);
    out.tris.reserve((size_t)n);
    out.nodes.emplace_back();
    queue.push_back({0, 0, 1});
    int max_level = 1;
    // One wide node from binary node pd.n2 into onodes[pd.n8]; its internal
    // children get contiguous slots appended to onodes and are queued on q,
    // its leaf triangles are appended to otris.
    auto emit = [&](const Pending& pd, std::vector<NodeSlot>& onodes, std::vector<Tri48>& otris,
                    std::vector<Pending>& q, int& lvl, std::string& emsg) -> bool {
        lvl = std::max(lvl, pd.level);
        int32_t kids[8];
        int nk = 0;
        const Node2& root2 = b.nodes[pd.n2];
        if (root2.left < 0) {
            kids[nk++] = pd.n2;  // a leaf root: one leaf child
        } else {
            kids[nk++] = root2.left;
            kids[nk++] = root2.left + 1;
            while (nk < 8) {
                int pick = -1;
                double pa = -1;
                for (int c = 0; c < nk; ++c) {
                    const Node2& cn = b.nodes[kids[c]];
                    if (cn.left >= 0 && cn.box.area() > pa) {
                        pa = cn.box.area();
                        pick = c;
                    }
                }
                if (pick < 0) break;
                const int32_t l = b.nodes[kids[pick]].left;
                kids[pick] = l;
                kids[nk++] = l + 1;
            }
        }

        // Octant slot assignment: greedily give each slot the child whose
        // centroid lies furthest toward that slot's octant direction.
        const Box& nb = root2.box;
        float ctr[3];
        for (int k = 0; k < 3; ++k) ctr[k] = 0.5f * nb.lo[k] + 0.5f * nb.hi[k];
        int slot_of[8];
        int child_in[8];
        for (int s = 0; s < 8; ++s) child_in[s] = -1;
        for (int c = 0; c < nk; ++c) slot_of[c] = -1;
        for (int round = 0; round < nk; ++round) {
            double best = -std::numeric_limits<double>::infinity();
            int bc = -1, bs = -1;
            for (int c = 0; c < nk; ++c) {
                if (slot_of[c] >= 0) continue;
                const Box& cb = b.nodes[kids[c]].box;
                double d[3];
                for (int k = 0; k < 3; ++k) d[k] = (0.5 * cb.lo[k] + 0.5 * cb.hi[k]) - ctr[k];
                for (int s = 0; s < 8; ++s) {
                    if (child_in[s] >= 0) continue;
                    const double v = ((s & 1) ? -d[0] : d[0]) + ((s & 2) ? -d[1] : d[1]) + ((s & 4) ? -d[2] : d[2]);
                    if (v > best) {
                        best = v;
                        bc = c;
                        bs = s;
                    }
                }
            }
            slot_of[bc] = bs;
            child_in[bs] = bc;
        }

        Node8 node;
        std::memset(&node, 0, sizeof(node));
        // Internal children get contiguous node indices in slot order; leaf
        // triangles are appended contiguously in slot order.
        uint32_t child_base = (uint32_t)onodes.size();
        uint32_t tri_base = (uint32_t)otris.size();
        int n_internal = 0;
        uint32_t tri_off = 0;
        float clo[3][8], chi[3][8];
        int slot_list[8];
        int ns = 0;
        for (int s = 0; s < 8; ++s) {
            const int c = child_in[s];
            if (c < 0) continue;
            const Node2& cn = b.nodes[kids[c]];
            for (int k = 0; k < 3; ++k) {
                clo[k][ns] = cn.box.lo[k];
                chi[k][ns] = cn.box.hi[k];
            }
            slot_list[ns++] = s;
            // A small subtree that would become a wide node of its own is
            // stored as one leaf of all its triangles instead.
            const bool as_leaf = cn.left < 0 || (cn.span <= (uint32_t)params.collapse_leaf && tri_off + cn.span <= 32);
            const uint32_t leaf_n = cn.left < 0 ? cn.count : cn.span;
            if (!as_leaf) {
                node.imask |= (uint8_t)(1u << s);
                node.meta[s] = 1;
                ++n_internal;
            } else {
                if (leaf_n < 1 || leaf_n > (uint32_t)kMaxLeafCount || tri_off + leaf_n > 32) {
                    emsg = "internal error: leaf layout exceeds node limits";
                    return false;
                }
                node.meta[s] = (uint8_t)((leaf_n << 5) | tri_off);
                for (uint32_t t = 0; t < leaf_n; ++t) {
                    const uint32_t p = b.idx[cn.first + t];
                    const float* f = tris + 9 * (size_t)p;
                    Tri48 tr;
                    for (int k = 0; k < 3; ++k) {
                        tr.v0[k] = f[k];
                        tr.v1[k] = f[3 + k];
                        tr.v2[k] = f[6 + k];
                    }
                    const int32_t id = ids ? ids[p] : (int32_t)p;
                    std::memcpy(&tr.v0[3], &id, 4);
                    tr.v1[3] = 0.0f;
                    tr.v2[3] = 0.0f;
                    otris.push_back(tr);
                }
                tri_off += leaf_n;
            }
        }
        node.child_base = child_base;
        node.tri_base = tri_base;
        // Internal children: allocate in slot order.
        for (int s = 0; s < 8; ++s) {
            const int c = child_in[s];
            if (c < 0 || !(node.imask & (1u << s))) continue;  // empty or stored as a leaf
            const uint32_t idx8 = (uint32_t)onodes.size();
            onodes.emplace_back();
            q.push_back({kids[c], idx8, pd.level + 1});
        }
        (void)n_internal;

        // Quantize child boxes relative to the node box.
        Box qb;
        qb.reset();
        for (int i = 0; i < ns; ++i)
            for (int k = 0; k < 3; ++k) {
                qb.lo[k] = std::min(qb.lo[k], clo[k][i]);
                qb.hi[k] = std::max(qb.hi[k], chi[k][i]);
            }
        float pp[3];
        int8_t ee[3];
        uint8_t ql[3][8], qh[3][8];
        for (int k = 0; k < 3; ++k) {
            pp[k] = qb.lo[k];
            const double span = (double)qb.hi[k] - qb.lo[k];
            int e = -126;
            if (span > 0) {
                int ex;
                std::frexp(span / 255.0, &ex);  // span/255 < 2^ex
                e = std::max(-126, ex);
            }
            for (;;) {
                if (e > 127) {
                    emsg = "internal error: child box quantization overflow";
                    return false;
                }
                if (quantize_axis(pp[k], e, clo[k], chi[k], ns, ql[k], qh[k])) break;
                ++e;
            }
            ee[k] = (int8_t)e;
        }
        node.px = pp[0];
        node.py = pp[1];
        node.pz = pp[2];
        node.ex = ee[0];
        node.ey = ee[1];
        node.ez = ee[2];
        for (int s = 0; s < 8; ++s) {
            node.qlox[s] = node.qloy[s] = node.qloz[s] = 255;
            node.qhix[s] = node.qhiy[s] = node.qhiz[s] = 0;
        }
        for (int i = 0; i < ns; ++i) {
            const int s = slot_list[i];
            node.qlox[s] = ql[0][i];
            node.qloy[s] = ql[1][i];
            node.qloz[s] = ql[2][i];
            node.qhix[s] = qh[0][i];
            node.qhiy[s] = qh[1][i];
            node.qhiz[s] = qh[2][i];
        }
        static_cast<Node8&>(onodes[pd.n8]) = node;
        return true;
    };
    // Breadth-first from the root until the frontier holds kFrontier subtrees,
    // then every frontier subtree is collapsed on its own (in parallel) and
    // the pieces are appended in frontier order: the layout depends only on
    // the tree, never on the thread count.
    constexpr size_t kFrontier = 256;
    size_t qi = 0;
    for (; qi < queue.size() && queue.size() - qi < kFrontier; ++qi) {
        const Pending pd = queue[qi];
        if (!emit(pd, out.nodes, out.tris, queue, max_level, err)) return false;
    }
    const size_t n_items = queue.size() - qi;
    if (n_items > 0) {
        struct Piece {
            std::vector<NodeSlot> nodes;
            std::vector<Tri48> tris;
            int level = 1;
            std::string err;
            bool ok = true;
        };
        std::vector<Piece> pieces(n_items);
        std::atomic<size_t> next_item{0};
        auto work = [&]() {
            for (size_t k; (k = next_item.fetch_add(1)) < n_items;) {
                Piece& P = pieces[k];
                std::vector<Pending> lq{{queue[qi + k].n2, 0u, queue[qi + k].level}};
                P.nodes.emplace_back();
                for (size_t j = 0; j < lq.size() && P.ok; ++j) {
                    const Pending pd = lq[j];
                    P.ok = emit(pd, P.nodes, P.tris, lq, P.level, P.err);
                }
            }
        };
        const int nt = std::max(1, std::min(threads_of(params), (int)n_items));
        std::vector<std::thread> th;
        for (int t = 1; t < nt; ++t) th.emplace_back(work);
        work();
        for (auto& t : th) t.join();
        for (size_t k = 0; k < n_items; ++k) {
            Piece& P = pieces[k];
            if (!P.ok) {
                err = P.err;
                return false;
            }
            // Local node 0 takes the slot its parent reserved; local j >= 1
            // goes to nbase + j - 1 (a child is never local node 0).
            const uint32_t nbase = (uint32_t)out.nodes.size(), tbase = (uint32_t)out.tris.size();
            for (size_t j = 0; j < P.nodes.size(); ++j) {
                Node8 nd = static_cast<const Node8&>(P.nodes[j]);
                nd.child_base = nbase + nd.child_base - 1u;
                nd.tri_base += tbase;
                if (j == 0) {
                    static_cast<Node8&>(out.nodes[queue[qi + k].n8]) = nd;
                } else {
                    out.nodes.emplace_back();
                    static_cast<Node8&>(out.nodes.back()) = nd;
                }
            }
            out.tris.insert(out.tris.end(), P.tris.begin(), P.tris.end());
            max_level = std::max(max_level, P.level);
        }
    }
    if (max_level > kMaxStack) {
        err = "BVH too deep for the traversal stack (" + std::to_string(max_level) + " levels)";
        return false;
    }
    out.stats.nodes = (int64_t)out.nodes.size();
    out.stats.tris = (int64_t)out.tris.size();
    out.stats.depth = max_level;
    out.stats.build_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return true;
}

}  // namespace rise_hip
