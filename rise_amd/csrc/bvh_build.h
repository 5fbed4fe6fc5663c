// bvh_build.h -- host-side build of the 8-wide quantized BVH.
//
// Stage 1 is a top-down binned-SAH BVH2 in the spirit of the reference's
// builder (Acceleration/BVH.h:772-1034: binned SAH, unit traversal and
// intersection costs, median fallback), evaluated on all three axes and run
// in parallel over subtrees.  Stage 2 collapses it into 8-wide nodes (the
// reference collapses to 4-wide, BVH.h:477-642), assigns child slots by
// octant, and quantizes child boxes outward to 8 bits per bound.
#pragma once

#include "bvh8.h"

#include <cstdint>
#include <string>
#include <vector>

#ifndef RISE_COLLAPSE_LEAF
#define RISE_COLLAPSE_LEAF 0
#endif

namespace rise_hip {

struct BuildParams {
    int bins = 32;              // SAH bins per axis (reference: 32)
    int max_leaf = kMaxLeafTris;
    // Collapse: a BVH2 subtree of at most this many triangles that would
    // become a wide node of its own is stored as one leaf (<= kMaxLeafTris,
    // so a node's leaves always fit its 32 triangle slots).  0: never.  On
    // the 1M-triangle soup, 2 and 3 cut node visits by 3% and 10% but add 39%
    // and 170% FP64 triangle tests, and trace 4% and 22% slower on MI355X.
    int collapse_leaf = RISE_COLLAPSE_LEAF;
    float sah_traversal = 1.0f;  // reference mesh costs (TriangleMeshGeometryIndexed.cpp:529-534)
    float sah_intersect = 1.0f;
    int threads = 0;            // 0 = hardware concurrency
};

struct BuildStats {
    int64_t bvh2_nodes = 0;
    int64_t nodes = 0;
    int64_t tris = 0;
    int depth = 0;             // wide-node levels (root = 1)
    double build_ms = 0.0;
    double sah_cost = 0.0;     // SAH cost of the BVH2, normalized by root area
    int64_t imported_tris = 0; // triangles under binary trees taken from the caller
};

struct WideBVH {
    std::vector<NodeSlot> nodes;
    std::vector<Tri48> tris;   // leaf order; v0.w carries the caller's id
    float scene_lo[3] = {0, 0, 0};
    float scene_hi[3] = {0, 0, 0};
    BuildStats stats;
};

// tris: n * 9 floats (v0, v1, v2).  ids: optional caller ids (NULL = index).
// Returns false and fills err on failure (too deep, bad input).
bool build_wide_bvh(const float* tris, int64_t n, const int32_t* ids, const BuildParams& params,
                    WideBVH& out, std::string& err);

// A binary tree in the reference's on-disk BVH2 layout (Acceleration/BVH.h:87-96,
// 32 bytes per node): node 0 is the root, an internal node's children are
// nodes first_or_left and first_or_left + 1, a leaf holds prims[first_or_left,
// first_or_left + count); prims[] are indices into the caller's triangles.
struct Bvh2Node {
    float lo[3];
    float hi[3];
    uint32_t first_or_left;
    uint16_t count;  // 0: internal
    uint8_t axis;
    uint8_t pad;
};
static_assert(sizeof(Bvh2Node) == 32, "Bvh2Node must be 32 bytes");

struct Bvh2Tree {
    std::vector<Bvh2Node> nodes;
    std::vector<uint32_t> prims;
};

// The binned-SAH binary stage alone (boxes padded as in the wide build);
// scene_lo / scene_hi receive the triangles' bounds.
bool build_bvh2(const float* tris, int64_t n, const BuildParams& params, Bvh2Tree& out, float scene_lo[3],
                float scene_hi[3], std::string& err);

// Checks a binary tree over `count` triangles: structure, and that prims is a
// permutation covered once by leaves.
bool check_bvh2(const Bvh2Tree& tree, uint32_t count, std::string& err);

// A range of the input triangles and, optionally, a binary tree over it.
struct Bvh2Part {
    int64_t first;
    int64_t count;
    const Bvh2Tree* tree;  // NULL: built by SAH
};

// Wide build whose binary stage takes the given trees for ranges of the input
// instead of building them (their boxes are refit to the triangles, so a tree
// built before an object's transform stays valid); parts without a tree are
// built by SAH, and a top level joins the parts.  parts must cover [0, n) in
// order.  Trees are checked (check_bvh2); a bad one fails the build.
bool build_wide_bvh_parts(const float* tris, int64_t n, const int32_t* ids, const std::vector<Bvh2Part>& parts,
                          const BuildParams& params, WideBVH& out, std::string& err);

// Decoded child box of slot s (exactly the arithmetic the kernels use).
void decode_child_box(const Node8& n, int s, float lo[3], float hi[3]);

}  // namespace rise_hip
