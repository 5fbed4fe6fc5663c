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

// Decoded child box of slot s (exactly the arithmetic the kernels use).
void decode_child_box(const Node8& n, int s, float lo[3], float hi[3]);

}  // namespace rise_hip
