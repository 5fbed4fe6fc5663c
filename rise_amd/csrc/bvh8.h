// bvh8.h -- device layout of the 8-wide BVH with quantized child boxes.
//
// The reference traverses a 4-wide SoA BVH of full-precision float boxes
// (128 B/node, Acceleration/BVH.h:124-137).  Here each node stores its eight
// child boxes as 8-bit offsets from a float origin with a power-of-two scale
// per axis (80 B/node, 10 B per child), so one wave64 lane reads a whole node
// with five 16-byte loads and tests eight children.  Boxes are rounded
// outward during quantization, so the decoded box always contains the exact
// one (the reference's 1-ULP outward rounding, BVH.h:1094-1102, plays the
// same role).
#pragma once

#include <cstdint>

namespace rise_hip {

// 80 bytes, 16-byte aligned.
struct alignas(16) Node8 {
    float px, py, pz;      // quantization origin (decoded = fma(q, 2^e, p))
    int8_t ex, ey, ez;     // per-axis scale exponents
    uint8_t imask;         // bit s set: slot s holds an internal child
    uint32_t child_base;   // node index of the first internal child
    uint32_t tri_base;     // triangle index of the first leaf triangle
    uint8_t meta[8];       // per slot: 0 empty; internal: 1; leaf: (count << 5) | offset
    uint8_t qlox[8], qloy[8], qloz[8];
    uint8_t qhix[8], qhiy[8], qhiz[8];
};
static_assert(sizeof(Node8) == 80, "Node8 must be 80 bytes");

// Nodes are stored RISE_NODE_STRIDE bytes apart.  80 packs them; 128 starts
// every node on a 128-byte cache line.  Measured on MI355X both run at the same
// speed on the 1M-triangle workload, so the packed layout is the default.
#ifndef RISE_NODE_STRIDE
#define RISE_NODE_STRIDE 80
#endif
#if RISE_NODE_STRIDE == 80
using NodeSlot = Node8;
#else
struct alignas(RISE_NODE_STRIDE) NodeSlot : Node8 {
};
#endif
static_assert(sizeof(NodeSlot) == RISE_NODE_STRIDE, "node stride");

// 48 bytes: three float4 loads.  w of v0 carries the caller's triangle id
// (bit pattern of an int32), w of v1 the mesh or object index, w of v2 flags
// (kTriNoShadow: the object casts no shadow; kTriDoubleSided: the mesh is
// double-sided; kTriNormals / kTriCoords: the triangle's per-corner vertex
// normals / texture coordinates are in the scene's attribute array).
constexpr unsigned kTriDoubleSided = 2u;
constexpr unsigned kTriNormals = 4u;
constexpr unsigned kTriCoords = 8u;
constexpr int kAttrDoubles = 15;  // per leaf record: 3 normals, 3 texture coordinates
struct alignas(16) Tri48 {
    float v0[4];
    float v1[4];
    float v2[4];
};
static_assert(sizeof(Tri48) == 48, "Tri48 must be 48 bytes");

// Slot s of a node holds the child lying furthest toward octant direction
// ((s&1)?-x:+x, (s&2)?-y:+y, (s&4)?-z:+z) from the node centre.  A ray whose
// direction sign bits form octant o visits slot (k ^ (~o & 7)) at step k, which
// is near-to-far to first order.
constexpr int kMaxLeafTris = 4;   // leaves of the binary build (SAH)
constexpr int kMaxLeafCount = 7;  // triangles of one wide-node leaf slot (3 bits of meta)
constexpr int kMaxStack = 64;  // node-group stack entries; depth is checked at build

}  // namespace rise_hip
