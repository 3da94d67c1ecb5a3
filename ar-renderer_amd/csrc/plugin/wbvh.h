// 8-wide quantised BVH for the MI355X traversal kernels.
//
// Layout in HBM:
//   WNode[]  128-byte nodes, one L2 line each. Child boxes are quantised to 8 bits per axis
//            against the node's origin and power-of-two scale, rounded outwards so every decoded
//            child box contains its primitives.
//   TriRec[] 48-byte leaf records in leaf order: a triangle's p0, e1 = p1 - p0, e2 = p2 - p0
//            (the exact float operands of the reference's Moller-Trumbore test) and its
//            original primitive index; a sphere stores its centre and radius.
#pragma once

#include <cstdint>
#include <vector>

#include "prender_gpu.h"

namespace pgpu {

constexpr uint32_t kLeafBit = 0x80000000u;
constexpr int kMaxLeafPrims = 8;  // primitives per leaf child (count - 1 in 3 bits)
constexpr uint32_t kSphereBit = 0x80000000u;  // TriRec::prim flag

struct alignas(16) WNode {
    float ox, oy, oz;            // quantisation origin
    int8_t ex, ey, ez;           // scale = 2^e per axis
    uint8_t nChildren;
    uint32_t child[8];           // internal: node index; leaf: kLeafBit | (count-1) << 27 | first record
    uint8_t qlo[3][8];           // [axis][child]
    uint8_t qhi[3][8];
    uint32_t pad[8];
};
static_assert(sizeof(WNode) == 128, "WNode must be one 128-byte line");

struct alignas(16) TriRec {
    float p0x, p0y, p0z;
    float e1x, e1y, e1z;
    float e2x, e2y, e2z;
    uint32_t prim;  // original primitive index | kSphereBit for spheres (p0 = centre, e1x = radius)
    uint32_t pad[2];
};
static_assert(sizeof(TriRec) == 48, "TriRec is 48 bytes");

struct WideBVH {
    std::vector<WNode> nodes;
    std::vector<TriRec> recs;
    int depth = 0;         // maximum node depth (root = 1)
    float lo[3], hi[3];    // scene bounds
};

// Builds the wide BVH over the scene's primitives (binned SAH binary build, collapsed to 8-wide).
void BuildWideBVH(const prgpu::Primitive *prims, int count, WideBVH *out);

}  // namespace pgpu
