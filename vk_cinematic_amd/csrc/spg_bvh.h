// Binned-SAH BVH2 builder (host side).
//
// Replaces the reference's greedy nearest-centroid 4-ary builder
// (bvh_CreateTree, src/bvh.cpp:51-200), which is O(n^2 log n) and whose
// 128-leaf query buffer overflows at the benchmark scene sizes.  The result
// is flattened into GpuNode records, two child boxes per node (Aila & Laine
// 2009 layout), in depth-first order so a parent's first child follows it.
#pragma once

#include <stdint.h>
#include <string.h>
#include <vector>

namespace spg {

// 64-byte node: both child boxes plus child references.
//   a = (c0.lo.xyz, c0 ref)   b = (c0.hi.xyz, c1 ref)
//   c = (c1.lo.xyz, 0)        d = (c1.hi.xyz, 0)
// A reference >= 0 is an inner node index; kEmptyChild is an absent child;
// otherwise ~ref = first*8 + (count-1), a leaf of `count` (1..8) triangles
// starting at `first` in the primitive order.
struct alignas(16) GpuNode
{
    float a[4], b[4], c[4], d[4];
};
static_assert(sizeof(GpuNode) == 64, "GpuNode must be 64 bytes");

constexpr int32_t kEmptyChild = 0x7FFFFFFF;

inline int32_t leaf_ref(uint32_t first, uint32_t count)
{
    return ~(int32_t)(first * 8u + (count - 1u));
}
inline bool ref_is_leaf(int32_t r) { return r < 0; }
inline uint32_t leaf_first(int32_t r) { return ((uint32_t)~r) >> 3; }
inline uint32_t leaf_count(int32_t r) { return (((uint32_t)~r) & 7u) + 1u; }

struct BvhBuildInput
{
    const float *lo;      // 3 floats per primitive
    const float *hi;      // 3 floats per primitive
    uint32_t count;
    uint32_t maxLeafSize; // 1..8
    uint32_t bins;        // per axis
    uint32_t threads;     // 0 = hardware concurrency
    uint32_t maxDepth;    // bound on inner-node depth (traversal stack)
};

struct BvhBuildResult
{
    std::vector<GpuNode> nodes;   // nodes[0] is the root (may hold one leaf)
    std::vector<uint32_t> prims;  // primitive order referenced by leaves
    uint32_t maxDepth = 0;
    uint32_t leafCount = 0;
    double sahCost = 0.0;         // C_trav = 1, C_tri = 1 per unit root area
};

// Builds the tree; boxes written to nodes are padded outward by a relative
// 2^-20 so that a ray whose Moller-Trumbore test accepts a triangle also
// passes the box tests of every node above it.  Returns false if count==0.
bool build_bvh(const BvhBuildInput &in, BvhBuildResult &out);

// 64-byte four-child node with 8-bit quantized child boxes:
//   p      quantization origin (float x3)
//   e      per axis scale s = 2^(e - 127) (byte x3)
//   ref    child references as in GpuNode (kEmptyChild: absent; present
//          children first)
//   q      field-major: q[f][c] is field f (lo.x lo.y lo.z hi.x hi.y hi.z)
//          of child c, so one 32-bit word holds one field of all four
//          children; the box is lo = p + q_lo * s, hi = p + q_hi * s in
//          FP32 (the product is exact), and contains the BVH2 child box it
//          replaces
struct alignas(16) WideNode
{
    float p[3];
    uint8_t e[3];
    uint8_t pad0;
    int32_t ref[4];
    uint8_t q[6][4];
    uint8_t pad1[8];
};
static_assert(sizeof(WideNode) == 64, "WideNode must be 64 bytes");

// Collapses a BVH2 into wide nodes: a node starts from the two children of
// a BVH2 node and opens the inner child of largest surface area until it
// has four children or only leaves (children keep their expansion order).
// Depth-first order, root 0; leaf references unchanged.  maxPending: the
// most references a closest-hit walk holds on its stack (over root-to-node
// paths, the sum of the nodes' present children less one).  Returns false
// if quantization fails (non-finite boxes).
bool collapse_wide(const BvhBuildResult &bvh2, std::vector<WideNode> &out,
                   uint32_t &maxPending);

// The decoded box of child c on one axis (what the traversal computes).
inline float wide_lo(const WideNode &n, int c, int k);
inline float wide_hi(const WideNode &n, int c, int k);

inline float wide_scale(uint8_t e)
{
    uint32_t bits = (uint32_t)e << 23;
    float s;
    memcpy(&s, &bits, 4);
    return s;
}
inline float wide_lo(const WideNode &n, int c, int k)
{
    return n.p[k] + (float)n.q[k][c] * wide_scale(n.e[k]);
}
inline float wide_hi(const WideNode &n, int c, int k)
{
    return n.p[k] + (float)n.q[3 + k][c] * wide_scale(n.e[k]);
}

}  // namespace spg
