// Binned-SAH BVH2 builder (host side).
//
// Replaces the reference's greedy nearest-centroid 4-ary builder
// (bvh_CreateTree, src/bvh.cpp:51-200), which is O(n^2 log n) and whose
// 128-leaf query buffer overflows at the benchmark scene sizes.  The result
// is flattened into GpuNode records, two child boxes per node (Aila & Laine
// 2009 layout), in depth-first order so a parent's first child follows it.
#pragma once

#include <stdint.h>
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

}  // namespace spg
