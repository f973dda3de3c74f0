// Binned-SAH BVH2 builder (Wald 2007, "On fast Construction of SAH-based
// Bounding Volume Hierarchies"), threaded over subtrees, flattened to the
// 64-byte two-box node layout the gfx950 traversal reads.
#include "spg_bvh.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <future>
#include <memory>
#include <thread>

namespace spg {
namespace {

struct Prim
{
    float lo[3], hi[3], c[3];
};

struct BNode
{
    float lo[3], hi[3];
    std::unique_ptr<BNode> kid[2];
    uint32_t first = 0, count = 0;  // count > 0: leaf
};

struct Box
{
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const float *l, const float *h)
    {
        for (int k = 0; k < 3; ++k)
        {
            lo[k] = std::min(lo[k], l[k]);
            hi[k] = std::max(hi[k], h[k]);
        }
    }
    void grow(const Box &b) { grow(b.lo, b.hi); }
    double area() const
    {
        if (!(hi[0] >= lo[0])) return 0.0;
        double dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        return 2.0 * (dx * dy + dy * dz + dz * dx);
    }
};

struct Builder
{
    const BvhBuildInput &in;
    std::vector<Prim> prims;
    std::vector<uint32_t> order;
    std::atomic<int> freeThreads{0};

    explicit Builder(const BvhBuildInput &i) : in(i) {}

    std::unique_ptr<BNode> make_leaf(uint32_t b, uint32_t e, const Box &bx)
    {
        auto n = std::make_unique<BNode>();
        memcpy(n->lo, bx.lo, 12);
        memcpy(n->hi, bx.hi, 12);
        n->first = b;
        n->count = e - b;
        return n;
    }

    static uint32_t ceil_log2(uint32_t x)
    {
        uint32_t r = 0;
        while ((1u << r) < x) ++r;
        return r;
    }

    // depth levels a balanced split of n primitives still needs below here
    uint32_t levels(uint32_t n) const
    {
        return ceil_log2((n + in.maxLeafSize - 1) / in.maxLeafSize);
    }

    std::unique_ptr<BNode> build(uint32_t b, uint32_t e, uint32_t depth)
    {
        Box bx, cb;
        for (uint32_t i = b; i < e; ++i)
        {
            const Prim &p = prims[order[i]];
            bx.grow(p.lo, p.hi);
            cb.grow(p.c, p.c);
        }
        uint32_t n = e - b;
        const uint32_t maxLeaf = in.maxLeafSize;
        if (n == 1) return make_leaf(b, e, bx);

        // SAH over K bins on each axis of the centroid bounds.
        const uint32_t K = in.bins;
        double bestCost = INFINITY;
        int bestAxis = -1;
        uint32_t bestSplit = 0;
        double parentArea = bx.area();
        std::vector<Box> binBox(K);
        std::vector<uint32_t> binCnt(K);
        std::vector<double> rightArea(K);
        std::vector<uint32_t> rightCnt(K);
        for (int axis = 0; axis < 3; ++axis)
        {
            float ext = cb.hi[axis] - cb.lo[axis];
            if (!(ext > 0.0f)) continue;
            for (uint32_t k = 0; k < K; ++k)
            {
                binBox[k] = Box();
                binCnt[k] = 0;
            }
            float scale = (float)K / ext;
            for (uint32_t i = b; i < e; ++i)
            {
                const Prim &p = prims[order[i]];
                uint32_t k = (uint32_t)std::min<float>(
                    (float)(K - 1), std::max(0.0f, (p.c[axis] - cb.lo[axis]) * scale));
                binBox[k].grow(p.lo, p.hi);
                binCnt[k]++;
            }
            Box acc;
            uint32_t cnt = 0;
            for (int k = (int)K - 1; k > 0; --k)
            {
                acc.grow(binBox[k]);
                cnt += binCnt[k];
                rightArea[k] = acc.area();
                rightCnt[k] = cnt;
            }
            acc = Box();
            cnt = 0;
            for (uint32_t k = 0; k + 1 < K; ++k)
            {
                acc.grow(binBox[k]);
                cnt += binCnt[k];
                if (cnt == 0 || rightCnt[k + 1] == 0) continue;
                double cost = 1.0 + (acc.area() * cnt +
                                     rightArea[k + 1] * rightCnt[k + 1]) /
                                        std::max(parentArea, 1e-30);
                if (cost < bestCost)
                {
                    bestCost = cost;
                    bestAxis = axis;
                    bestSplit = k;
                }
            }
        }
        if (n <= maxLeaf && (bestAxis < 0 || (double)n <= bestCost))
            return make_leaf(b, e, bx);

        uint32_t mid;
        if (bestAxis >= 0)
        {
            float ext = cb.hi[bestAxis] - cb.lo[bestAxis];
            float scale = (float)K / ext;
            auto it = std::partition(
                order.begin() + b, order.begin() + e, [&](uint32_t pi) {
                    const Prim &p = prims[pi];
                    uint32_t k = (uint32_t)std::min<float>(
                        (float)(K - 1),
                        std::max(0.0f, (p.c[bestAxis] - cb.lo[bestAxis]) * scale));
                    return k <= bestSplit;
                });
            mid = (uint32_t)(it - order.begin());
        }
        else
            mid = b;
        bool tooDeep = mid > b && mid < e &&
                       (depth + 1 + levels(mid - b) > in.maxDepth ||
                        depth + 1 + levels(e - mid) > in.maxDepth);
        if (mid == b || mid == e || tooDeep)
        {
            // Coincident centroids, or an SAH split that would exceed the
            // depth bound: split the index range in half on the longest
            // axis of the centroid box.
            int axis = 0;
            for (int k = 1; k < 3; ++k)
                if (cb.hi[k] - cb.lo[k] > cb.hi[axis] - cb.lo[axis]) axis = k;
            mid = b + n / 2;
            std::nth_element(order.begin() + b, order.begin() + mid,
                             order.begin() + e, [&](uint32_t x, uint32_t y) {
                                 if (prims[x].c[axis] != prims[y].c[axis])
                                     return prims[x].c[axis] < prims[y].c[axis];
                                 return x < y;
                             });
        }

        auto node = std::make_unique<BNode>();
        memcpy(node->lo, bx.lo, 12);
        memcpy(node->hi, bx.hi, 12);
        if (n > 16384 && freeThreads.fetch_sub(1) > 0)
        {
            auto fut = std::async(std::launch::async,
                                  [this, b, mid, depth] {
                                      return build(b, mid, depth + 1);
                                  });
            node->kid[1] = build(mid, e, depth + 1);
            node->kid[0] = fut.get();
            freeThreads.fetch_add(1);
        }
        else
        {
            if (n > 16384) freeThreads.fetch_add(1);
            node->kid[0] = build(b, mid, depth + 1);
            node->kid[1] = build(mid, e, depth + 1);
        }
        return node;
    }
};

struct Flattener
{
    BvhBuildResult &out;
    double rootArea;
    float absPad;

    void pad_box(const float *lo, const float *hi, float *plo, float *phi) const
    {
        for (int k = 0; k < 3; ++k)
        {
            float p = (fabsf(lo[k]) + fabsf(hi[k]) + (hi[k] - lo[k])) *
                          (1.0f / 1048576.0f) +
                      absPad;
            plo[k] = lo[k] - p;
            phi[k] = hi[k] + p;
        }
    }

    int32_t child_ref(const BNode *c, uint32_t depth)
    {
        if (!c) return kEmptyChild;
        if (c->count > 0)
        {
            out.leafCount++;
            out.maxDepth = std::max(out.maxDepth, depth);
            Box b;
            b.grow(c->lo, c->hi);
            out.sahCost += b.area() / rootArea * c->count;
            return leaf_ref(c->first, c->count);
        }
        return (int32_t)emit(c, depth);
    }

    uint32_t emit(const BNode *n, uint32_t depth)
    {
        uint32_t me = (uint32_t)out.nodes.size();
        out.nodes.push_back(GpuNode{});
        Box b;
        b.grow(n->lo, n->hi);
        out.sahCost += b.area() / rootArea;
        const BNode *k0 = n->kid[0].get(), *k1 = n->kid[1].get();
        int32_t r0, r1;
        if (n->count > 0)
        {
            // root that is a single leaf
            k0 = n;
            r0 = child_ref(n, depth + 1);
            r1 = kEmptyChild;
            k1 = nullptr;
        }
        else
        {
            r0 = child_ref(k0, depth + 1);
            r1 = child_ref(k1, depth + 1);
        }
        GpuNode &g = out.nodes[me];
        float lo[3], hi[3];
        pad_box(k0->lo, k0->hi, lo, hi);
        memcpy(g.a, lo, 12);
        memcpy(g.b, hi, 12);
        memcpy(&g.a[3], &r0, 4);
        memcpy(&g.b[3], &r1, 4);
        if (k1)
        {
            pad_box(k1->lo, k1->hi, lo, hi);
            memcpy(g.c, lo, 12);
            memcpy(g.d, hi, 12);
        }
        else
        {
            // absent child: NaN bounds fail the slab test (entry <= exit is
            // false for a NaN entry); traversal also checks kEmptyChild.
            for (int k = 0; k < 3; ++k)
            {
                g.c[k] = NAN;
                g.d[k] = NAN;
            }
        }
        g.c[3] = 0.0f;
        g.d[3] = 0.0f;
        return me;
    }
};

}  // namespace

bool build_bvh(const BvhBuildInput &in, BvhBuildResult &out)
{
    out = BvhBuildResult();
    if (in.count == 0) return false;
    BvhBuildInput cfg = in;
    if (cfg.maxLeafSize < 1) cfg.maxLeafSize = 1;
    if (cfg.maxLeafSize > 8) cfg.maxLeafSize = 8;
    if (cfg.bins < 4) cfg.bins = 16;
    Builder bld(cfg);
    bld.prims.resize(in.count);
    bld.order.resize(in.count);
    float gmax = 0.0f;
    for (uint32_t i = 0; i < in.count; ++i)
    {
        Prim &p = bld.prims[i];
        for (int k = 0; k < 3; ++k)
        {
            p.lo[k] = in.lo[3 * i + k];
            p.hi[k] = in.hi[3 * i + k];
            p.c[k] = 0.5f * (p.lo[k] + p.hi[k]);
            gmax = std::max(gmax, std::max(fabsf(p.lo[k]), fabsf(p.hi[k])));
        }
        bld.order[i] = i;
    }
    unsigned hw = cfg.threads ? cfg.threads : std::thread::hardware_concurrency();
    bld.freeThreads = (int)(hw > 1 ? hw - 1 : 0);
    if (cfg.maxDepth == 0) cfg.maxDepth = 64;
    std::unique_ptr<BNode> root = bld.build(0, in.count, 0);
    Box rb;
    rb.grow(root->lo, root->hi);
    Flattener fl{out, std::max(rb.area(), 1e-30), gmax * (1.0f / 16777216.0f)};
    out.nodes.reserve(2 * (size_t)in.count / std::max(1u, cfg.maxLeafSize) + 2);
    fl.emit(root.get(), 0);
    out.prims = std::move(bld.order);
    return true;
}

}  // namespace spg
