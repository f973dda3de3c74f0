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

namespace {

struct WideChild
{
    float lo[3], hi[3];
    int32_t ref;
    double area() const
    {
        double dx = (double)hi[0] - lo[0], dy = (double)hi[1] - lo[1],
               dz = (double)hi[2] - lo[2];
        return 2.0 * (dx * dy + dy * dz + dz * dx);
    }
};

// The two child slots of a BVH2 node (its padded boxes).
void bvh2_children(const GpuNode &g, std::vector<WideChild> &out)
{
    WideChild a;
    memcpy(a.lo, g.a, 12);
    memcpy(a.hi, g.b, 12);
    memcpy(&a.ref, &g.a[3], 4);
    out.push_back(a);
    int32_t r1;
    memcpy(&r1, &g.b[3], 4);
    if (r1 != kEmptyChild)
    {
        WideChild b;
        memcpy(b.lo, g.c, 12);
        memcpy(b.hi, g.d, 12);
        b.ref = r1;
        out.push_back(b);
    }
}

// Outward quantization of the children's boxes on axis k; false if the
// boxes are not finite.
bool quantize_axis(WideNode &w, const std::vector<WideChild> &kids, int k)
{
    float lo = INFINITY, hi = -INFINITY;
    for (const WideChild &c : kids)
    {
        if (!std::isfinite(c.lo[k]) || !std::isfinite(c.hi[k])) return false;
        lo = std::min(lo, c.lo[k]);
        hi = std::max(hi, c.hi[k]);
    }
    w.p[k] = lo;
    const double ext = (double)hi - (double)lo;
    int e = 1;  // scale 2^-126 at least (normal, so q * s is exact)
    while (e < 254 && 255.0 * std::ldexp(1.0, e - 127) < ext) ++e;
    for (; e < 255; ++e)
    {
        w.e[k] = (uint8_t)e;
        const float s = wide_scale(w.e[k]);
        bool ok = true;
        for (size_t c = 0; c < kids.size() && ok; ++c)
        {
            double ql = std::floor(((double)kids[c].lo[k] - lo) / s);
            double qh = std::ceil(((double)kids[c].hi[k] - lo) / s);
            long a = (long)std::max(0.0, std::min(255.0, ql));
            long b = (long)std::max(0.0, std::min(256.0, qh));
            // decoded as the traversal decodes: p + q * s in FP32
            while (a > 0 && lo + (float)a * s > kids[c].lo[k]) --a;
            while (b <= 255 && lo + (float)b * s < kids[c].hi[k]) ++b;
            if (b > 255 || lo + (float)a * s > kids[c].lo[k]) ok = false;
            else
            {
                w.q[k][c] = (uint8_t)a;
                w.q[3 + k][c] = (uint8_t)b;
            }
        }
        if (ok) return true;
    }
    return false;
}

struct WideBuilder
{
    const std::vector<GpuNode> &in;
    std::vector<WideNode> &out;
    bool ok = true;

    // returns the node's index; pending: its subtree's stack bound
    uint32_t emit(uint32_t n2, uint32_t &pending)
    {
        std::vector<WideChild> kids;
        bvh2_children(in[n2], kids);
        while (kids.size() < 4)
        {
            int best = -1;
            double bestArea = -1.0;
            for (size_t c = 0; c < kids.size(); ++c)
                if (kids[c].ref >= 0 && kids[c].area() > bestArea)
                {
                    best = (int)c;
                    bestArea = kids[c].area();
                }
            if (best < 0) break;
            std::vector<WideChild> sub;
            bvh2_children(in[(size_t)kids[best].ref], sub);
            if (kids.size() + sub.size() - 1 > 4) break;
            kids[best] = sub[0];
            for (size_t j = 1; j < sub.size(); ++j)
                kids.insert(kids.begin() + best + (long)j, sub[j]);
        }
        const uint32_t me = (uint32_t)out.size();
        out.push_back(WideNode{});
        WideNode w{};
        for (int c = 0; c < 4; ++c) w.ref[c] = kEmptyChild;
        for (int k = 0; k < 3 && ok; ++k) ok = quantize_axis(w, kids, k);
        uint32_t below = 0;
        for (size_t c = 0; c < kids.size(); ++c)
        {
            uint32_t sub = 0;
            w.ref[c] = kids[c].ref >= 0 ? (int32_t)emit((uint32_t)kids[c].ref, sub)
                                        : kids[c].ref;
            below = std::max(below, sub);
        }
        pending = (uint32_t)kids.size() - 1 + below;
        out[me] = w;
        return me;
    }
};

}  // namespace

bool collapse_wide(const BvhBuildResult &bvh2, std::vector<WideNode> &out,
                   uint32_t &maxPending)
{
    out.clear();
    maxPending = 0;
    if (bvh2.nodes.empty()) return false;
    out.reserve(bvh2.nodes.size() / 2 + 1);
    WideBuilder wb{bvh2.nodes, out};
    wb.emit(0, maxPending);
    return wb.ok;
}

}  // namespace spg
