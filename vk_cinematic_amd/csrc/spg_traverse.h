// Closest-hit traversal as a per-lane step machine, run by persistent waves
// that refill idle lanes with new rays.
//
// One loop iteration of a wave = one step per active lane: either a BVH2
// node visit (both child slab tests) or one triangle of the lane's current
// leaf.  A hit leaf child becomes the lane's leaf range (the nearer child
// first, a farther one waits on the stack like an inner node), so an
// iteration costs one node visit plus one triangle test however the lanes
// are mixed, instead of a node visit plus a whole leaf.  A lane whose ray
// finishes takes the next ray from its wave's chunk of the ray list, so the
// wave's SIMD lanes stay busy instead of idling behind the wave's longest
// ray.  The two levels of a scene -- instance hierarchy (TLAS) and mesh
// hierarchy (BLAS) -- are walked in the same loop: entering an instance
// transforms the ray into mesh space (sp_RayIntersectScene,
// src/sp_scene.cpp:263-276) and the mesh walk continues on the same stack.
//
// Stack: the top kShortStack entries of each lane live in LDS (4 KB per
// wave); deeper entries spill to a global array (rare: the builder bounds
// tree depth, and a closest-hit walk holds at most one pending far child per
// level).
#pragma once

#include "spg_internal.h"

namespace spg {

constexpr int kShortStack = 16;
constexpr uint32_t kRayChunk = 64;
constexpr int32_t kNoRef = kEmptyChild;

__device__ __forceinline__ bool d_is_leaf(int32_t r) { return r < 0; }
__device__ __forceinline__ uint32_t d_leaf_first(int32_t r)
{
    return ((uint32_t)~r) >> 3;
}
__device__ __forceinline__ uint32_t d_leaf_count(int32_t r)
{
    return (((uint32_t)~r) & 7u) + 1u;
}

struct Hit
{
    float t, u, v;
    uint32_t prim;
};

struct SceneHit
{
    float t, u, v;
    uint32_t inst, prim;
};

struct Counts
{
    uint32_t box, tlasBox, tri, mesh;
};

struct LaneStack
{
    int *lds;        // this wave's LDS block, entry i of lane l at lds[i*64 + l]
    int *spill;      // entry i >= kShortStack at spill[(i-kShortStack)*stride + gl]
    uint32_t stride; // lanes in the grid
    uint32_t gl;     // this lane's index in the grid
    uint32_t lane;
    __device__ __forceinline__ void push(int &sp, int v)
    {
        if (sp < kShortStack) lds[sp * 64 + lane] = v;
        else spill[(size_t)(sp - kShortStack) * stride + gl] = v;
        ++sp;
    }
    __device__ __forceinline__ int pop(int &sp)
    {
        --sp;
        return sp < kShortStack ? lds[sp * 64 + lane]
                                : spill[(size_t)(sp - kShortStack) * stride + gl];
    }
};

// The instance level of one scene as a lane sees it (batched launches mix
// scenes within a wave).
struct SceneView
{
    const float4 *tlasNodes;
    const uint32_t *tlasInst;
    const DevInstance *inst;
    const DevMesh *meshes;
};

__device__ __forceinline__ SceneView scene_view(const DevScene &S)
{
    return SceneView{S.tlasNodes, S.tlasInst, S.inst, S.meshes};
}

struct Trav
{
    v3 o, d, inv;              // world ray
    v3 lo, ld, linv;           // ray in the current mesh's space
    const float4 *bnodes, *btris;
    float lim;                 // bound on the mesh-space t from the world best
    Hit h;                     // best hit in the current mesh
    SceneHit out;              // best hit in the scene
    int node, sp, spBase, level;
    uint32_t inst;
    uint32_t triNext, triEnd;  // pending triangles of the current leaf
};

// One triangle (leaf order index k): Moller-Trumbore
// (src/ray_intersection.cpp:156-190), keep t > 0 and smaller
// (src/sp_scene.cpp:189-193); equal t goes to the lower triangle index so
// the result does not depend on the visit order.
__device__ __forceinline__ void test_triangle(const float4 *tris, uint32_t k, v3 o,
                                              v3 d, Hit &best)
{
    const float4 *T = tris + 3 * k;
    float4 t0 = T[0], t1 = T[1], t2 = T[2];
    v3 a{t0.x, t0.y, t0.z}, e1v{t0.w, t1.x, t1.y}, e2v{t1.z, t1.w, t2.x};
    uint32_t prim = __float_as_uint(t2.y);
    float u, v;
    float t = mt_test(o, d, a, e1v, e2v, &u, &v);
    if (t > 0.0f && (t < best.t || (t == best.t && prim < best.prim)))
    {
        best.t = t;
        best.u = u;
        best.v = v;
        best.prim = prim;
    }
}

// Enter the instance of a TLAS leaf: o' = M^-1 o, d' = norm(M^-1 d)
// (src/sp_scene.cpp:272-276); the mesh walk starts above the current stack.
template <bool kCount>
__device__ __forceinline__ void enter_instance(const SceneView &sv, Trav &s,
                                               int32_t ref, Counts &cnt)
{
    uint32_t ii = sv.tlasInst[d_leaf_first(ref)];
    const DevInstance &I = sv.inst[ii];
    s.lo = xform_point(I.inv, s.o);
    s.ld = normalize(xform_vector(I.inv, s.d));
    s.linv = inverse(s.ld);
    s.lim = s.out.t > 0.0f ? s.out.t * I.cullScale * kCullSlack : INFINITY;
    const DevMesh &M = sv.meshes[I.mesh];
    s.bnodes = M.nodes;
    s.btris = M.tris;
    s.h = Hit{INFINITY, 0.0f, 0.0f, SPG_MISS};
    s.level = 1;
    s.spBase = s.sp;
    s.node = 0;
    s.inst = ii;
    s.triNext = s.triEnd = 0;
    if (kCount) cnt.mesh += 1;
}

// Leave the instance: world t = dot(M (o' + t' d') - o, d) and the closest
// world hit wins (src/sp_scene.cpp:296-322); equal t goes to the lower
// instance index.
__device__ __forceinline__ void finish_instance(const SceneView &sv, Trav &s)
{
    if (s.h.prim == SPG_MISS) return;
    const DevInstance &I = sv.inst[s.inst];
    v3 lh = s.lo + s.ld * s.h.t;
    v3 wh = xform_point(I.model, lh);
    float t = dot(wh - s.o, s.d);
    if (t < s.out.t || s.out.t < 0.0f || (t == s.out.t && s.inst < s.out.inst))
    {
        s.out.t = t;
        s.out.u = s.h.u;
        s.out.v = s.h.v;
        s.out.inst = s.inst;
        s.out.prim = s.h.prim;
    }
}

// One step (a node visit or one leaf triangle).  Returns true when the
// ray's walk is complete.  kTwoLevel = false walks a single mesh (level
// stays 1, spBase 0).
template <bool kCount, bool kTwoLevel>
__device__ __forceinline__ bool trav_step(const SceneView &sv, Trav &s,
                                         LaneStack &st, Counts &cnt)
{
    int32_t next = kNoRef;
    if (s.triNext < s.triEnd)
    {
        test_triangle(s.btris, s.triNext, s.lo, s.ld, s.h);
        if (++s.triNext < s.triEnd) return false;
    }
    else
    {
        const bool blas = !kTwoLevel || s.level != 0;
        const float4 *N = (blas ? s.bnodes : sv.tlasNodes) + 4 * s.node;
        v3 o = blas ? s.lo : s.o;
        v3 inv = blas ? s.linv : s.inv;
        float tc;
        if (blas)
        {
            tc = s.h.t * kCullSlack;
            tc = tc < s.lim ? tc : s.lim;
        }
        else
            tc = s.out.t > 0.0f ? s.out.t * kCullSlack : INFINITY;
        float4 A = N[0], B = N[1], C = N[2], D = N[3];
        int32_t r0 = __float_as_int(A.w), r1 = __float_as_int(B.w);
        float e0, e1;
        bool h0 = slab_hit(A.x, A.y, A.z, B.x, B.y, B.z, o, inv, &e0) && e0 <= tc;
        bool h1 = r1 != kEmptyChild &&
                  slab_hit(C.x, C.y, C.z, D.x, D.y, D.z, o, inv, &e1) && e1 <= tc;
        if (kCount)
        {
            uint32_t n = r1 != kEmptyChild ? 2u : 1u;
            if (blas) cnt.box += n;
            else cnt.tlasBox += n;
        }
        if (h0 && h1)
        {
            int32_t nearN = r0, farN = r1;
            if (e1 < e0)
            {
                nearN = r1;
                farN = r0;
            }
            st.push(s.sp, farN);
            next = nearN;
        }
        else if (h0)
            next = r0;
        else if (h1)
            next = r1;
    }
    for (;;)
    {
        if (next != kNoRef)
        {
            if (d_is_leaf(next))
            {
                if (kTwoLevel && s.level == 0)
                {
                    enter_instance<kCount>(sv, s, next, cnt);
                    return false;
                }
                s.triNext = d_leaf_first(next);
                s.triEnd = s.triNext + d_leaf_count(next);
                if (kCount) cnt.tri += d_leaf_count(next);
                return false;
            }
            s.node = next;
            return false;
        }
        if (!kTwoLevel)
        {
            if (s.sp == 0) return true;
            next = st.pop(s.sp);
            continue;
        }
        if (s.level != 0)
        {
            if (s.sp > s.spBase)
            {
                next = st.pop(s.sp);
                continue;
            }
            finish_instance(sv, s);
            s.level = 0;
        }
        if (s.sp == 0) return true;
        next = st.pop(s.sp);
    }
}

__device__ __forceinline__ void trav_begin_scene(Trav &s, v3 o, v3 d)
{
    s.o = o;
    s.d = d;
    s.inv = inverse(d);
    s.out = SceneHit{-1.0f, 0.0f, 0.0f, SPG_MISS, SPG_MISS};
    s.level = 0;
    s.sp = 0;
    s.spBase = 0;
    s.node = 0;
    s.h = Hit{INFINITY, 0.0f, 0.0f, SPG_MISS};
    s.triNext = s.triEnd = 0;
}

__device__ __forceinline__ void trav_begin_mesh(const DevMesh &M, Trav &s, v3 o,
                                                v3 d)
{
    s.lo = o;
    s.ld = d;
    s.linv = inverse(d);
    s.bnodes = M.nodes;
    s.btris = M.tris;
    s.lim = INFINITY;
    s.h = Hit{INFINITY, 0.0f, 0.0f, SPG_MISS};
    s.level = 1;
    s.sp = 0;
    s.spBase = 0;
    s.node = 0;
    s.triNext = s.triEnd = 0;
}

// Persistent driver: every wave walks rays [0, n) taken in chunks of
// kRayChunk from *counter (zero on entry).  begin(ray, s) starts a ray and
// returns false if it needs no walk (it is then stored at once);
// store(ray, s) writes the result.  Must be called by every lane of a
// one-wave block.
template <bool kCount, bool kTwoLevel, class Begin, class Store>
__device__ __forceinline__ void persistent_walk(SceneView &sv, uint32_t n,
                                                uint32_t *counter, int *spill,
                                                Begin begin, Store store,
                                                Counts &cnt, uint32_t &rays)
{
    __shared__ int lds[kShortStack * 64];
    const uint32_t lane = threadIdx.x;
    LaneStack st{lds, spill, gridDim.x * 64u, blockIdx.x * 64u + lane, lane};
    const unsigned long long below = (1ull << lane) - 1ull;
    uint32_t chunkNext = 0, chunkEnd = 0;
    bool exhausted = false;
    uint32_t ray = SPG_MISS;
    Trav s;
    for (;;)
    {
        bool need = ray == SPG_MISS;
        unsigned long long needMask = __ballot(need);
        if (needMask && !(exhausted && chunkNext >= chunkEnd))
        {
            if (chunkNext >= chunkEnd)
            {
                uint32_t base = 0;
                if (lane == 0) base = atomicAdd(counter, kRayChunk);
                base = __shfl(base, 0);
                if (base >= n)
                {
                    exhausted = true;
                    chunkNext = chunkEnd = 0;
                }
                else
                {
                    chunkNext = base;
                    chunkEnd = base + kRayChunk < n ? base + kRayChunk : n;
                    if (base + kRayChunk >= n) exhausted = true;
                }
            }
            uint32_t want = (uint32_t)__popcll(needMask);
            uint32_t avail = chunkEnd - chunkNext;
            uint32_t rank = (uint32_t)__popcll(needMask & below);
            if (need && rank < avail)
            {
                ray = chunkNext + rank;
                if (!begin(ray, s))
                {
                    store(ray, s);
                    ray = SPG_MISS;
                    rays++;
                }
            }
            chunkNext += want < avail ? want : avail;
        }
        if (__ballot(ray != SPG_MISS) == 0ull)
        {
            if (exhausted && chunkNext >= chunkEnd) break;
            continue;
        }
        if (ray != SPG_MISS && trav_step<kCount, kTwoLevel>(sv, s, st, cnt))
        {
            store(ray, s);
            ray = SPG_MISS;
            rays++;
        }
    }
}

}  // namespace spg
