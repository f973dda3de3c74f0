// Closest-hit traversal as a per-lane step machine, run by persistent waves
// that refill idle lanes with new rays.
//
// One loop iteration of a wave = one step per active lane, and every step
// is one 64-byte fetch plus the arithmetic on it: a four-wide node visit
// (decode and slab-test its quantized child boxes, order the hits by entry
// distance), one triangle of the lane's current leaf, entering an
// instance (its inverse transform and BVH) or leaving one after a hit (its
// model transform, for the world-space t).  All lanes issue their fetch
// together at the top of the step, so a wave whose lanes hold different
// kinds of work waits for one memory round trip per iteration, not one per
// kind.  A hit leaf child becomes the lane's leaf range (the nearest child
// first, farther ones wait on the stack like inner nodes).  A lane whose
// ray finishes takes the next ray from its wave's chunk of the ray list, so
// the wave's SIMD lanes stay busy instead of idling behind the wave's longest
// ray.  The two levels of a scene -- instance hierarchy (TLAS) and mesh
// hierarchy (BLAS) -- are walked in the same loop: entering an instance
// transforms the ray into mesh space (sp_RayIntersectScene,
// src/sp_scene.cpp:263-276) and the mesh walk continues on the same stack.
//
// Stack: the top entries of each lane live in LDS (4 or 8 KB per wave);
// deeper entries spill to a global array (rare: the builder bounds
// tree depth, and a closest-hit walk holds at most three pending children
// per wide level).
#pragma once

#include "spg_internal.h"

namespace spg {

// Rays a wave takes from the shared counter with one atomic: the scene walk
// of path-tracing rays is fastest at 64 (128: same, 256: -8% on the 4M-
// triangle box scene), the single-mesh walk of incoherent rays at 128
// (+23% over 64 on the 1M-triangle soup: with 64, the 65,536 same-address
// atomics of a 2^22-ray launch serialize in L2; 256: -2% against 128).
constexpr uint32_t kSceneChunk = 64, kMeshChunk = 128;
// A wave refills its idle lanes once kRefill of them wait (or all do): each
// refill stalls the wave on the new rays' loads and setup, so refilling one
// lane per iteration adds a memory round trip to nearly every step, while
// waiting for all 64 leaves the wave running its longest ray alone.  The
// top kShort entries of each lane's stack live in LDS (kShort * 256 bytes
// per wave), deeper ones spill to global memory.  Measured with four-wide
// nodes: the scene walk (TLAS + BLAS, path-tracing rays of the 4M-triangle
// box scene) is fastest at 40 / 16 (48: -4%, 56: -11%; 24 or 32 LDS
// entries: -1%); the single-mesh walk of random rays (the 1M-triangle soup,
// deeper stacks, longer rays) at 56 / 32 (+4% and +10% over 48 / 16).
constexpr uint32_t kSceneRefill = 40, kMeshRefill = 56;
constexpr int kSceneShortStack = 16, kMeshShortStack = 32;
constexpr int kMinShortStack = 16;  // sizes the spill area
constexpr int32_t kNoRef = kEmptyChild;

// Node, triangle and stack-spill arrays are reached through pointers read
// from memory (mesh and scene tables), so the compiler cannot prove them
// global and would emit flat loads, which also hold the LDS wait counter
// and serialize with the LDS stack.  These loads name the global address
// space explicitly (a native vector type, so no HIP vector constructor
// takes the address back to the generic space).
typedef float f4v __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(1))) f4v gf4v;
typedef __attribute__((address_space(1))) int gint;
typedef __attribute__((address_space(3))) int lint;

__device__ __forceinline__ float4 ld_g4(const float4 *p)
{
    f4v v = *(gf4v *)p;
    return make_float4(v.x, v.y, v.z, v.w);
}

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

template <int kShort>
struct LaneStack
{
    lint *lds;       // this wave's LDS block, entry i of lane l at lds[i*64 + l]
    gint *spill;     // entry i >= kShort at spill[(i-kShort)*stride + gl]
    uint32_t stride; // lanes in the grid
    uint32_t gl;     // this lane's index in the grid
    uint32_t lane;
    __device__ __forceinline__ void write(int i, int v)
    {
        if (i < kShort) lds[i * 64 + lane] = v;
        else spill[(size_t)(i - kShort) * stride + gl] = v;
    }
    __device__ __forceinline__ int read(int i)
    {
        int v;
        if (i < kShort) v = lds[i * 64 + lane];
        else v = spill[(size_t)(i - kShort) * stride + gl];
        return v;
    }
};

// The instance level of one scene as a lane sees it (batched launches mix
// scenes within a wave).
struct SceneView
{
    const float4 *tlasNodes;
    const DevInstance *inst;
};

__device__ __forceinline__ SceneView scene_view(const DevScene &S)
{
    return SceneView{S.tlasNodes, S.inst};
}

struct Trav
{
    v3 o, d, inv;              // world ray
    v3 lo, ld, linv;           // ray in the current mesh's space
    const float4 *bnodes;      // the mesh's records: nodes, then triangles
    float lim;                 // bound on the mesh-space t from the world best
    Hit h;                     // best hit in the current mesh
    SceneHit out;              // best hit in the scene
    int node, sp, spBase, level;
    uint32_t inst;
    uint32_t triNext, triEnd;  // pending triangles of the current leaf
    uint32_t mode;             // the next step's work (kMode*)
    bool mmW, mmL;             // world / mesh-space ray is finite
    bool anyHit;               // occlusion query: stop at the first instance
                               // hit with a positive world t
};

// Kinds of step.
constexpr uint32_t kModeNode = 0;    // visit node `node` of the current level
constexpr uint32_t kModeTri = 1;     // test triangle triNext of the leaf range
constexpr uint32_t kModeEnter = 2;   // enter instance `inst`
constexpr uint32_t kModeFinish = 3;  // leave instance `inst` with its hit

// Moller-Trumbore on one fetched triangle record (a.xyz, e1.x) (e1.yz,
// e2.xy) (e2.z, prim, n.xy) (n.z, ...) with n = cross(e1, e2) stored
// (src/ray_intersection.cpp:156-190).  Returns whether the hit c beats
// `best`: t > 0 and smaller (src/sp_scene.cpp:189-193); equal t goes to the
// lower triangle index so the result does not depend on the visit order.
__device__ __forceinline__ bool triangle_candidate(float4 t0, float4 t1, float4 t2,
                                                   float4 t3, v3 o, v3 d,
                                                   const Hit &best, Hit &c)
{
    v3 a{t0.x, t0.y, t0.z}, e1v{t0.w, t1.x, t1.y}, e2v{t1.z, t1.w, t2.x};
    v3 n{t2.z, t2.w, t3.x};
    c.prim = __float_as_uint(t2.y);
    c.u = c.v = 0.0f;
    c.t = mt_test_n(o, d, a, e1v, e2v, n, &c.u, &c.v);
    return c.t > 0.0f && (c.t < best.t || (c.t == best.t && c.prim < best.prim));
}

__device__ __forceinline__ m34 rows3(float4 a, float4 b, float4 c)
{
    return m34{{{a.x, a.y, a.z, a.w}, {b.x, b.y, b.z, b.w}, {c.x, c.y, c.z, c.w}}};
}

// One step of an active lane (an inactive lane's state is left as it is).
// Returns true when the ray's walk is complete.  kTwoLevel = false walks a
// single mesh (level stays 1, spBase 0; only node and triangle steps).
//
// Loads and arithmetic sit in branches; the lane state is written through
// selects after them, and the step runs on every lane with `active` in its
// predicates.  Written inside divergent branches, every state value needs a
// second register at each join and the compiler copies the whole state
// around every branch of the step.
template <bool kCount, bool kTwoLevel, int kShort>
__device__ __forceinline__ bool trav_step(const SceneView &sv, Trav &s,
                                         LaneStack<kShort> &st, Counts &cnt,
                                         bool active)
{
    const uint32_t mode = s.mode;
    const bool blas = !kTwoLevel || s.level != 0;

    // (1) the step's 64-byte fetch: a node or triangle record of the mesh,
    // a TLAS node, instance bytes 48..111 (enter) or 0..63 (leave)
    const float4 *src;
    if (!kTwoLevel || mode <= kModeTri)
        src = (blas ? s.bnodes : sv.tlasNodes) +
              4 * (mode == kModeTri ? (int)s.triNext : s.node);
    else
        src = reinterpret_cast<const float4 *>(sv.inst + s.inst) +
              (mode == kModeEnter ? 3 : 0);
    float4 A, B, C, D;  // read only by active lanes
    if (active)
    {
        A = ld_g4(src);
        B = ld_g4(src + 1);
        C = ld_g4(src + 2);
        D = ld_g4(src + 3);
    }

    // (2) the work on it
    const bool isTri = active && mode == kModeTri;
    const bool isNode = active && mode == kModeNode;
    const bool isEnter = kTwoLevel && active && mode == kModeEnter;
    const bool isFin = kTwoLevel && active && mode == kModeFinish;
    Hit c{0.0f, 0.0f, 0.0f, 0u};
    bool take = false;
    int32_t next = kNoRef;
    // pending children of a node visit, stored at sp, sp + 1, sp + 2
    int32_t push0 = kNoRef, push1 = kNoRef, push2 = kNoRef;
    int npush = 0;
    if (isTri)
        take = triangle_candidate(A, B, C, D, s.lo, s.ld, s.h, c);
    else if (isNode)
    {
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
        // A: origin p and scale exponents; B: child references; C, D: the
        // child boxes, one field of all four children per 32-bit word.  A
        // coordinate decodes to p + q * s: q * s is exact (8-bit integer
        // times a power of two), so the FMA rounds the same sum once.
        const uint32_t eb = __float_as_uint(A.w);
        const float sx = __uint_as_float((eb & 0xFFu) << 23);
        const float sy = __uint_as_float(((eb >> 8) & 0xFFu) << 23);
        const float sz = __uint_as_float(((eb >> 16) & 0xFFu) << 23);
        const uint32_t q0 = __float_as_uint(C.x), q1 = __float_as_uint(C.y),
                       q2 = __float_as_uint(C.z), q3 = __float_as_uint(C.w),
                       q4 = __float_as_uint(D.x), q5 = __float_as_uint(D.y);
        int32_t ref[4] = {__float_as_int(B.x), __float_as_int(B.y),
                          __float_as_int(B.z), __float_as_int(B.w)};
        auto dq = [](uint32_t w, int k, float sc, float p) {
            return fmaf((float)((w >> (8 * k)) & 0xFFu), sc, p);
        };
        // a hit child's key: the entry distance's bits with the sign
        // cleared (-0 = +0) and the two low bits replaced by the child slot,
        // so keys are distinct and near-ties go to the lower slot; a missed
        // or absent child gets the largest key
        auto keyOf = [&](int k, bool h, float e) {
            return (h & (e <= tc) & (ref[k] != kEmptyChild))
                       ? ((__float_as_uint(e) & 0x7FFFFFFCu) | (uint32_t)k)
                       : 0xFFFFFFFFu;
        };
        uint32_t key[4];
        // a plane at p + q s (q the 8-bit field) is crossed at
        // t = (p + q s - o) inv = q (s inv) + (p - o) inv: per node and axis
        // a = s inv (exact: s is a power of two) and b = (p - o) inv, then
        // one FMA per plane
        const float ax = sx * inv.x, ay = sy * inv.y, az = sz * inv.z;
        const float bx = (A.x - o.x) * inv.x, by = (A.y - o.y) * inv.y,
                    bz = (A.z - o.z) * inv.z;
        const bool fold = (blas ? s.mmL : s.mmW) && finite3(v3{ax, ay, az}) &&
                          finite3(v3{bx, by, bz});
        {
            // per axis the near plane is the lo field where inv > 0 (a
            // positive a makes t increase with q) and the hi field where
            // inv < 0, so choosing the fields once per node orders each
            // axis's two distances
            const bool px = inv.x > 0.0f, py = inv.y > 0.0f, pz = inv.z > 0.0f;
            const uint32_t nx = px ? q0 : q3, fx = px ? q3 : q0;
            const uint32_t ny = py ? q1 : q4, fy = py ? q4 : q1;
            const uint32_t nz = pz ? q2 : q5, fz = pz ? q5 : q2;
#pragma unroll
            for (int k = 0; k < 4; ++k)
            {
                float e;
                bool h = slab_hit_t(dq(nx, k, ax, bx), dq(ny, k, ay, by), dq(nz, k, az, bz),
                                    dq(fx, k, ax, bx), dq(fy, k, ay, by), dq(fz, k, az, bz),
                                    &e);
                key[k] = keyOf(k, h, e);
            }
        }
        if (!fold)
        {
            // a non-finite origin, inverse direction or per-axis factor (an
            // axis-parallel ray, or s inv overflowing): the reference's
            // compare-and-select form on the decoded box
#pragma unroll
            for (int k = 0; k < 4; ++k)
            {
                float e;
                bool h = slab_hit(dq(q0, k, sx, A.x), dq(q1, k, sy, A.y),
                                  dq(q2, k, sz, A.z), dq(q3, k, sx, A.x),
                                  dq(q4, k, sy, A.y), dq(q5, k, sz, A.z), o, inv, &e);
                key[k] = keyOf(k, h, e);
            }
        }
        if (kCount)
        {
            uint32_t present = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) present += ref[k] != kEmptyChild ? 1u : 0u;
            if (blas) cnt.box += present;
            else cnt.tlasBox += present;
        }
        // sort (key, reference) pairs by key: five compare-exchanges
        auto cx = [&](int a, int b) {
            const bool sw = key[b] < key[a];
            const uint32_t lo = sw ? key[b] : key[a], hi = sw ? key[a] : key[b];
            const int32_t ra = sw ? ref[b] : ref[a], rb = sw ? ref[a] : ref[b];
            key[a] = lo;
            key[b] = hi;
            ref[a] = ra;
            ref[b] = rb;
        };
        cx(0, 1);
        cx(2, 3);
        cx(0, 2);
        cx(1, 3);
        cx(1, 2);
        // the nearest is next; the others wait on the stack, the nearest
        // of them on top: slot 3 at sp, slot 2 above it, slot 1 on top
        const bool v1 = key[1] != 0xFFFFFFFFu, v2 = key[2] != 0xFFFFFFFFu,
                   v3 = key[3] != 0xFFFFFFFFu;
        next = key[0] != 0xFFFFFFFFu ? ref[0] : kNoRef;
        push0 = v3 ? ref[3] : v2 ? ref[2] : ref[1];
        push1 = v3 ? ref[2] : ref[1];
        push2 = ref[1];
        npush = (v1 ? 1 : 0) + (v2 ? 1 : 0) + (v3 ? 1 : 0);
    }
    s.h.t = take ? c.t : s.h.t;
    s.h.u = take ? c.u : s.h.u;
    s.h.v = take ? c.v : s.h.v;
    s.h.prim = take ? c.prim : s.h.prim;
    s.triNext += isTri ? 1u : 0u;
    if (npush > 0) st.write(s.sp, push0);
    if (npush > 1) st.write(s.sp + 1, push1);
    if (npush > 2) st.write(s.sp + 2, push2);
    s.sp += npush;
    const bool busy = isTri && s.triNext < s.triEnd;

    bool stop = false;  // an occlusion query found a hit
    if (kTwoLevel)
    {
        // entering: o' = M^-1 o, d' = norm(M^-1 d) (src/sp_scene.cpp:272-276);
        // the mesh walk starts at its root above the current stack
        v3 nlo{0.0f, 0.0f, 0.0f}, nld{0.0f, 0.0f, 0.0f}, nlinv{0.0f, 0.0f, 0.0f};
        float nlim = 0.0f;
        bool nmm = false;
        const float4 *nb = nullptr;
        if (isEnter)
        {
            const m34 I = rows3(A, B, C);
            nlo = xform_point(I, s.o);
            nld = normalize(xform_vector(I, s.d));
            nlinv = inverse(nld);
            nmm = finite3(nlo) && finite3(nlinv);
            nlim = s.out.t > 0.0f ? s.out.t * D.x * kCullSlack : INFINITY;
            nb = reinterpret_cast<const float4 *>(
                ((uint64_t)__float_as_uint(D.w) << 32) | __float_as_uint(D.z));
        }
        s.lo = v3{isEnter ? nlo.x : s.lo.x, isEnter ? nlo.y : s.lo.y,
                  isEnter ? nlo.z : s.lo.z};
        s.ld = v3{isEnter ? nld.x : s.ld.x, isEnter ? nld.y : s.ld.y,
                  isEnter ? nld.z : s.ld.z};
        s.linv = v3{isEnter ? nlinv.x : s.linv.x, isEnter ? nlinv.y : s.linv.y,
                    isEnter ? nlinv.z : s.linv.z};
        s.lim = isEnter ? nlim : s.lim;
        s.mmL = isEnter ? nmm : s.mmL;
        s.bnodes = isEnter ? nb : s.bnodes;
        s.h.t = isEnter ? INFINITY : s.h.t;
        s.h.prim = isEnter ? SPG_MISS : s.h.prim;
        s.spBase = isEnter ? s.sp : s.spBase;
        if (kCount) cnt.mesh += isEnter ? 1u : 0u;

        // leaving with a hit: world t = dot(M (o' + t' d') - o, d), the
        // closest world hit wins (src/sp_scene.cpp:296-322), equal t goes to
        // the lower instance index
        bool better = false;
        float tw = 0.0f;
        if (isFin)
        {
            v3 lh = s.lo + s.ld * s.h.t;
            v3 wh = xform_point(rows3(A, B, C), lh);
            tw = dot(wh - s.o, s.d);
            // an occlusion query takes the first instance hit at a positive
            // world t (what k_resolve_shadow counts as a blocker) and stops;
            // a hit whose world t rounds to <= 0 neither stops nor blocks
            better = s.anyHit ? tw > 0.0f
                              : tw < s.out.t || s.out.t < 0.0f ||
                                    (tw == s.out.t && s.inst < s.out.inst);
        }
        stop = s.anyHit & better;
        s.out.t = better ? tw : s.out.t;
        s.out.u = better ? s.h.u : s.out.u;
        s.out.v = better ? s.h.v : s.out.v;
        s.out.inst = better ? s.inst : s.out.inst;
        s.out.prim = better ? s.h.prim : s.out.prim;
        s.level = isEnter ? 1 : isFin ? 0 : s.level;
    }

    // (3) the next work.  Nothing to descend into: leave a finished mesh
    // walk (at once without a hit, by a leave step with one), then pop.  A
    // popped reference is never empty, so one pop gives the next step work.
    bool needPop = (isNode && next == kNoRef) || (isTri && !busy) || isFin;
    bool leaveStep = false;
    if (kTwoLevel)
    {
        const bool leave = needPop && s.level != 0 && s.sp <= s.spBase;
        leaveStep = leave && s.h.prim != SPG_MISS;
        s.level = leave && !leaveStep ? 0 : s.level;
        needPop = needPop && !leaveStep;
    }
    const bool done = (needPop && s.sp == 0) || stop;
    const bool doPop = needPop && s.sp != 0;
    int32_t popped = kNoRef;
    if (doPop) popped = st.read(s.sp - 1);
    s.sp -= doPop ? 1 : 0;
    next = doPop ? popped : next;

    const bool leaf = d_is_leaf(next);
    const bool enter = kTwoLevel && leaf && s.level == 0;
    const bool range = leaf && !enter;
    const bool inner = !leaf && next != kNoRef;
    const uint32_t first = d_leaf_first(next), count = d_leaf_count(next);
    s.triNext = range ? first : s.triNext;
    s.triEnd = range ? first + count : s.triEnd;
    if (kCount) cnt.tri += range ? count : 0u;
    s.inst = enter ? first : s.inst;
    s.node = inner ? next : isEnter ? 0 : s.node;
    s.mode = leaveStep ? kModeFinish
             : enter   ? kModeEnter
             : range   ? kModeTri
             : (inner || isEnter) ? kModeNode
             : busy    ? kModeTri
                       : s.mode;
    return done;
}

__device__ __forceinline__ void trav_begin_scene(Trav &s, v3 o, v3 d)
{
    s.o = o;
    s.d = d;
    s.inv = inverse(d);
    s.mmW = finite3(o) && finite3(s.inv);
    s.mmL = false;
    s.anyHit = false;
    s.out = SceneHit{-1.0f, 0.0f, 0.0f, SPG_MISS, SPG_MISS};
    s.level = 0;
    s.sp = 0;
    s.spBase = 0;
    s.node = 0;
    s.h = Hit{INFINITY, 0.0f, 0.0f, SPG_MISS};
    s.triNext = s.triEnd = 0;
    s.mode = kModeNode;
}

__device__ __forceinline__ void trav_begin_mesh(const DevMesh &M, Trav &s, v3 o,
                                                v3 d)
{
    s.lo = o;
    s.ld = d;
    s.linv = inverse(d);
    s.mmL = finite3(o) && finite3(s.linv);
    s.mmW = false;
    s.anyHit = false;
    s.bnodes = M.nodes;
    s.lim = INFINITY;
    s.h = Hit{INFINITY, 0.0f, 0.0f, SPG_MISS};
    s.level = 1;
    s.sp = 0;
    s.spBase = 0;
    s.node = 0;
    s.triNext = s.triEnd = 0;
    s.mode = kModeNode;
}

// Persistent driver: every wave walks rays [0, n) taken in chunks of
// kChunk from *counter (zero on entry), refilling its idle lanes once
// `refill` (1..64, wave-uniform) of them wait.  begin(ray, s) starts a ray
// and returns false if it needs no walk (it is then stored at once);
// store(ray, s) writes the result.  Must be called by every lane of a
// one-wave block.
template <bool kCount, bool kTwoLevel, int kShort, uint32_t kChunk, class Begin,
          class Store>
__device__ __forceinline__ void persistent_walk(SceneView &sv, uint32_t n,
                                                uint32_t *counter, int *spill,
                                                uint32_t refill, Begin begin,
                                                Store store, Counts &cnt,
                                                uint32_t &rays)
{
    __shared__ int lds[kShort * 64];
    const uint32_t lane = threadIdx.x;
    LaneStack<kShort> st{(lint *)lds, (gint *)spill, gridDim.x * 64u,
                         blockIdx.x * 64u + lane,
                 lane};
    const unsigned long long below = (1ull << lane) - 1ull;
    uint32_t chunkNext = 0, chunkEnd = 0;
    bool exhausted = false;
    uint32_t ray = SPG_MISS;
    Trav s;
    for (;;)
    {
        bool need = ray == SPG_MISS;
        unsigned long long needMask = __ballot(need);
        const uint32_t idle = (uint32_t)__popcll(needMask);
        if ((idle >= refill || idle == 64u) && idle != 0u &&
            !(exhausted && chunkNext >= chunkEnd))
        {
            if (chunkNext >= chunkEnd)
            {
                uint32_t base = 0;
                if (lane == 0) base = atomicAdd(counter, kChunk);
                // wave-uniform (a scalar register), so the loop's exit is
                // uniform and the compiler keeps no per-lane copies of the
                // lane state across it
                base = __builtin_amdgcn_readfirstlane(base);
                if (base >= n)
                {
                    exhausted = true;
                    chunkNext = chunkEnd = 0;
                }
                else
                {
                    chunkNext = base;
                    chunkEnd = base + kChunk < n ? base + kChunk : n;
                    if (base + kChunk >= n) exhausted = true;
                }
            }
            uint32_t want = idle;
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
        // one path round the loop: with no active lane the step's branches
        // are all skipped, and the lane state needs no copies for a second
        // path back to the loop head
        if (__ballot(ray != SPG_MISS) == 0ull && exhausted && chunkNext >= chunkEnd)
            break;
        if (trav_step<kCount, kTwoLevel, kShort>(sv, s, st, cnt, ray != SPG_MISS))
        {
            store(ray, s);
            ray = SPG_MISS;
            rays++;
        }
    }
}

}  // namespace spg
