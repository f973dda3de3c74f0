/* ORACLE -- test infrastructure only (see spo.h).
 *
 * CPU restatement of the GPU closest-hit traversal over the product's
 * exported BVH2 (two child boxes per 64-byte node), used to check the GPU's
 * results and to count the slab and triangle tests that the roofline's
 * algorithmic bytes are computed from.  No reference counterpart: the
 * reference walks its greedy 4-ary tree without t-max culling
 * (src/bvh.cpp:203-311); the per-box slab test (src/simd.h:198-271) and the
 * Moller-Trumbore test (src/ray_intersection.cpp:156-190) are the
 * reference's. */
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include "spo_math.h"

#define EMPTY_CHILD 0x7FFFFFFF
#define CULL_SLACK (1.0f + 1.0f / 4096.0f)

static int slab(const float *lo, const float *hi, spo_v3 o, spo_v3 inv,
                float *entry)
{
    float t0x = (lo[0] - o.x) * inv.x, t1x = (hi[0] - o.x) * inv.x;
    float t0y = (lo[1] - o.y) * inv.y, t1y = (hi[1] - o.y) * inv.y;
    float t0z = (lo[2] - o.z) * inv.z, t1z = (hi[2] - o.z) * inv.z;
    float mnx = t0x < t1x ? t0x : t1x, mxx = t0x > t1x ? t0x : t1x;
    float mny = t0y < t1y ? t0y : t1y, mxy = t0y > t1y ? t0y : t1y;
    float mnz = t0z < t1z ? t0z : t1z, mxz = t0z > t1z ? t0z : t1z;
    float e = spo_max(0.0f, spo_max(mnx, spo_max(mny, mnz)));
    float x = spo_min(mxx, spo_min(mxy, mxz));
    *entry = e;
    return e <= x;
}

static int32_t ref_of(const float *f)
{
    int32_t r;
    memcpy(&r, f, 4);
    return r;
}

typedef struct {
    const float *nodes;
    const uint32_t *order;
    const float *tris;
    const float *rays;
    uint32_t n;
    uint32_t *prim;
    float *t;
    atomic_uint next;
    atomic_ullong box, tri;
} bvh2_job;

/* The GPU's step order: a step is a node visit (near child next, far child
 * pushed, whether leaf or inner) or one triangle of the current leaf; leaves
 * are tested in full once reached, triangles in leaf order. */
static void trace_one(const bvh2_job *j, uint32_t r, uint64_t *box,
                      uint64_t *tri)
{
    const float *ro = j->rays + 6 * (size_t)r;
    spo_v3 o = v3p(ro), d = v3p(ro + 3);
    spo_v3 inv = v3inv(d);
    float bestT = INFINITY;
    uint32_t bestP = 0xFFFFFFFFu;
    int32_t stack[128];
    int sp = 0;
    int32_t node = 0;
    uint32_t triNext = 0, triEnd = 0;
    for (;;)
    {
        int32_t next = EMPTY_CHILD;
        if (triNext < triEnd)
        {
            uint32_t p = j->order[triNext];
            const float *tr = j->tris + 9 * (size_t)p;
            float tt = spo_rt_mt(ro, ro + 3, tr, tr + 3, tr + 6, NULL);
            if (tt > 0.0f && (tt < bestT || (tt == bestT && p < bestP)))
            {
                bestT = tt;
                bestP = p;
            }
            if (++triNext < triEnd) continue;
        }
        else
        {
            const float *N = j->nodes + 16 * (size_t)node;
            int32_t r0 = ref_of(N + 3), r1 = ref_of(N + 7);
            float tc = bestT * CULL_SLACK;
            float e0 = 0, e1 = 0;
            int h0 = slab(N, N + 4, o, inv, &e0) && e0 <= tc;
            int h1 = r1 != EMPTY_CHILD && slab(N + 8, N + 12, o, inv, &e1) && e1 <= tc;
            *box += r1 != EMPTY_CHILD ? 2 : 1;
            if (h0 && h1)
            {
                int32_t nearN = r0, farN = r1;
                if (e1 < e0) { nearN = r1; farN = r0; }
                stack[sp++] = farN;
                next = nearN;
            }
            else if (h0) next = r0;
            else if (h1) next = r1;
        }
        if (next == EMPTY_CHILD)
        {
            if (sp == 0) break;
            next = stack[--sp];
        }
        if (next < 0)
        {
            uint32_t v = (uint32_t)~next;
            triNext = v >> 3;
            triEnd = triNext + (v & 7u) + 1u;
            *tri += (v & 7u) + 1u;
        }
        else
            node = next;
    }
    j->prim[r] = bestP;
    j->t[r] = bestP == 0xFFFFFFFFu ? -1.0f : bestT;
}

static void *bvh2_worker(void *p)
{
    bvh2_job *j = (bvh2_job *)p;
    uint64_t box = 0, tri = 0;
    for (;;)
    {
        uint32_t r0 = atomic_fetch_add(&j->next, 256u);
        if (r0 >= j->n) break;
        uint32_t r1 = r0 + 256 < j->n ? r0 + 256 : j->n;
        for (uint32_t r = r0; r < r1; ++r) trace_one(j, r, &box, &tri);
    }
    atomic_fetch_add(&j->box, box);
    atomic_fetch_add(&j->tri, tri);
    return NULL;
}

/* nodes16: the exported nodes (16 floats each, root = 0); order: the leaf
 * primitive order; tris9: triangles a,b,c by primitive id. */
void spo_bvh2_closest(const float *nodes16, const uint32_t *order,
                      const float *tris9, const float *rays6, uint32_t n,
                      uint32_t threads, uint32_t *primOut, float *tOut,
                      uint64_t *boxTests, uint64_t *triTests)
{
    bvh2_job j;
    memset(&j, 0, sizeof j);
    j.nodes = nodes16; j.order = order; j.tris = tris9; j.rays = rays6;
    j.n = n; j.prim = primOut; j.t = tOut;
    atomic_init(&j.next, 0u);
    atomic_init(&j.box, 0ull);
    atomic_init(&j.tri, 0ull);
    if (threads < 1) threads = 1;
    if (threads > 64) threads = 64;
    pthread_t th[64];
    for (uint32_t i = 0; i < threads; ++i)
        pthread_create(&th[i], NULL, bvh2_worker, &j);
    for (uint32_t i = 0; i < threads; ++i) pthread_join(th[i], NULL);
    if (boxTests) *boxTests = atomic_load(&j.box);
    if (triTests) *triTests = atomic_load(&j.tri);
}

/* Structural invariants of an exported BVH2: every primitive referenced by
 * exactly one leaf, every child box contains its subtree's triangles'
 * bounds, child indices in range and after their parent.  Returns 0 if all
 * hold, else a code naming the first violation. */
int spo_bvh2_check(const float *nodes16, uint32_t nNodes, const uint32_t *order,
                   const float *tris9, uint32_t nTris)
{
    uint8_t *seen = (uint8_t *)calloc(nTris ? nTris : 1, 1);
    int rc = 0;
    for (uint32_t i = 0; i < nNodes && !rc; ++i)
    {
        const float *N = nodes16 + 16 * (size_t)i;
        for (int c = 0; c < 2 && !rc; ++c)
        {
            int32_t r = ref_of(N + 3 + 4 * c);
            const float *lo = N + 8 * c, *hi = N + 8 * c + 4;
            if (r == EMPTY_CHILD) continue;
            if (r >= 0)
            {
                if ((uint32_t)r >= nNodes || (uint32_t)r <= i) { rc = 1; break; }
                const float *K = nodes16 + 16 * (size_t)r;
                for (int cc = 0; cc < 2; ++cc)
                {
                    if (ref_of(K + 3 + 4 * cc) == EMPTY_CHILD) continue;
                    for (int a = 0; a < 3; ++a)
                        if (!(K[8 * cc + a] >= lo[a] && K[8 * cc + 4 + a] <= hi[a]))
                            rc = 2;
                }
            }
            else
            {
                uint32_t v = (uint32_t)~r, first = v >> 3, count = (v & 7u) + 1u;
                if (first + count > nTris) { rc = 3; break; }
                for (uint32_t k = 0; k < count; ++k)
                {
                    uint32_t p = order[first + k];
                    if (p >= nTris || seen[p]) { rc = 4; break; }
                    seen[p] = 1;
                    const float *tr = tris9 + 9 * (size_t)p;
                    for (int vtx = 0; vtx < 3; ++vtx)
                        for (int a = 0; a < 3; ++a)
                            if (!(tr[3 * vtx + a] >= lo[a] && tr[3 * vtx + a] <= hi[a]))
                                rc = 5;
                }
            }
        }
    }
    for (uint32_t p = 0; p < nTris && !rc; ++p)
        if (!seen[p]) rc = 6;
    free(seen);
    return rc;
}
