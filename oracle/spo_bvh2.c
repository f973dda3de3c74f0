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
#include <math.h>
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

/* ---- the four-wide quantized hierarchy ----
 *
 * 64-byte node: bytes 0..11 origin p (3 floats), 12..14 per-axis scale
 * exponents e (s = 2^(e-127)), 16..31 four child references (as in the
 * BVH2; EMPTY_CHILD marks an absent child, absent children come last),
 * 32..55 the quantized child boxes field-major: byte 32 + 4 f + c is field
 * f (lo.x lo.y lo.z hi.x hi.y hi.z) of child c.  A child box decodes as
 * p + q * s per coordinate in FP32 (the product is exact). */

static float wide_scale(uint8_t e)
{
    uint32_t bits = (uint32_t)e << 23;
    float s;
    memcpy(&s, &bits, 4);
    return s;
}

static void wide_box(const uint8_t *N, int c, float *lo, float *hi)
{
    float p[3];
    memcpy(p, N, 12);
    for (int k = 0; k < 3; ++k)
    {
        float s = wide_scale(N[12 + k]);
        lo[k] = p[k] + (float)N[32 + 4 * k + c] * s;
        hi[k] = p[k] + (float)N[32 + 4 * (3 + k) + c] * s;
    }
}

static uint32_t f2u(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

/* One node visit: the hit children's references in visit order (nearest
 * first) and their count.  Hits are ordered by a key of the entry distance
 * with its two low bits replaced by the child slot (ties and near-ties go
 * to the lower slot); the sign bit is cleared so -0 and +0 agree. */
static int finite3f(float x, float y, float z)
{
    return isfinite(x) && isfinite(y) && isfinite(z);
}

/* The GPU's slab distances of child c: a plane at p + q s is crossed at
 * fma(q, s inv, (p - o) inv) (s inv exact, s a power of two), near and far
 * picked per axis by the sign of inv; entry clamped at 0. */
static int wide_slab_fold(const uint8_t *N, int c, const float *a, const float *b,
                          spo_v3 inv, float *entry)
{
    const float iv[3] = {inv.x, inv.y, inv.z};
    float nr[3], fr[3];
    for (int k = 0; k < 3; ++k)
    {
        float tl = fmaf((float)N[32 + 4 * k + c], a[k], b[k]);
        float th = fmaf((float)N[32 + 4 * (3 + k) + c], a[k], b[k]);
        nr[k] = iv[k] > 0.0f ? tl : th;
        fr[k] = iv[k] > 0.0f ? th : tl;
    }
    float e = fmaxf(fmaxf(0.0f, nr[0]), fmaxf(nr[1], nr[2]));
    float x = fminf(fminf(fr[0], fr[1]), fr[2]);
    *entry = e;
    return e <= x;
}

static int wide_visit(const uint8_t *N, spo_v3 o, spo_v3 inv, float tc,
                      int32_t *refs, uint64_t *box)
{
    int32_t ref[4];
    memcpy(ref, N + 16, 16);
    uint32_t key[4];
    int m = 0;
    float p[3], a[3], b[3];
    memcpy(p, N, 12);
    const float iv[3] = {inv.x, inv.y, inv.z}, ov[3] = {o.x, o.y, o.z};
    for (int k = 0; k < 3; ++k)
    {
        a[k] = wide_scale(N[12 + k]) * iv[k];
        b[k] = (p[k] - ov[k]) * iv[k];
    }
    /* finite ray and per-axis factors: the GPU's folded form; otherwise
     * the decoded box and the reference's compare-and-select test */
    const int fold = finite3f(o.x, o.y, o.z) && finite3f(inv.x, inv.y, inv.z) &&
                     finite3f(a[0], a[1], a[2]) && finite3f(b[0], b[1], b[2]);
    for (int c = 0; c < 4; ++c)
    {
        key[c] = 0xFFFFFFFFu;
        if (ref[c] == EMPTY_CHILD) continue;
        ++*box;
        float lo[3], hi[3], e = 0.0f;
        int hit;
        if (fold)
            hit = wide_slab_fold(N, c, a, b, inv, &e);
        else
        {
            wide_box(N, c, lo, hi);
            hit = slab(lo, hi, o, inv, &e);
        }
        if (hit && e <= tc)
        {
            key[c] = (f2u(e) & 0x7FFFFFFCu) | (uint32_t)c;
            ++m;
        }
    }
    for (int i = 1; i < 4; ++i)  /* insertion sort, keys are distinct */
        for (int j = i; j > 0 && key[j] < key[j - 1]; --j)
        {
            uint32_t t = key[j]; key[j] = key[j - 1]; key[j - 1] = t;
        }
    for (int i = 0; i < m; ++i) refs[i] = ref[key[i] & 3u];
    return m;
}

static void trace_one_wide(const bvh2_job *j, uint32_t r, uint64_t *box,
                           uint64_t *tri)
{
    const float *ro = j->rays + 6 * (size_t)r;
    spo_v3 o = v3p(ro), d = v3p(ro + 3);
    spo_v3 inv = v3inv(d);
    const uint8_t *nodes = (const uint8_t *)j->nodes;
    float bestT = INFINITY;
    uint32_t bestP = 0xFFFFFFFFu;
    int32_t stack[256];
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
            int32_t refs[4];
            int m = wide_visit(nodes + 64 * (size_t)node, o, inv,
                               bestT * CULL_SLACK, refs, box);
            /* the rest wait on the stack, the nearest of them on top */
            for (int i = m - 1; i >= 1; --i) stack[sp++] = refs[i];
            if (m > 0) next = refs[0];
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

static void *wide_worker(void *p)
{
    bvh2_job *j = (bvh2_job *)p;
    uint64_t box = 0, tri = 0;
    for (;;)
    {
        uint32_t r0 = atomic_fetch_add(&j->next, 256u);
        if (r0 >= j->n) break;
        uint32_t r1 = r0 + 256 < j->n ? r0 + 256 : j->n;
        for (uint32_t r = r0; r < r1; ++r) trace_one_wide(j, r, &box, &tri);
    }
    atomic_fetch_add(&j->box, box);
    atomic_fetch_add(&j->tri, tri);
    return NULL;
}

/* The GPU's walk of the wide hierarchy: a step is a node visit (the nearest
 * hit child next, the others pushed farthest first) or one triangle of the
 * current leaf.  nodes64: the exported wide nodes (root 0); order, tris9,
 * rays6 and the outputs as for spo_bvh2_closest; boxTests counts the
 * present children of visited nodes. */
void spo_wide_closest(const uint8_t *nodes64, const uint32_t *order,
                      const float *tris9, const float *rays6, uint32_t n,
                      uint32_t threads, uint32_t *primOut, float *tOut,
                      uint64_t *boxTests, uint64_t *triTests)
{
    bvh2_job j;
    memset(&j, 0, sizeof j);
    j.nodes = (const float *)nodes64; j.order = order; j.tris = tris9;
    j.rays = rays6; j.n = n; j.prim = primOut; j.t = tOut;
    atomic_init(&j.next, 0u);
    atomic_init(&j.box, 0ull);
    atomic_init(&j.tri, 0ull);
    if (threads < 1) threads = 1;
    if (threads > 64) threads = 64;
    pthread_t th[64];
    for (uint32_t i = 0; i < threads; ++i)
        pthread_create(&th[i], NULL, wide_worker, &j);
    for (uint32_t i = 0; i < threads; ++i) pthread_join(th[i], NULL);
    if (boxTests) *boxTests = atomic_load(&j.box);
    if (triTests) *triTests = atomic_load(&j.tri);
}

/* Checks of an exported wide hierarchy: every child reference in range and
 * after its parent (or a leaf range inside the primitive order), present
 * children before absent ones, every primitive in exactly one leaf, and
 * every triangle's vertices inside the decoded box of every node child
 * above it.  maxPending: the largest number of references a walk can hold
 * on its stack (the sum over a root-to-node path of the visited nodes'
 * present children less one).  Returns 0 if all hold, else a code naming
 * the first violation. */
typedef struct {
    const uint8_t *nodes;
    uint32_t nNodes;
    const uint32_t *order;
    const float *tris;
    uint32_t nTris;
    uint8_t *seen;
    float lo[64][3], hi[64][3];  /* decoded boxes on the path */
    int rc;
} wide_chk;

static uint32_t wide_check_node(wide_chk *w, uint32_t node, int depth)
{
    const uint8_t *N = w->nodes + 64 * (size_t)node;
    int32_t ref[4];
    memcpy(ref, N + 16, 16);
    int present = 0;
    uint32_t worst = 0;
    for (int c = 0; c < 4 && !w->rc; ++c)
    {
        if (ref[c] == EMPTY_CHILD) continue;
        if (present != c) { w->rc = 7; break; }
        ++present;
        if (depth >= 64) { w->rc = 8; break; }
        wide_box(N, c, w->lo[depth], w->hi[depth]);
        if (ref[c] >= 0)
        {
            if ((uint32_t)ref[c] >= w->nNodes || (uint32_t)ref[c] <= node)
            {
                w->rc = 1;
                break;
            }
            uint32_t sub = wide_check_node(w, (uint32_t)ref[c], depth + 1);
            if (sub > worst) worst = sub;
        }
        else
        {
            uint32_t v = (uint32_t)~ref[c], first = v >> 3, count = (v & 7u) + 1u;
            if (first + count > w->nTris) { w->rc = 3; break; }
            for (uint32_t k = 0; k < count && !w->rc; ++k)
            {
                uint32_t p = w->order[first + k];
                if (p >= w->nTris || w->seen[p]) { w->rc = 4; break; }
                w->seen[p] = 1;
                const float *tr = w->tris + 9 * (size_t)p;
                for (int lv = 0; lv <= depth && !w->rc; ++lv)
                    for (int vtx = 0; vtx < 3; ++vtx)
                        for (int a = 0; a < 3; ++a)
                            if (!(tr[3 * vtx + a] >= w->lo[lv][a] &&
                                  tr[3 * vtx + a] <= w->hi[lv][a]))
                                w->rc = 5;
            }
        }
    }
    if (present == 0) w->rc = w->rc ? w->rc : 9;
    return (uint32_t)(present - 1) + worst;
}

int spo_wide_check(const uint8_t *nodes64, uint32_t nNodes, const uint32_t *order,
                   const float *tris9, uint32_t nTris, uint32_t *maxPending)
{
    wide_chk *w = (wide_chk *)calloc(1, sizeof(wide_chk));
    w->nodes = nodes64; w->nNodes = nNodes; w->order = order;
    w->tris = tris9; w->nTris = nTris;
    w->seen = (uint8_t *)calloc(nTris ? nTris : 1, 1);
    uint32_t pend = nNodes ? wide_check_node(w, 0, 0) : (w->rc = 1, 0);
    for (uint32_t p = 0; p < nTris && !w->rc; ++p)
        if (!w->seen[p]) w->rc = 6;
    int rc = w->rc;
    if (maxPending) *maxPending = pend;
    free(w->seen);
    free(w);
    return rc;
}
