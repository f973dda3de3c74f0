/* ORACLE -- test infrastructure only (see spo.h).
 * The reference's greedy 4-ary BVH: bvh_CreateTree (src/bvh.cpp:51-200) and
 * the ping-pong two-stack query bvh_IntersectRay (src/bvh.cpp:203-311).
 * Nodes live in an index-addressed array instead of a pointer pool. */
#include <stdlib.h>
#include <string.h>
#include "spo_math.h"

#define SPO_BVH_STACK_SIZE 256 /* src/bvh.h BVH_STACK_SIZE */

typedef struct { float distSq; int32_t node; } spo_DistPair;

/* CompareNodeDistSqPair, src/bvh.cpp:11-33: descending distance. */
static int cmp_pair(const spo_DistPair *a, const spo_DistPair *b)
{
    if (a->distSq < b->distSq) return 1;
    if (a->distSq > b->distSq) return -1;
    return 0;
}

/* The reference sorts with libc qsort; glibc (2.35 here, and the versions the
 * reference was built with) implements it as a top-down merge sort that
 * splits n into n/2 and n-n/2 and takes from the left run when cmp <= 0.
 * That exact merge order is restated so equal keys (and NaN keys) land where
 * the reference puts them. */
static void msort_pairs(spo_DistPair *b, size_t n, spo_DistPair *tmp)
{
    if (n <= 1) return;
    size_t n1 = n / 2, n2 = n - n1;
    spo_DistPair *b1 = b, *b2 = b + n1;
    msort_pairs(b1, n1, tmp);
    msort_pairs(b2, n2, tmp);
    spo_DistPair *out = tmp;
    while (n1 > 0 && n2 > 0)
    {
        if (cmp_pair(b1, b2) <= 0) { *out++ = *b1++; --n1; }
        else { *out++ = *b2++; --n2; }
    }
    if (n1 > 0) memcpy(out, b1, n1 * sizeof *b1);
    memcpy(b, tmp, (n - n2) * sizeof *b);
}

static int32_t alloc_node(spo_Bvh *t)
{
    if (t->count >= t->capacity) return -1;
    int32_t i = (int32_t)t->count++;
    memset(&t->nodes[i], 0, sizeof t->nodes[i]);
    for (int k = 0; k < 4; ++k) t->nodes[i].children[k] = -1;
    return i;
}

spo_Bvh *spo_bvh_create(const float *mins, const float *maxs, uint32_t count)
{
    spo_Bvh *t = (spo_Bvh *)calloc(1, sizeof *t);
    t->root = -1;
    if (count == 0) return t;
    t->capacity = count * 10; /* src/bvh.cpp:57 */
    t->nodes = (spo_BvhNode *)malloc(sizeof(spo_BvhNode) * t->capacity);
    spo_DistPair *unmerged[2];
    unmerged[0] = (spo_DistPair *)calloc(count, sizeof(spo_DistPair));
    unmerged[1] = (spo_DistPair *)calloc(count, sizeof(spo_DistPair));
    spo_DistPair *tmp = (spo_DistPair *)malloc(count * sizeof(spo_DistPair));
    uint32_t n[2] = {count, 0};
    uint32_t rd = 0, wr = 1;
    for (uint32_t leaf = 0; leaf < count; ++leaf)
    {
        int32_t node = alloc_node(t);
        t->nodes[node].min = v3p(mins + 3 * leaf);
        t->nodes[node].max = v3p(maxs + 3 * leaf);
        t->nodes[node].leafIndex = leaf;
        unmerged[rd][leaf].node = node;
    }
    int32_t lastAllocated = unmerged[rd][0].node;
    for (;;)
    {
        n[wr] = 0;
        while (n[rd] > 0)
        {
            uint32_t index = n[rd] - 1;
            int32_t node = unmerged[rd][index].node;
            spo_BvhNode *nd = &t->nodes[node];
            spo_v3 centroid = v3mul(v3add(nd->max, nd->min), 0.5f);
            /* bvh_SortNodesByDistanceSq, src/bvh.cpp:36-49 */
            for (uint32_t i = 0; i < n[rd] - 1; ++i)
            {
                spo_BvhNode *o = &t->nodes[unmerged[rd][i].node];
                spo_v3 c = v3mul(v3add(o->max, o->min), 0.5f);
                spo_v3 d = v3sub(c, centroid);
                unmerged[rd][i].distSq = v3dot(d, d);
            }
            msort_pairs(unmerged[rd], n[rd] - 1, tmp);
            uint32_t nc = spo_minu(3, n[rd] - 1);
            int32_t neighbors[3];
            for (uint32_t i = 0; i < nc; ++i)
                neighbors[i] = unmerged[rd][index - (i + 1)].node;
            if (nc > 0)
            {
                int32_t nn = alloc_node(t);
                spo_BvhNode *newNode = &t->nodes[nn];
                newNode->leafIndex = 0xFFFFFFFFu;
                newNode->min = t->nodes[node].min;
                newNode->max = t->nodes[node].max;
                newNode->children[0] = node;
                for (uint32_t i = 0; i < nc; ++i)
                {
                    newNode->min = v3min(newNode->min, t->nodes[neighbors[i]].min);
                    newNode->max = v3max(newNode->max, t->nodes[neighbors[i]].max);
                    newNode->children[i + 1] = neighbors[i];
                }
                lastAllocated = nn;
                unmerged[wr][n[wr]++].node = nn;
                n[rd] -= nc + 1;
            }
            else
            {
                unmerged[wr][n[wr]++].node = node;
                n[rd]--;
            }
        }
        if (n[wr] > 1)
        {
            uint32_t tswap = rd; rd = wr; wr = tswap;
        }
        else break;
    }
    free(unmerged[0]);
    free(unmerged[1]);
    free(tmp);
    t->root = lastAllocated;
    return t;
}

void spo_bvh_destroy(spo_Bvh *t)
{
    if (!t) return;
    free(t->nodes);
    free(t);
}

/* bvh_IntersectRay, src/bvh.cpp:203-311 (the USE_BVH_SIMD_RAY_INTERSECT_AABB
 * branch).  leaves receives node leaf indices in visit order. */
uint32_t spo_bvh_intersect(const spo_Bvh *t, const float *o, const float *d,
                           uint32_t maxHits, uint32_t *leaves,
                           uint32_t *aabbTests, uint32_t *errorOccurred)
{
    static __thread int32_t stack[2][SPO_BVH_STACK_SIZE + 1];
    uint32_t sizes[2] = {0, 0};
    uint32_t rd = 0, wr = 1, count = 0, tests = 0, err = 0;
    spo_v3 inv = v3inv(v3p(d));
    float invD[3];
    v3put(invD, inv);
    if (t && t->root >= 0)
    {
        const spo_BvhNode *r = &t->nodes[t->root];
        float mn[12] = {0}, mx[12] = {0};
        v3put(mn, r->min);
        v3put(mx, r->max);
        if (spo_rt_aabb4(mn, mx, o, invD) & 1u)
        {
            stack[0][0] = t->root;
            sizes[0] = 1;
        }
    }
    while (sizes[rd] > 0)
    {
        uint32_t top = --sizes[rd];
        const spo_BvhNode *node = &t->nodes[stack[rd][top]];
        if (node->children[0] >= 0)
        {
            float mn[12] = {0}, mx[12] = {0};
            uint32_t childCount = 0;
            for (int i = 0; i < 4; ++i)
                if (node->children[i] >= 0)
                {
                    v3put(mn + 3 * i, t->nodes[node->children[i]].min);
                    v3put(mx + 3 * i, t->nodes[node->children[i]].max);
                    childCount++;
                }
            tests += childCount;
            uint32_t mask = spo_rt_aabb4(mn, mx, o, invD);
            for (uint32_t i = 0; i < childCount; ++i)
                if (mask & (1u << i))
                {
                    if (sizes[wr] <= SPO_BVH_STACK_SIZE)
                        stack[wr][sizes[wr]] = node->children[i];
                    sizes[wr] += 1;
                }
        }
        else
        {
            if (count < maxHits) leaves[count++] = node->leafIndex;
            else { err = 1; break; }
        }
        if (sizes[rd] == 0) { uint32_t s = rd; rd = wr; wr = s; }
    }
    if (aabbTests) *aabbTests = tests;
    if (errorOccurred) *errorOccurred = err;
    return count;
}

static uint32_t dump_node(const spo_Bvh *t, int32_t ni, float *boxes,
                          int32_t *children, uint32_t *leafIndex,
                          uint32_t *count, uint32_t cap)
{
    uint32_t me = (*count)++;
    if (me >= cap) return me;
    const spo_BvhNode *n = &t->nodes[ni];
    v3put(boxes + 6 * me, n->min);
    v3put(boxes + 6 * me + 3, n->max);
    leafIndex[me] = n->leafIndex;
    for (int i = 0; i < 4; ++i) children[4 * me + i] = -1;
    for (int i = 0; i < 4; ++i)
        if (n->children[i] >= 0)
            children[4 * me + i] = (int32_t)dump_node(
                t, n->children[i], boxes, children, leafIndex, count, cap);
    return me;
}

uint32_t spo_bvh_dump(const spo_Bvh *t, float *boxes, int32_t *children,
                      uint32_t *leafIndex, uint32_t cap)
{
    uint32_t count = 0;
    if (t && t->root >= 0)
        dump_node(t, t->root, boxes, children, leafIndex, &count, cap);
    return count;
}
