/* ORACLE -- test infrastructure only (see spo.h).
 *
 * The reference's parked Vulkan compute tracer (src/shaders/
 * test_compute.comp.glsl: main :616-671, PathTraceTile :523-614,
 * RayIntersectScene :367-445, RayIntersectTriangleMesh :278-323,
 * RayIntersectTriangleMT :168-201, ComputeRadianceForPath :456-489) restated
 * in C with the FP32 operation order written out where GLSL leaves it open:
 *   - a 256 x 192 image in 16 x 16 tiles, one invocation per tile index,
 *     tile (index % 16, index / 16); xorshift32 seeded (index + 1) *
 *     0xF51C0E49; RandomUnilateral = float(x) / float(2^32 - 1) (= 2^32 in
 *     float), RandomBilateral = 2 U - 1; constructor arguments drawn left to
 *     right;
 *   - film half-extents 0.5 x (1, 192/256), half-pixel 0.5/W, 0.5/H, film
 *     centre position + forward * 0.8, p = ((x, y) + 0.5 + jitter) / (W, H),
 *     y flipped, mapped to [-1, 1];
 *   - normalize(v) = v / sqrt(dot(v, v)); mat4 * vec4 summed column by
 *     column (x, y, z, w); every entity's mesh tested triangle by triangle
 *     (Moller-Trumbore without back-face culling, the first of equal t
 *     kept), the flat normal normalize(cross(e1, e2)) and barycentric uv;
 *     an entity whose mesh has no hit is skipped (the shader goes on with
 *     t = FLT_MAX and an uninitialized normal, and its world t can land
 *     just below FLT_MAX and count as a hit: undefined there, excluded
 *     here);
 *   - three bounces, direction normalize(n + (B, B, B)) flipped into the
 *     normal's hemisphere, origin offset 0.001 along n;
 *   - materials by id: 0 red (albedo 0.18, 0.1, 0.1), 2 checkerboard (the
 *     bilinear, repeating lookup at level 0 of CreateCheckerBoardImage,
 *     src/image.cpp:134-160), 5 white light (emission 1), any other id
 *     black; radiance_i = emission + (albedo * L_i+1) * max(0, n.wi).
 * Pinned by closed forms (tests/test_compute_tracer.py): a light seen
 * directly is exactly 1 and an empty view 0, a red surface lit by the
 * white light keeps green equal to blue and below the albedo bound, the
 * checkerboard's texels and bilinear midpoints, each tile writing its own
 * 16 x 16 pixels. */
#include <float.h>
#include <math.h>
#include <string.h>

#include "spo_math.h"

#define CT_W 256u
#define CT_H 192u
#define CT_TILE 16u
#define CT_BOUNCES 3

typedef struct { uint32_t s; } ct_rng;

static uint32_t ct_next(ct_rng *r)
{
    uint32_t x = r->s;
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    r->s = x;
    return x;
}
static float ct_uni(ct_rng *r) { return (float)ct_next(r) / (float)0xFFFFFFFFu; }
static float ct_bi(ct_rng *r) { return -1.0f + 2.0f * ct_uni(r); }

static spo_v3 gl_normalize(spo_v3 v)
{
    float len = sqrtf(v3dot(v, v));
    return v3(v.x / len, v.y / len, v.z / len);
}

/* CreateCheckerBoardImage, src/image.cpp:134-160: 256 x 256, 16-texel
 * squares of 0.18 and 0.04 */
static float checker(uint32_t x, uint32_t y)
{
    return ((x / 16u) + (y / 16u)) % 2u ? 0.04f : 0.18f;
}

/* bilinear, repeat addressing, texel centres at (i + 0.5) / 256 */
float spo_checker_bilinear(float u, float v)
{
    float fu = u * 256.0f - 0.5f, fv = v * 256.0f - 0.5f;
    float iu = floorf(fu), iv = floorf(fv);
    float au = fu - iu, av = fv - iv;
    int32_t x0 = (int32_t)iu, y0 = (int32_t)iv;
    uint32_t x0w = (uint32_t)(((x0 % 256) + 256) % 256), y0w = (uint32_t)(((y0 % 256) + 256) % 256);
    uint32_t x1w = (x0w + 1u) % 256u, y1w = (y0w + 1u) % 256u;
    float top = checker(x0w, y0w) * (1.0f - au) + checker(x1w, y0w) * au;
    float bot = checker(x0w, y1w) * (1.0f - au) + checker(x1w, y1w) * au;
    return top * (1.0f - av) + bot * av;
}

typedef struct { float t; spo_v2 uv; spo_v3 normal; } ct_tri;

static ct_tri ct_mt(spo_v3 o, spo_v3 d, spo_v3 a, spo_v3 b, spo_v3 c)
{
    ct_tri r;
    r.t = -1.0f;
    r.uv.x = r.uv.y = 0.0f;
    r.normal = v3(0, 0, 0);
    spo_v3 T = v3sub(o, a), e1 = v3sub(b, a), e2 = v3sub(c, a);
    spo_v3 p = v3cross(d, e2), q = v3cross(T, e1);
    float mx = v3dot(q, e2), my = v3dot(p, T), mz = v3dot(q, d);
    float det = 1.0f / v3dot(p, e1);
    float t = det * mx, u = det * my, v = det * mz;
    float w = 1.0f - u - v;
    if (u >= 0.0f && u <= 1.0f && v >= 0.0f && v <= 1.0f && w >= 0.0f && w <= 1.0f)
    {
        r.t = t;
        r.normal = gl_normalize(v3cross(e1, e2));
        r.uv.x = u;
        r.uv.y = v;
    }
    return r;
}

static ct_tri ct_mesh(const spo_Mesh *m, spo_v3 o, spo_v3 d)
{
    ct_tri best;
    best.t = FLT_MAX;
    best.uv.x = best.uv.y = 0.0f;
    best.normal = v3(0, 0, 0);
    for (uint32_t k = 0; k + 2 < m->indexCount; k += 3)
    {
        const spo_Vertex *v0 = &m->vertices[m->indices[k]];
        const spo_Vertex *v1 = &m->vertices[m->indices[k + 1]];
        const spo_Vertex *v2 = &m->vertices[m->indices[k + 2]];
        ct_tri h = ct_mt(o, d, v0->p, v1->p, v2->p);
        if (h.t >= 0.0f && h.t < FLT_MAX && h.t < best.t)
        {
            float w = 1.0f - h.uv.x - h.uv.y;
            best = h;
            best.uv.x = v0->uv.x * w + v1->uv.x * h.uv.x + v2->uv.x * h.uv.y;
            best.uv.y = v0->uv.y * w + v1->uv.y * h.uv.x + v2->uv.y * h.uv.y;
        }
    }
    return best;
}

typedef struct { float t; uint32_t material; spo_v3 normal; spo_v2 uv; } ct_hit;

static ct_hit ct_scene(const spo_Scene *s, spo_v3 o, spo_v3 d)
{
    ct_hit r;
    r.t = FLT_MAX;
    r.material = 6; /* Material_Black */
    r.normal = v3(0, 0, 0);
    r.uv.x = r.uv.y = 0.0f;
    for (uint32_t i = 0; i < s->objectCount; ++i)
    {
        const spo_Object *e = &s->objects[i];
        spo_v3 lo = m4point(e->invModel, o);
        spo_v3 ld = gl_normalize(m4vector(e->invModel, d));
        ct_tri h = ct_mesh(&s->meshes[e->mesh], lo, ld);
        if (!(h.t < FLT_MAX)) continue;  /* no triangle hit */
        spo_v3 wh = m4point(e->model, v3add(lo, v3mul(ld, h.t)));
        float t = v3dot(v3sub(wh, o), d);
        spo_v3 wn = gl_normalize(m4vector(e->model, h.normal));
        if (t >= 0.0f && t < r.t)
        {
            r.t = t;
            r.material = e->material;
            r.normal = wn;
            r.uv = h.uv;
        }
    }
    return r;
}

/* One tile of the compute tracer into the 256 x 192 RGBA32F image.
 * cam16: the uniform's camera transform, columns right, up, forward,
 * position (column-major mat4). */
void spo_compute_tracer_tile(const spo_Scene *s, const float *cam16, uint32_t tileIndex,
                             float *image)
{
    const spo_v3 right = v3(cam16[0], cam16[1], cam16[2]);
    const spo_v3 up = v3(cam16[4], cam16[5], cam16[6]);
    const spo_v3 fwd = v3(cam16[8], cam16[9], cam16[10]);
    const spo_v3 pos = v3(cam16[12], cam16[13], cam16[14]);
    const float W = (float)CT_W, H = (float)CT_H;
    const float filmW = 1.0f, filmH = H / W;
    const float halfW = filmW * 0.5f, halfH = filmH * 0.5f;
    const float hpw = 0.5f / W, hph = 0.5f / H;
    const spo_v3 center = v3add(pos, v3mul(fwd, 0.8f));
    const float invW = 1.0f / W, invH = 1.0f / H;
    const uint32_t tx = tileIndex % (CT_W / CT_TILE), ty = tileIndex / (CT_W / CT_TILE);
    const uint32_t x0 = tx * CT_TILE, y0 = ty * CT_TILE;
    const uint32_t x1 = spo_minu(CT_W, (tx + 1) * CT_TILE), y1 = spo_minu(CT_H, (ty + 1) * CT_TILE);
    ct_rng rng = {(tileIndex + 1u) * 0xF51C0E49u};
    for (uint32_t y = y0; y < y1; ++y)
        for (uint32_t x = x0; x < x1; ++x)
        {
            float px = (float)x + 0.5f, py = (float)y + 0.5f;
            float jx = hpw * ct_bi(&rng);
            float jy = hph * ct_bi(&rng);
            px = px + jx;
            py = py + jy;
            float p0 = px * invW, p1 = py * invH;
            p1 = 1.0f - p1;
            p0 = p0 * 2.0f - 1.0f;
            p1 = p1 * 2.0f - 1.0f;
            spo_v3 filmP = center;
            filmP = v3add(filmP, v3mul(right, halfW * p0));
            filmP = v3add(filmP, v3mul(up, halfH * p1));
            spo_v3 o = pos, d = gl_normalize(v3sub(filmP, o));
            uint32_t mats[CT_BOUNCES];
            spo_v3 nrm[CT_BOUNCES], inc[CT_BOUNCES];
            spo_v2 uvs[CT_BOUNCES];
            int len = 0;
            for (int b = 0; b < CT_BOUNCES; ++b)
            {
                ct_hit h = ct_scene(s, o, d);
                if (h.t > 0.0f && h.t < FLT_MAX)
                {
                    float ox = ct_bi(&rng), oy = ct_bi(&rng), oz = ct_bi(&rng);
                    spo_v3 dir = gl_normalize(v3add(h.normal, v3(ox, oy, oz)));
                    if (v3dot(dir, h.normal) < 0.0f) dir = v3neg(dir);
                    mats[len] = h.material;
                    nrm[len] = h.normal;
                    inc[len] = dir;
                    uvs[len] = h.uv;
                    ++len;
                    spo_v3 wp = v3add(o, v3mul(d, h.t));
                    o = v3add(wp, v3mul(h.normal, 0.001f));
                    d = dir;
                }
                else
                {
                    mats[len] = 6;
                    nrm[len] = v3(0, 0, 0);
                    inc[len] = v3(0, 0, 0);
                    uvs[len].x = uvs[len].y = 0.0f;
                    ++len;
                    break;
                }
            }
            spo_v3 L = v3(0, 0, 0);
            for (int i = len - 1; i >= 0; --i)
            {
                spo_v3 e = v3(0, 0, 0), a = v3(0, 0, 0);
                if (mats[i] == 0) a = v3(0.18f, 0.1f, 0.1f);
                else if (mats[i] == 2)
                {
                    float c = spo_checker_bilinear(uvs[i].x, uvs[i].y);
                    a = v3(c, c, c);
                }
                else if (mats[i] == 5) e = v3(1.0f, 1.0f, 1.0f);
                float cosine = spo_max(0.0f, v3dot(nrm[i], inc[i]));
                L = v3add(e, v3mul(v3had(a, L), cosine));
            }
            float *out = image + 4 * ((size_t)y * CT_W + x);
            out[0] = L.x;
            out[1] = L.y;
            out[2] = L.z;
            out[3] = 1.0f;
        }
}
