// The reference's parked Vulkan compute tracer (src/shaders/
// test_compute.comp.glsl) on the GPU, as a parity harness: one lane per
// tile index, as the shader runs one invocation per tile, every entity's
// mesh tested triangle by triangle, the shader's hard-coded materials.  The
// FP32 operation order where GLSL leaves it open is the oracle's
// (oracle/spo_compute.c): constructor arguments drawn left to right,
// normalize(v) = v / sqrt(dot(v, v)), mat4 * vec4 summed x, y, z, w; an
// entity whose mesh has no hit is skipped (the shader would read an
// uninitialized normal there).
#include <float.h>

#include "spg_internal.h"

namespace spg {
namespace {

constexpr uint32_t kCtW = 256, kCtH = 192, kCtTile = 16;
constexpr int kCtBounces = 3;

struct CtRng
{
    uint32_t s;
    __device__ uint32_t next()
    {
        uint32_t x = s;
        x ^= x << 13;
        x ^= x >> 17;
        x ^= x << 5;
        s = x;
        return x;
    }
    __device__ float uni() { return (float)next() / (float)0xFFFFFFFFu; }
    __device__ float bi() { return -1.0f + 2.0f * uni(); }
};

__device__ v3 gl_normalize(v3 v)
{
    float len = sqrtf(dot(v, v));
    return v3{v.x / len, v.y / len, v.z / len};
}

__device__ float checker(uint32_t x, uint32_t y)
{
    return ((x / 16u) + (y / 16u)) % 2u ? 0.04f : 0.18f;
}

// CreateCheckerBoardImage (src/image.cpp:134-160) sampled bilinearly with
// repeat addressing, texel centres at (i + 0.5) / 256
__device__ float checker_bilinear(float u, float v)
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

struct CtTri
{
    float t;
    v2 uv;
    v3 n;
};

__device__ CtTri ct_mt(v3 o, v3 d, v3 a, v3 b, v3 c)
{
    CtTri r{-1.0f, {0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
    v3 T = o - a, e1 = b - a, e2 = c - a;
    v3 p = cross(d, e2), q = cross(T, e1);
    float mx = dot(q, e2), my = dot(p, T), mz = dot(q, d);
    float det = 1.0f / dot(p, e1);
    float t = det * mx, u = det * my, v = det * mz;
    float w = 1.0f - u - v;
    if (u >= 0.0f && u <= 1.0f && v >= 0.0f && v <= 1.0f && w >= 0.0f && w <= 1.0f)
    {
        r.t = t;
        r.n = gl_normalize(cross(e1, e2));
        r.uv = v2{u, v};
    }
    return r;
}

__device__ v3 vtx_p(const DevMesh &m, uint32_t i)
{
    const float *p = m.verts + 8 * (size_t)i;
    return v3{p[0], p[1], p[2]};
}

__device__ CtTri ct_mesh(const DevMesh &m, v3 o, v3 d)
{
    CtTri best{FLT_MAX, {0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
    for (uint32_t k = 0; k < m.triCount; ++k)
    {
        const uint32_t i0 = m.idx[3 * k], i1 = m.idx[3 * k + 1], i2 = m.idx[3 * k + 2];
        CtTri h = ct_mt(o, d, vtx_p(m, i0), vtx_p(m, i1), vtx_p(m, i2));
        if (h.t >= 0.0f && h.t < FLT_MAX && h.t < best.t)
        {
            const float *t0 = m.verts + 8 * (size_t)i0 + 6;
            const float *t1 = m.verts + 8 * (size_t)i1 + 6;
            const float *t2 = m.verts + 8 * (size_t)i2 + 6;
            float w = 1.0f - h.uv.x - h.uv.y;
            best = h;
            best.uv = v2{t0[0] * w + t1[0] * h.uv.x + t2[0] * h.uv.y,
                         t0[1] * w + t1[1] * h.uv.x + t2[1] * h.uv.y};
        }
    }
    return best;
}

struct CtHit
{
    float t;
    uint32_t material;
    v3 n;
    v2 uv;
};

__device__ CtHit ct_scene(const DevScene &S, v3 o, v3 d)
{
    CtHit r{FLT_MAX, 6u, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f}};
    for (uint32_t i = 0; i < S.instCount; ++i)
    {
        const DevInstance &e = S.inst[i];
        v3 lo = xform_point(e.inv, o);
        v3 ld = gl_normalize(xform_vector(e.inv, d));
        CtTri h = ct_mesh(S.meshes[e.mesh], lo, ld);
        if (!(h.t < FLT_MAX)) continue;  // no triangle hit
        v3 wh = xform_point(e.model, lo + ld * h.t);
        float t = dot(wh - o, d);
        v3 wn = gl_normalize(xform_vector(e.model, h.n));
        if (t >= 0.0f && t < r.t)
        {
            r.t = t;
            r.material = e.materialId;
            r.n = wn;
            r.uv = h.uv;
        }
    }
    return r;
}

struct CtCamera
{
    float c[16];
};

__global__ __launch_bounds__(64) void k_compute_tracer(DevScene S, CtCamera cam,
                                                       const uint32_t *tiles,
                                                       uint32_t count, float4 *image)
{
    const uint32_t j = blockIdx.x * 64 + threadIdx.x;
    if (j >= count) return;
    const uint32_t tileIndex = tiles[j];
    const v3 right{cam.c[0], cam.c[1], cam.c[2]}, up{cam.c[4], cam.c[5], cam.c[6]};
    const v3 fwd{cam.c[8], cam.c[9], cam.c[10]}, pos{cam.c[12], cam.c[13], cam.c[14]};
    const float W = (float)kCtW, H = (float)kCtH;
    const float filmW = 1.0f, filmH = H / W;
    const float halfW = filmW * 0.5f, halfH = filmH * 0.5f;
    const float hpw = 0.5f / W, hph = 0.5f / H;
    const v3 center = pos + fwd * 0.8f;
    const float invW = 1.0f / W, invH = 1.0f / H;
    const uint32_t tx = tileIndex % (kCtW / kCtTile), ty = tileIndex / (kCtW / kCtTile);
    const uint32_t x0 = tx * kCtTile, y0 = ty * kCtTile;
    const uint32_t x1 = min(kCtW, (tx + 1) * kCtTile), y1 = min(kCtH, (ty + 1) * kCtTile);
    CtRng rng{(tileIndex + 1u) * 0xF51C0E49u};
    for (uint32_t y = y0; y < y1; ++y)
        for (uint32_t x = x0; x < x1; ++x)
        {
            float px = (float)x + 0.5f, py = (float)y + 0.5f;
            float jx = hpw * rng.bi();
            float jy = hph * rng.bi();
            px = px + jx;
            py = py + jy;
            float p0 = px * invW, p1 = py * invH;
            p1 = 1.0f - p1;
            p0 = p0 * 2.0f - 1.0f;
            p1 = p1 * 2.0f - 1.0f;
            v3 filmP = center;
            filmP = filmP + right * (halfW * p0);
            filmP = filmP + up * (halfH * p1);
            v3 o = pos, d = gl_normalize(filmP - o);
            uint32_t mats[kCtBounces];
            v3 nrm[kCtBounces], inc[kCtBounces];
            v2 uvs[kCtBounces];
            int len = 0;
            for (int b = 0; b < kCtBounces; ++b)
            {
                CtHit h = ct_scene(S, o, d);
                if (h.t > 0.0f && h.t < FLT_MAX)
                {
                    float ox = rng.bi(), oy = rng.bi(), oz = rng.bi();
                    v3 dir = gl_normalize(h.n + v3{ox, oy, oz});
                    if (dot(dir, h.n) < 0.0f) dir = -dir;
                    mats[len] = h.material;
                    nrm[len] = h.n;
                    inc[len] = dir;
                    uvs[len] = h.uv;
                    ++len;
                    v3 wp = o + d * h.t;
                    o = wp + h.n * 0.001f;
                    d = dir;
                }
                else
                {
                    mats[len] = 6u;
                    nrm[len] = v3{0.0f, 0.0f, 0.0f};
                    inc[len] = v3{0.0f, 0.0f, 0.0f};
                    uvs[len] = v2{0.0f, 0.0f};
                    ++len;
                    break;
                }
            }
            v3 L{0.0f, 0.0f, 0.0f};
            for (int i = len - 1; i >= 0; --i)
            {
                v3 e{0.0f, 0.0f, 0.0f}, a{0.0f, 0.0f, 0.0f};
                if (mats[i] == 0u) a = v3{0.18f, 0.1f, 0.1f};
                else if (mats[i] == 2u)
                {
                    float c = checker_bilinear(uvs[i].x, uvs[i].y);
                    a = v3{c, c, c};
                }
                else if (mats[i] == 5u) e = v3{1.0f, 1.0f, 1.0f};
                float cd = dot(nrm[i], inc[i]);
                float cosine = 0.0f > cd ? 0.0f : cd;
                L = e + had(a, L) * cosine;
            }
            image[(size_t)y * kCtW + x] = make_float4(L.x, L.y, L.z, 1.0f);
        }
}

}  // namespace
}  // namespace spg

using namespace spg;

extern "C" int spg_compute_tracer_tiles(spg_Scene *s, const float cameraTransform[16],
                                        const uint32_t *tileIndices, uint32_t count,
                                        float *image)
{
    if (!s || !cameraTransform || (count && (!tileIndices || !image)))
        return fail(SPG_E_INVALID, "spg_compute_tracer_tiles: NULL argument");
    const uint32_t tilesTotal = (kCtW / kCtTile) * (kCtH / kCtTile);
    for (uint32_t k = 0; k < count; ++k)
        if (tileIndices[k] >= tilesTotal)
            return fail(SPG_E_INVALID, "spg_compute_tracer_tiles: tile index out of range");
    if (count == 0) return SPG_OK;
    spg_Device *d = s->dev;
    std::lock_guard<std::mutex> g(d->lock);
    HIPCHK(hipSetDevice(d->ordinal));
    CHK(scene_sync(s));
    void *buf = nullptr;
    const size_t imgBytes = (size_t)kCtW * kCtH * 16;
    HIPCHK(hipMalloc(&buf, imgBytes + 4 * (size_t)count));
    float4 *dImg = (float4 *)buf;
    uint32_t *dTiles = (uint32_t *)((char *)buf + imgBytes);
    CtCamera cam;
    memcpy(cam.c, cameraTransform, sizeof cam.c);
    int r = hip_check(hipMemcpyAsync(dImg, image, imgBytes, hipMemcpyHostToDevice, d->stream),
                      "hipMemcpy");
    if (!r)
        r = hip_check(hipMemcpyAsync(dTiles, tileIndices, 4 * (size_t)count,
                                     hipMemcpyHostToDevice, d->stream), "hipMemcpy");
    if (!r)
    {
        hipLaunchKernelGGL(k_compute_tracer, dim3((count + 63) / 64), dim3(64), 0, d->stream,
                           s->dev_scene, cam, dTiles, count, dImg);
        r = hip_check(hipGetLastError(), "k_compute_tracer");
    }
    if (!r)
        r = hip_check(hipMemcpyAsync(image, dImg, imgBytes, hipMemcpyDeviceToHost, d->stream),
                      "hipMemcpy");
    if (!r) r = hip_check(hipStreamSynchronize(d->stream), "k_compute_tracer");
    hipFree(buf);
    return r;
}
