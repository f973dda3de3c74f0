// gfx950 kernels of the wavefront path tracer.
//
//   closest hit   k_closest_mesh   one ray per lane, BVH2 with both child
//                                  boxes per 64-byte node, LDS stack, t-max
//                                  culling, strict-FP slab and MT tests
//   scene query   k_intersect_scene  TLAS over instances -> per-instance
//                                  local ray -> BLAS (sp_RayIntersectScene)
//   wavefront     k_raygen -> (k_extend -> k_shade)* -> k_finalize
//                 rays live in SoA queues; k_shade appends continuing rays
//                 to the next queue with a wave ballot + one atomic per wave
//
// All arithmetic that the reference defines is written in its operation
// order (spg_math.h) and built with -ffp-contract=off and correctly rounded
// division/sqrt, so a path computes the same FP32 bits as the reference CPU
// tracer given the same random numbers and transcendental results.
#include "spg_mlp.h"
#include "spg_traverse.h"
#include "spg_tx.h"

namespace spg {

// ------------------------------------------------------------- helpers
__device__ __forceinline__ uint32_t lane_id()
{
    return __lane_id();
}

// Sum over the active lanes of the wave, result valid in every lane.
__device__ __forceinline__ unsigned long long wave_sum(unsigned long long x)
{
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    return x;
}

__device__ __forceinline__ void wave_add(unsigned long long *dst,
                                         unsigned long long x)
{
    // all 64 lanes must be active here
    unsigned long long s = wave_sum(x);
    if (lane_id() == 0 && s) atomicAdd(dst, s);
}

// Queue positions for a 256-thread block: every wave's count of entries
// for queue A (and B) is added with one atomic per block and queue, so a
// launch makes one atomic per block-iteration instead of one per
// wave-iteration (same-address atomics serialize in L2).  Returns the
// wave's first position in each queue.  Called by every thread of the
// block (block-uniform control flow); B may be null.
__device__ __forceinline__ void block_queue_bases(uint32_t *qa, uint32_t na,
                                                  uint32_t *qb, uint32_t nb,
                                                  uint32_t &baseA, uint32_t &baseB)
{
    __shared__ uint32_t cnt[2][4], bases[2][4];
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0)
    {
        cnt[0][w] = na;
        cnt[1][w] = nb;
    }
    __syncthreads();
    if (threadIdx.x < 2)
    {
        const int k = threadIdx.x;
        uint32_t *q = k ? qb : qa;
        const uint32_t c0 = cnt[k][0], c1 = cnt[k][1], c2 = cnt[k][2], c3 = cnt[k][3];
        const uint32_t all = c0 + c1 + c2 + c3;
        const uint32_t b = (q && all) ? atomicAdd(q, all) : 0u;
        bases[k][0] = b;
        bases[k][1] = b + c0;
        bases[k][2] = b + c0 + c1;
        bases[k][3] = b + c0 + c1 + c2;
    }
    __syncthreads();
    baseA = bases[0][w];
    baseB = bases[1][w];
}

// dst[0] += a, dst[1] += b summed over a 256-thread block, one atomic each.
__device__ __forceinline__ void block_add2(unsigned long long *dst, unsigned long long a,
                                           unsigned long long b)
{
    __shared__ unsigned long long part[2][4];
    const unsigned long long sa = wave_sum(a), sb = wave_sum(b);
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0)
    {
        part[0][w] = sa;
        part[1][w] = sb;
    }
    __syncthreads();
    if (threadIdx.x < 2)
    {
        const unsigned long long t = part[threadIdx.x][0] + part[threadIdx.x][1] +
                                     part[threadIdx.x][2] + part[threadIdx.x][3];
        if (t) atomicAdd(&dst[threadIdx.x], t);
    }
}

// Surface at a hit (src/sp_scene.cpp:196-213 and :302-312): uv from the
// barycentrics, smooth normal normalize(w n0 + u n1 + v n2) or the MT normal
// cross(e1, e2), then world normal normalize(M n).
__device__ __forceinline__ void surface_at(const DevScene &S, uint32_t inst,
                                           uint32_t prim, float u, float v,
                                           v3 &worldN, v2 &uv)
{
    const DevInstance &I = S.inst[inst];
    const DevMesh &M = S.meshes[I.mesh];
    // the triangle's shading record: vertex normals (or the flat normal)
    // and uvs in one 64-byte fetch
    const float4 *r = M.shade + 4 * (size_t)prim;
    const float4 a = r[0], b = r[1], c = r[2], d = r[3];
    float w = 1.0f - u - v;
    uv.x = (c.y * w + c.w * u) + d.y * v;
    uv.y = (c.z * w + d.x * u) + d.z * v;
    v3 ln;
    if (M.smooth)
    {
        v3 n0{a.x, a.y, a.z}, n1{a.w, b.x, b.y}, n2{b.z, b.w, c.x};
        ln = normalize((n0 * w + n1 * u) + n2 * v);
    }
    else
        ln = v3{a.x, a.y, a.z};
    worldN = normalize(xform_vector(I.model, ln));
}

// ------------------------------------------------------------- materials
__device__ __forceinline__ float4 texel(const DevScene &S, const DevTexture &T,
                                        uint64_t i)
{
    return S.texels[T.offset + i];
}

// SampleImageNearest, src/image.h:3-17, with (u32)Floor taking NaN and
// negatives to 0 and the flat index clamped to the last texel.
__device__ __forceinline__ float4 sample_nearest(const DevScene &S,
                                                 const DevTexture &T, float u,
                                                 float v)
{
    float fx = floorf(u * (float)T.width);
    float fy = floorf(v * (float)T.height);
    uint32_t x = fx >= 0.0f ? (fx < 4294967296.0f ? (uint32_t)fx : 0xFFFFFFFFu) : 0u;
    uint32_t y = fy >= 0.0f ? (fy < 4294967296.0f ? (uint32_t)fy : 0xFFFFFFFFu) : 0u;
    uint64_t idx = (uint64_t)y * T.width + x;
    uint64_t last = (uint64_t)T.width * T.height - 1;
    if (idx > last) idx = last;
    return texel(S, T, idx);
}

// Face and face-plane coordinates of a direction: the inverse of
// CreateCubeMap's texel-to-direction map (src/cubemap.cpp:237-291) with the
// face bases of src/cubemap.cpp:55-103.
struct CubeCoord
{
    uint32_t face;
    float fx, fy;
};

__device__ __forceinline__ CubeCoord cube_coord(v3 d)
{
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    uint32_t face;
    float ma, sc, tc;
    if (ax >= ay && ax >= az)
    {
        face = d.x >= 0.0f ? 0 : 1;
        ma = ax;
        sc = d.x >= 0.0f ? -d.z : d.z;
        tc = d.y;
    }
    else if (ay >= az)
    {
        face = d.y >= 0.0f ? 2 : 3;
        ma = ay;
        sc = d.x;
        tc = d.y >= 0.0f ? -d.z : d.z;
    }
    else
    {
        face = d.z >= 0.0f ? 4 : 5;
        ma = az;
        sc = d.z >= 0.0f ? d.x : -d.x;
        tc = d.y;
    }
    return CubeCoord{face, sc / ma, tc / ma};
}

// Texel index of the nearest cube-map texel (face-major, rows, columns).
__device__ __forceinline__ uint64_t cube_index(const DevTexture &T, CubeCoord c)
{
    float px = (c.fx + 1.0f) * (0.5f * (float)T.width) + 0.5f;
    float py = (1.0f - c.fy) * (0.5f * (float)T.height) + 0.5f;
    px = px >= 0.0f ? px : 0.0f;
    py = py >= 0.0f ? py : 0.0f;
    px = px < (float)(T.width - 1) ? px : (float)(T.width - 1);
    py = py < (float)(T.height - 1) ? py : (float)(T.height - 1);
    uint32_t ix = (uint32_t)floorf(px), iy = (uint32_t)floorf(py);
    return ((uint64_t)c.face * T.height + iy) * T.width + ix;
}

// Nearest cube-map texel by direction.
__device__ __forceinline__ float4 sample_cube(const DevScene &S,
                                              const DevTexture &T, v3 d)
{
    return texel(S, T, cube_index(T, cube_coord(d)));
}

// Density per steradian with which the environment sampler draws direction
// d: the texel's P_i / area_i times (1 + fx^2 + fy^2)^(3/2); 0 without a
// table.
__device__ __forceinline__ float env_pdf(const DevScene &S, v3 d)
{
    if (S.envTex < 0) return 0.0f;
    const CubeCoord c = cube_coord(d);
    const float t = (1.0f + c.fx * c.fx) + c.fy * c.fy;
    return S.envPdf[cube_index(S.texs[S.envTex], c)] * (t * sqrtf(t));
}

// A direction drawn from the environment table: texel by the alias table
// (index from the 32-bit draw r, alias test with e1), then uniform in the
// texel's face-plane rectangle (e2, e3).
__device__ __forceinline__ v3 env_direction(const DevScene &S, uint32_t r, float e1,
                                            float e2, float e3)
{
    const DevTexture &T = S.texs[S.envTex];
    uint32_t i = (uint32_t)(((uint64_t)r * S.envCount) >> 32);
    if (!(e1 < S.envProb[i])) i = S.envAlias[i];
    const uint32_t w = T.width, h = T.height;
    const uint32_t ix = i % w, iy = (i / w) % h, face = i / (w * h);
    const float x0 = ix == 0 ? -1.0f : (float)(2 * ix - 1) / (float)w - 1.0f;
    const float x1 = ix + 1 == w ? 1.0f : (float)(2 * ix + 1) / (float)w - 1.0f;
    const float y1 = iy == 0 ? 1.0f : 1.0f - (float)(2 * iy - 1) / (float)h;
    const float y0 = iy + 1 == h ? -1.0f : 1.0f - (float)(2 * iy + 1) / (float)h;
    const float fx = x0 + e2 * (x1 - x0), fy = y0 + e3 * (y1 - y0);
    v3 v;
    switch (face)
    {
    case 0: v = v3{1.0f, fy, -fx}; break;
    case 1: v = v3{-1.0f, fy, fx}; break;
    case 2: v = v3{fx, 1.0f, -fy}; break;
    case 3: v = v3{fx, -1.0f, fy}; break;
    case 4: v = v3{fx, fy, 1.0f}; break;
    default: v = v3{-fx, fy, -1.0f}; break;
    }
    return normalize(v);
}

// World-space vertices of triangle `prim` of instance `inst`.
__device__ __forceinline__ void world_triangle(const DevScene &S, uint32_t inst,
                                               uint32_t prim, v3 &A, v3 &B, v3 &C)
{
    const DevInstance &I = S.inst[inst];
    const DevMesh &M = S.meshes[I.mesh];
    const float *p0 = M.verts + 8 * (size_t)M.idx[3 * prim];
    const float *p1 = M.verts + 8 * (size_t)M.idx[3 * prim + 1];
    const float *p2 = M.verts + 8 * (size_t)M.idx[3 * prim + 2];
    A = xform_point(I.model, v3{p0[0], p0[1], p0[2]});
    B = xform_point(I.model, v3{p1[0], p1[1], p1[2]});
    C = xform_point(I.model, v3{p2[0], p2[1], p2[2]});
}

// The spatial emitter choice (S.emitInstCount > 0): emitter instance k seen
// from x with shading normal n weighs power x 1 / max(d^2, r^2) (d the
// distance to its bounding sphere's centre, r the radius; the reciprocal
// correctly rounded), 0 when the sphere lies wholly behind the tangent
// plane; the weights are summed in index order.
__device__ __forceinline__ float emit_weight(const DevScene &S, uint32_t k, v3 x, v3 n)
{
    const float4 a = S.emitInst[2 * k], b = S.emitInst[2 * k + 1];
    const v3 v{a.x - x.x, a.y - x.y, a.z - x.z};
    const float d2 = dot(v, v), r2 = a.w * a.w;
    const float w = b.x * rcp_cr(fmaxf(d2, r2));
    return dot(n, v) > -a.w ? w : 0.0f;
}

// One scan of the spatial choice's weights from x (shading normal n): their
// sum W in index order, the running sum at the end of each of 8 equal
// groups of emitters (the last ones may be short or empty) and the last
// positive weight.  The emitter drawn is the first whose running sum
// exceeds e4 W; emit_direction finds its group from the group sums and
// rescans only that group from the sum before it, which repeats the full
// scan's additions and so its result.
constexpr int kEmitGroups = 8;
struct EmitScan
{
    float W;
    float G[kEmitGroups];
    uint32_t lastPos;
    float lastPosW;
};

__device__ __forceinline__ void emit_scan(const DevScene &S, v3 x, v3 n, EmitScan &e)
{
    const uint32_t cnt = S.emitInstCount, gs = (cnt + kEmitGroups - 1) / kEmitGroups;
    e.W = 0.0f;
    e.lastPos = 0;
    e.lastPosW = 0.0f;
#pragma unroll
    for (int g = 0; g < kEmitGroups; ++g)
    {
        const uint32_t k1 = min(cnt, (uint32_t)(g + 1) * gs);
        for (uint32_t k = (uint32_t)g * gs; k < k1; ++k)
        {
            const float w = emit_weight(S, k, x, n);
            e.W += w;
            e.lastPos = w > 0.0f ? k : e.lastPos;
            e.lastPosW = w > 0.0f ? w : e.lastPosW;
        }
        e.G[g] = e.W;
    }
}

// Density per unit area of the emitter sampler on instance `inst`'s
// triangle `prim`: the spatial choice's P(instance | x, n) / area (W the
// weight sum at x, kept from the light sample drawn there), or the global
// table's entry; 0 if not sampled.
__device__ __forceinline__ float emit_pdf_area(const DevScene &S, uint32_t inst,
                                               uint32_t prim, v3 x, v3 n, float W)
{
    const int32_t b = S.emitBase[inst];
    if (b < 0) return 0.0f;
    if (S.emitInstCount == 0) return S.emitPdfA[b + prim];
    const int32_t k = S.emitSlot[inst];
    if (k < 0) return 0.0f;
    if (!(W > 0.0f)) return 0.0f;
    return emit_weight(S, (uint32_t)k, x, n) / W * S.emitInst[2 * k + 1].y;
}

// Density per steradian with which the emitter sampler at x (shading normal
// n, weight sum W) reaches triangle `prim` of instance `inst` at distance t
// along unit direction d: pdfA * t^2 / |cos| at the light, 0 if it is not
// sampled.
__device__ __forceinline__ float emit_pdf(const DevScene &S, uint32_t inst, uint32_t prim,
                                          float t, v3 d, v3 x, v3 n, float W)
{
    const float pa = emit_pdf_area(S, inst, prim, x, n, W);
    if (!(pa > 0.0f)) return 0.0f;
    v3 A, B, C;
    world_triangle(S, inst, prim, A, B, C);
    const v3 tn = cross(B - A, C - A);
    const float c = fabsf(dot(tn, d)) / sqrtf(dot(tn, tn));
    return c > 0.0f ? pa * (t * t) / c : 0.0f;
}

// A point on an emissive triangle (uniform on the triangle by e2, e3): the
// unit direction from x, its density per steradian, the triangle.  With
// the spatial choice (the scan `e` from x and nx) an emitter instance is
// drawn by its weight (the first whose running sum exceeds e4 x the total),
// then a triangle of it by area (entry by its alias table from r and e1);
// else an entry of the global table (r and e1).  pdf 0: nothing to sample.
__device__ __forceinline__ v3 emit_direction(const DevScene &S, const EmitScan &e,
                                             uint32_t r, float e1, float e2, float e3,
                                             float e4, v3 x, v3 nx, float &pdf,
                                             float &dist, uint32_t &inst, uint32_t &prim)
{
    uint32_t g;
    float pa;
    if (S.emitInstCount)
    {
        const float W = e.W;
        if (!(W > 0.0f))
        {
            pdf = 0.0f;
            dist = 0.0f;
            inst = prim = SPG_MISS;
            return v3{0.0f, 0.0f, 1.0f};
        }
        const float target = e4 * W;
        // the first group whose running sum exceeds the target and the sum
        // before it; none (the target rounds to W): the last positive weight
        int gsel = kEmitGroups;
        float base = 0.0f;
#pragma unroll
        for (int j = kEmitGroups - 1; j >= 0; --j)
        {
            const bool in = e.G[j] > target;
            gsel = in ? j : gsel;
            base = in ? (j > 0 ? e.G[j - 1] : 0.0f) : base;
        }
        uint32_t pick = e.lastPos;
        float wp = e.lastPosW;
        if (gsel < kEmitGroups)
        {
            // the crossing weight is positive (the running sum grows only
            // by positive weights), so it is the pick
            const uint32_t cnt = S.emitInstCount, gs = (cnt + kEmitGroups - 1) / kEmitGroups;
            const uint32_t k0 = (uint32_t)gsel * gs, k1 = min(cnt, k0 + gs);
            float acc = base;
            bool found = false;
            for (uint32_t k = k0; k < k1 && !found; ++k)
            {
                const float w = emit_weight(S, k, x, nx);
                acc += w;
                found = acc > target;
                pick = found ? k : pick;
                wp = found ? w : wp;
            }
        }
        const float4 b = S.emitInst[2 * pick + 1];
        const uint32_t first = __float_as_uint(b.z), count = __float_as_uint(b.w);
        uint32_t j = (uint32_t)(((uint64_t)r * count) >> 32);
        if (!(e1 < S.emitProbL[first + j])) j = S.emitAliasL[first + j];
        g = first + j;
        pa = wp / W * b.y;
    }
    else
    {
        g = (uint32_t)(((uint64_t)r * S.emitCount) >> 32);
        if (!(e1 < S.emitProb[g])) g = S.emitAlias[g];
        pa = S.emitPdfA[g];
    }
    inst = S.emitPair[2 * g];
    prim = S.emitPair[2 * g + 1];
    v3 A, B, C;
    world_triangle(S, inst, prim, A, B, C);
    const float su = sqrtf(e2);
    const v3 P = (A * (1.0f - su) + B * (su * (1.0f - e3))) + C * (su * e3);
    const v3 v = P - x;
    const float dist2 = dot(v, v);
    const v3 w = normalize(v);
    const v3 n = cross(B - A, C - A);
    const float fw = dot(n, w);
    const float c = fabsf(fw) / sqrtf(dot(n, n));
    // the closest-hit test accepts a triangle only when the ray meets its
    // front (Moller-Trumbore's dot(d, n) < 0 in the mesh's frame): one
    // clearly facing away (|cos| > 1e-3 at the light) is never reached, so
    // its sample casts no shadow ray (pdf 0, no contribution, as the ray's
    // answer would have been)
    const int32_t orient = S.inst[inst].orient;
    const bool away = orient != 0 && (orient > 0 ? fw > 0.0f : fw < 0.0f) && c > 1e-3f;
    pdf = c > 0.0f && dist2 > 0.0f && !away ? pa * dist2 / c : 0.0f;
    dist = sqrtf(dist2);
    return w;
}

struct MatOut
{
    v3 albedo, emission;
    float roughness;
};

// sp_EvaluateMaterial, src/sp_material_system.cpp:59-105 (emission texture:
// equirect by ToSphericalCoordinates/MapToEquirectangular, or a cube map by
// direction).  A missing material gives emission (1,0,1),
// src/simd_path_tracer.cpp:125-134.
template <bool kDouble>
__device__ __forceinline__ MatOut eval_material(const DevScene &S, int32_t slot,
                                                v2 uv, v3 outgoingDir)
{
    MatOut r;
    r.albedo = v3{0.0f, 0.0f, 0.0f};
    r.emission = v3{1.0f, 0.0f, 1.0f};
    r.roughness = 0.0f;
    if (slot < 0) return r;
    const DevMaterial &M = S.mats[slot];
    if (M.albedoTex >= 0 && !S.texs[M.albedoTex].isCube)
    {
        float4 t = sample_nearest(S, S.texs[M.albedoTex], uv.x, uv.y);
        r.albedo = v3{t.x, t.y, t.z};
    }
    else
        r.albedo = v3{M.albedo[0], M.albedo[1], M.albedo[2]};
    if (M.emissionTex >= 0)
    {
        const DevTexture &T = S.texs[M.emissionTex];
        v3 dir = -outgoingDir;
        float4 t;
        if (T.isCube)
            t = sample_cube(S, T, dir);
        else
        {
            // ToSphericalCoordinates, src/math_lib.h:849-860
            float inc = Tx<kDouble>::atan2_(sqrtf(dir.x * dir.x + dir.z * dir.z),
                                            dir.y);
            float az = Tx<kDouble>::atan2_(dir.z, dir.x);
            // MapToEquirectangular, src/math_lib.h:873-884
            if (az < 0.0f) az += 2.0f * kPi;
            float eu = az / (2.0f * kPi);
            float ev = Tx<kDouble>::cos_(inc) * 0.5f + 0.5f;
            ev = 1.0f - ev;
            t = sample_nearest(S, T, eu, ev);
        }
        r.emission = v3{t.x * 1.0f, t.y * 1.0f, t.z * 1.0f};
    }
    else
        r.emission = v3{M.emission[0], M.emission[1], M.emission[2]};
    r.roughness = M.roughness;
    return r;
}

// The bracket of ComputeRadianceForPath (src/simd_path_tracer.cpp:136-165):
// X = kD o albedo / pi + D G F / (4 (N.V)(N.L) + 1e-4); the vertex radiance
// is E + (X o Lin) * max(0, N.L).  noDeltaSpecular drops the GGX term at
// roughness 0 (unbiased estimator only; there D is 0 except on a measure-
// zero set where the reference computes 0/0).
template <bool kDouble>
__device__ __forceinline__ v3 brdf_bracket(v3 N, v3 V, v3 L, const MatOut &m,
                                           bool noDeltaSpecular)
{
    v3 H = normalize(L + V);
    float cosTheta = fmax_sel(dot(H, V), 0.0f);
    float p5 = Tx<kDouble>::pow5_(1.0f - cosTheta);
    v3 F0 = splat(0.04f);
    v3 F = F0 + (splat(1.0f) - F0) * p5;
    v3 kD = splat(1.0f) - F;
    float oneOverPI = 1.0f / kPi;
    float roughness = m.roughness;
    v3 specular{0.0f, 0.0f, 0.0f};
    if (!(noDeltaSpecular && roughness == 0.0f))
    {
        // DistributionGGX, src/simd_path_tracer.cpp:72-84
        float a = roughness * roughness;
        float a2 = a * a;
        float NdotH = fmax_sel(dot(N, H), 0.0f);
        float NdotH2 = NdotH * NdotH;
        float denom = (NdotH2 * (a2 - 1.0f) + 1.0f);
        denom = kPi * denom * denom;
        float NDF = a2 / denom;
        // GeometrySmith / GeometrySchlickGGX, src/simd_path_tracer.cpp:86-105
        float NdotV = fmax_sel(dot(N, V), 0.0f);
        float NdotL = fmax_sel(dot(N, L), 0.0f);
        float r = (roughness + 1.0f);
        float k = (r * r) / 8.0f;
        float ggx2 = NdotV / (NdotV * (1.0f - k) + k);
        float ggx1 = NdotL / (NdotL * (1.0f - k) + k);
        float G = ggx1 * ggx2;
        v3 numerator = F * (NDF * G);
        float denominator =
            4.0f * fmax_sel(dot(N, V), 0.0f) * fmax_sel(dot(N, L), 0.0f) +
            0.0001f;
        specular = numerator * (1.0f / denominator);
    }
    return had(kD, m.albedo) * oneOverPI + specular;
}

// --------------------------------------------------------------- RNG
struct Rng
{
    uint32_t xs;
    uint64_t state, inc;
    bool pcg;
    __device__ __forceinline__ uint32_t next()
    {
        if (pcg) return pcg32_next(state, inc);
        return xorshift32(xs);
    }
    __device__ __forceinline__ float uni() { return unilateral(next()); }
    __device__ __forceinline__ float bi() { return -1.0f + 2.0f * uni(); }
    __device__ __forceinline__ float u24()
    {
        return (float)(next() >> 8) * (1.0f / 16777216.0f);
    }
};

__device__ __forceinline__ Rng load_rng(const TraceParams &p, uint32_t path,
                                        uint32_t pixel)
{
    Rng r;
    r.pcg = p.rngMode == SPG_RNG_PCG;
    if (r.pcg)
    {
        r.state = p.paths.rngPcg[path];
        r.inc = ((uint64_t)pixel << 1u) | 1u;
        r.xs = 0;
    }
    else
    {
        r.xs = p.paths.rngXs[path];
        r.state = r.inc = 0;
    }
    return r;
}

__device__ __forceinline__ void store_rng(const TraceParams &p, uint32_t path,
                                          const Rng &r)
{
    if (r.pcg) p.paths.rngPcg[path] = r.state;
    else p.paths.rngXs[path] = r.xs;
}

// ------------------------------------------------------------- kernels
// Closest hit of a ray list against one mesh (mesh space).
template <bool kCount>
__global__ __launch_bounds__(kTraceBlock) void k_closest_mesh(
    DevMesh m, const float *__restrict__ ox, const float *__restrict__ oy,
    const float *__restrict__ oz, const float *__restrict__ dx,
    const float *__restrict__ dy, const float *__restrict__ dz, uint32_t n,
    uint32_t *__restrict__ primOut, float *__restrict__ tOut,
    float *__restrict__ uOut, float *__restrict__ vOut,
    unsigned long long *__restrict__ stats, uint32_t *counter, int *spill)
{
    Counts cnt{0, 0, 0, 0};
    uint32_t rays = 0, hits = 0;
    auto begin = [&](uint32_t i, Trav &s) {
        trav_begin_mesh(m, s, v3{ox[i], oy[i], oz[i]}, v3{dx[i], dy[i], dz[i]});
        return true;
    };
    auto store = [&](uint32_t i, const Trav &s) {
        primOut[i] = s.h.prim;
        tOut[i] = s.h.prim == SPG_MISS ? -1.0f : s.h.t;
        if (uOut) uOut[i] = s.h.u;
        if (vOut) vOut[i] = s.h.v;
        if (kCount && s.h.prim != SPG_MISS) hits++;
    };
    SceneView sv{};
    persistent_walk<kCount, false, kMeshShortStack, kMeshChunk>(
        sv, n, counter, spill, kMeshRefill, begin, store, cnt, rays);
    if (kCount)
    {
        wave_add(&stats[0], rays);
        wave_add(&stats[1], cnt.box);
        wave_add(&stats[2], cnt.tri);
        wave_add(&stats[4], hits);
    }
}

// sp_RayIntersectScene for a ray list, with the surface at the hit.
__global__ __launch_bounds__(kTraceBlock) void k_intersect_scene(
    DevScene S, const float *__restrict__ rays6, uint32_t n,
    float *__restrict__ out7, uint32_t *__restrict__ objOut,
    uint32_t *__restrict__ primOut, uint32_t *counter, int *spill)
{
    Counts cnt{0, 0, 0, 0};
    uint32_t rays = 0;
    SceneView sv = scene_view(S);
    auto begin = [&](uint32_t i, Trav &s) {
        trav_begin_scene(s, v3{rays6[6 * i], rays6[6 * i + 1], rays6[6 * i + 2]},
                         v3{rays6[6 * i + 3], rays6[6 * i + 4], rays6[6 * i + 5]});
        return S.instCount > 0;
    };
    auto store = [&](uint32_t i, const Trav &s) {
        const SceneHit &h = s.out;
        float *r = out7 + 7 * (size_t)i;
        if (h.inst == SPG_MISS)
        {
            r[0] = -1.0f;
            for (int k = 1; k < 7; ++k) r[k] = 0.0f;
        }
        else
        {
            v3 N;
            v2 uv;
            surface_at(S, h.inst, h.prim, h.u, h.v, N, uv);
            r[0] = h.t;
            r[1] = __uint_as_float(S.inst[h.inst].materialId);
            r[2] = N.x;
            r[3] = N.y;
            r[4] = N.z;
            r[5] = uv.x;
            r[6] = uv.y;
        }
        if (objOut) objOut[i] = h.inst;
        if (primOut) primOut[i] = h.prim;
    };
    persistent_walk<false, true, kSceneShortStack, kSceneChunk>(
        sv, n, counter, spill, kSceneRefill, begin, store, cnt, rays);
}

// Logical block index spread so consecutive ranges of work stay on one XCD
// (blocks are dispatched to the 8 XCDs round-robin).
__device__ __forceinline__ uint32_t xcd_block()
{
    uint32_t g = gridDim.x, b = blockIdx.x;
    if (g % 8u) return b;
    return (b % 8u) * (g / 8u) + b / 8u;
}

// Camera ray of one path: the jitter draws (y drawn first: the reference's
// two RandomBilateral arguments, src/simd_path_tracer.cpp:219-222, evaluate
// right to left as g++ on x86-64 builds them), sp_CalculateFilmPositions
// (src/simd_path_tracer.cpp:38-63), dir = Normalize(filmP - position).
template <bool kBatch>
__global__ __launch_bounds__(256) void k_raygen(TraceParams p)
{
    const uint32_t nPaths = p.slots * p.sppPass;
    if (blockIdx.x == 0 && threadIdx.x == 0)
    {
        // this pass's camera rays fill queue 0; the first k_extend starts
        // its work counter at zero
        p.paths.counters[0] = nPaths;
        *p.work = 0u;
    }
    for (uint32_t path = xcd_block() * 256 + threadIdx.x; path < nPaths;
         path += gridDim.x * 256)
    {
        uint32_t slot = path / p.sppPass;
        uint32_t sample = p.sampleBase + path % p.sppPass;
        uint32_t pixel = p.slotPixel[slot];
        const DevCamera &cam = kBatch ? p.cams[p.slotItem[slot]] : p.cam;
        uint32_t x = pixel % cam.width, y = pixel / cam.width;
        Rng rng;
        rng.pcg = p.rngMode == SPG_RNG_PCG;
        if (rng.pcg)
        {
            pcg32_seed(rng.state, rng.inc, p.seed, pixel);
            pcg32_advance(rng.state, rng.inc, (uint64_t)sample << 20);
            rng.xs = 0;
        }
        else
        {
            rng.xs = sample == 0 ? pixel_seed(p.seed, pixel)
                                 : p.paths.slotRng[slot];
            rng.state = rng.inc = 0;
        }
        float b1 = rng.bi();
        float b2 = rng.bi();
        float jy = cam.halfPixelHeight * b1;
        float jx = cam.halfPixelWidth * b2;
        float px = (float)x + 0.5f, py = (float)y + 0.5f;
        px = px + jx;
        py = py + jy;
        float fx = px / (float)cam.width;
        float fy = py / (float)cam.height;
        fy = 1.0f - fy;
        fx = fx * 2.0f - 1.0f;
        fy = fy * 2.0f - 1.0f;
        v3 film = cam.right * (cam.halfFilmWidth * fx);
        film = film + cam.up * (cam.halfFilmHeight * fy);
        film = film + cam.filmCenter;
        v3 o = cam.position;
        v3 d = normalize(film - cam.position);
        const uint32_t cap = p.paths.capacity;
        p.paths.qo[0][path] = o.x;
        p.paths.qo[0][cap + path] = o.y;
        p.paths.qo[0][2 * cap + path] = o.z;
        p.paths.qd[0][path] = d.x;
        p.paths.qd[0][cap + path] = d.y;
        p.paths.qd[0][2 * cap + path] = d.z;
        p.paths.qpath[0][path] = path;
        p.paths.pathSlot[path] = slot;
        p.paths.pathSample[path] = sample;
        p.paths.vcount[path] = 0;
        p.paths.beta[path] = 1.0f;
        p.paths.beta[cap + path] = 1.0f;
        p.paths.beta[2 * cap + path] = 1.0f;
        p.paths.radiance[path] = 0.0f;
        p.paths.radiance[cap + path] = 0.0f;
        p.paths.radiance[2 * cap + path] = 0.0f;
        store_rng(p, path, rng);
    }
}

// Closest hit of every queued ray.  (Forcing a register budget of 6, 7 or
// 8 resident waves per SIMD instead of the 5 that 84 registers give was
// measured slower on the 4M-triangle box scene: 1.44, 1.80 and 2.43 ms per
// launch against 1.41; the step is bound by instruction issue, and 7 and 8
// spill.)
// kShadow: the shadow rays of next-event estimation (queue 2).  A ray
// toward the sky only asks whether anything is hit, so it stops at the
// first instance with a hit; a ray toward an emissive triangle starts with
// its best t at 1.001 x the light's distance, so nothing beyond the light
// is visited (the closest hit below it is still found: the ray reaches its
// light iff that hit is the light's triangle).
template <bool kCount, bool kBatch, bool kShadow>
__global__ __launch_bounds__(kTraceBlock) void k_extend(TraceParams p, int q,
                                                        uint32_t *counter,
                                                        int *spill)
{
    const uint32_t n = p.paths.counters[q];
    const uint32_t cap = p.paths.capacity;
    const float *qo = p.paths.qo[q], *qd = p.paths.qd[q];
    if (!kShadow && blockIdx.x == 0 && threadIdx.x == 0)
    {
        // the queues the following k_shade appends to start empty (their
        // previous contents were consumed by the bounce before)
        p.paths.counters[q ^ 1] = 0u;
        p.paths.counters[2] = 0u;
    }
    Counts cnt{0, 0, 0, 0};
    uint32_t rays = 0;
    SceneView sv = scene_view(p.scene);
    auto begin = [&](uint32_t i, Trav &s) {
        trav_begin_scene(s, v3{qo[i], qo[cap + i], qo[2 * cap + i]},
                         v3{qd[i], qd[cap + i], qd[2 * cap + i]});
        if (kShadow)
        {
            const uint32_t path = p.paths.qpath[2][i];
            if (p.paths.neeTarget[path] == SPG_MISS) s.anyHit = true;
            else s.out.t = p.paths.neeDist[path];
        }
        if (kBatch)
        {
            const DevScene &S =
                p.scenes[p.slotItem[p.paths.pathSlot[p.paths.qpath[q][i]]]];
            sv = scene_view(S);
            return S.instCount > 0;
        }
        return p.scene.instCount > 0;
    };
    auto store = [&](uint32_t i, const Trav &s) {
        p.paths.ht[i] = s.out.t;
        p.paths.hu[i] = s.out.u;
        p.paths.hv[i] = s.out.v;
        p.paths.hinst[i] = s.out.inst;
        p.paths.hprim[i] = s.out.prim;
    };
    const uint32_t refill = p.refill >= 1u && p.refill <= 64u ? p.refill : kSceneRefill;
    persistent_walk<kCount, true, kSceneShortStack, kSceneChunk>(
        sv, n, counter, spill, refill, begin, store, cnt, rays);
    if (kCount)
    {
        wave_add(&p.paths.metrics[3], cnt.box);
        wave_add(&p.paths.metrics[4], cnt.mesh);
        wave_add(&p.paths.metrics[5], cnt.tri);
        wave_add(&p.paths.metrics[6], cnt.tlasBox);
    }
    wave_add(&p.paths.metrics[0], rays);
}

// Probability of sampling the GGX lobe instead of the cosine lobe for a
// material with roughness > 0 (one-sample MIS of the two).
constexpr float kSpecProb = 0.25f;

// Direction sampling of the two lobes with one sincos of phi = 2 pi u2 and
// one orthonormal frame about n (Duff et al. 2017, "Building an
// Orthonormal Basis, Revisited"):
//   cosine lobe: r = sqrt(u1), z = sqrt(max(0, 1 - u1)), direction
//     r cos phi t1 + r sin phi t2 + z n;
//   GGX half-vector sampling (Walter et al. 2007, "Microfacet Models for
//     Refraction through Rough Surfaces"): cos^2 theta_h = (1 - u1) /
//     (1 + (a^2 - 1) u1) with a = roughness^2 (the reference's
//     DistributionGGX parameter), H in the same frame, L = 2 (V.H) H - V.
template <bool kDouble>
__device__ __forceinline__ v3 lobe_direction(v3 n, v3 V, float roughness, bool ggx,
                                             float u1, float u2)
{
    float phi = 6.28318548f * u2;
    float sp, cp;
    Tx<kDouble>::sincos_(phi, &sp, &cp);
    float rad, ax;
    if (ggx)
    {
        float a = roughness * roughness;
        float a2 = a * a;
        float c2 = (1.0f - u1) / (1.0f + (a2 - 1.0f) * u1);
        ax = sqrtf(c2);
        rad = sqrtf(fmax_sel(0.0f, 1.0f - c2));
    }
    else
    {
        rad = sqrtf(u1);
        ax = sqrtf(fmax_sel(0.0f, 1.0f - u1));
    }
    float x = rad * cp;
    float y = rad * sp;
    float sign = copysignf(1.0f, n.z);
    float aa = -1.0f / (sign + n.z);
    float b = n.x * n.y * aa;
    v3 t1{1.0f + sign * n.x * n.x * aa, sign * b, -sign * n.x};
    v3 t2{b, sign + n.y * n.y * aa, -n.y};
    v3 D = (t1 * x + t2 * y) + n * ax;
    return ggx ? D * (2.0f * dot(V, D)) - V : D;
}

// Density of ggx_direction: D(H) (N.H) / (4 V.H), H = normalize(L + V).
__device__ __forceinline__ float ggx_pdf(v3 n, v3 V, v3 L, float roughness)
{
    v3 H = normalize(L + V);
    float a = roughness * roughness;
    float a2 = a * a;
    float NdotH = fmax_sel(dot(n, H), 0.0f);
    float NdotH2 = NdotH * NdotH;
    float denom = (NdotH2 * (a2 - 1.0f) + 1.0f);
    denom = kPi * denom * denom;
    float D = a2 / denom;
    float VdotH = dot(V, H);
    return VdotH > 0.0f ? D * NdotH / (4.0f * VdotH) : 0.0f;
}

// The BSDF sampling density: (1 - ps) cos/pi + ps ggx.
__device__ __forceinline__ float bsdf_pdf(v3 n, v3 V, v3 L, float cosv, float roughness,
                                          float ps)
{
    float pc = (1.0f - ps) * cosv * (1.0f / kPi);
    return ps > 0.0f ? pc + ps * ggx_pdf(n, V, L, roughness) : pc;
}

// Verification hook of the guided estimator: appends one scatter event
// (pn, N, V, roughness, the three uniforms, the mixture, the sampled
// direction and its density, the hit point) to the guide's record buffer.
__device__ __forceinline__ void record_scatter(const DevGuide &G, v3 pn, v3 N, v3 V,
                                            float rough, float u0, float u1, float u2,
                                            const LobeRef &L, v3 wi, float pdf, v3 P)
{
    const uint32_t at = atomicAdd(G.recCount, 1u);
    if (at >= G.recCap) return;
    float *r = G.rec + (size_t)at * kGRecordFloats;
    const float head[13] = {pn.x, pn.y, pn.z, N.x, N.y, N.z, V.x, V.y, V.z,
                            rough, u0, u1, u2};
    for (int k = 0; k < 13; ++k) r[k] = head[k];
    for (int k = 0; k < kGLobeFloats; ++k) r[13 + k] = L.at(k);
    r[53] = wi.x;
    r[54] = wi.y;
    r[55] = wi.z;
    r[56] = pdf;
    r[57] = P.x;
    r[58] = P.y;
    r[59] = P.z;
    for (int k = 60; k < kGRecordFloats; ++k) r[k] = 0.0f;
}

// One light sample of next-event estimation at x (shading normal N, view V,
// material m): the sky with probability envSelect, else an emitter
// (emit_direction).  Draw order: a 24-bit uniform choosing sky or emitter,
// one 32-bit draw for the table entry, four 24-bit uniforms.  With a usable
// direction the contribution beta o (f cos Le w / p_light) -- w the power
// heuristic against the scattering density scatterPdf(wl, cos) -- waits in
// the path's NEE slot with its target (SPG_MISS: the sky) and the function
// returns true with the shadow ray (so, sd).
template <bool kDouble, class ScatterPdf>
__device__ __forceinline__ bool nee_light_sample(const TraceParams &p, const DevScene &S,
                                                 uint32_t path, Rng &rng, v3 x, v3 N,
                                                 v3 V, const MatOut &m, v3 beta,
                                                 ScatterPdf scatterPdf, v3 &so, v3 &sd)
{
    const uint32_t cap = p.paths.capacity;
    float e0 = rng.u24();
    uint32_t r = rng.next();
    float e1 = rng.u24(), e2 = rng.u24(), e3 = rng.u24();
    float e4 = rng.u24();
    // the spatial choice's weights from x, whichever light is drawn: their
    // sum is kept for the MIS weight of an emitter that the scattered ray
    // may reach next (the same x and N there give the same sum)
    EmitScan scan;
    scan.W = 0.0f;
    if (S.emitInstCount)
    {
        emit_scan(S, x, N, scan);
        p.paths.lastW[path] = scan.W;
    }
    const bool useEnv = e0 < S.envSelect;
    if (!(useEnv ? S.envTex >= 0 : S.emitCount > 0)) return false;
    v3 wl, Le;
    float pl;
    uint32_t ti = SPG_MISS, tp = 0;
    if (useEnv)
    {
        wl = env_direction(S, r, e1, e2, e3);
        pl = S.envSelect * env_pdf(S, wl);
        v2 uv0{0.0f, 0.0f};
        Le = eval_material<kDouble>(S, S.background, uv0, -wl).emission;
    }
    else
    {
        float pg, dl;
        wl = emit_direction(S, scan, r, e1, e2, e3, e4, x, N, pg, dl, ti, tp);
        p.paths.neeDist[path] = dl * 1.001f;
        pl = (1.0f - S.envSelect) * pg;
        Le = v3{0.0f, 0.0f, 0.0f};
        // pg = 0: nothing was sampled and ti is SPG_MISS
        if (pg > 0.0f)
        {
            const float *em = S.mats[S.inst[ti].material].emission;
            Le = v3{em[0], em[1], em[2]};
        }
    }
    float cl = dot(N, wl);
    if (!(cl > 0.0f && pl > 0.0f)) return false;
    v3 Xl = brdf_bracket<kDouble>(N, V, wl, m, true);
    float pb = scatterPdf(wl, cl);
    float pl2 = pl * pl;
    float wmis = pl2 / (pl2 + pb * pb);
    v3 C = had(had(beta, Xl * ((cl * wmis) / pl)), Le);
    p.paths.nee[path] = C.x;
    p.paths.nee[cap + path] = C.y;
    p.paths.nee[2 * cap + path] = C.z;
    p.paths.neeTarget[path] = ti;
    p.paths.neeTarget[cap + path] = tp;
    so = x;
    sd = wl;
    return true;
}

// Power-heuristic share of emission reached by a scattered ray (density pb
// at the previous vertex) against the light sampler: toward the emitter
// triangle (inst, prim) at distance t along d from o, or (inst SPG_MISS)
// the sky along d.  The ray left the previous vertex from the point where
// that vertex's light sample was drawn (o), with the shading normal and
// the spatial choice's weight sum stored there.
__device__ __forceinline__ float nee_hit_weight(const TraceParams &p, const DevScene &S,
                                                uint32_t path, uint32_t inst, uint32_t prim,
                                                float t, v3 d, v3 o)
{
    const uint32_t cap = p.paths.capacity;
    float pl;
    if (inst == SPG_MISS)
        pl = S.envSelect * env_pdf(S, d);
    else
    {
        const v3 nPrev{p.paths.lastN[path], p.paths.lastN[cap + path],
                       p.paths.lastN[2 * cap + path]};
        pl = (1.0f - S.envSelect) *
             emit_pdf(S, inst, prim, t, d, o, nPrev, p.paths.lastW[path]);
    }
    float pb = p.paths.lastPdf[path];
    float pb2 = pb * pb;
    return pb2 / (pb2 + pl * pl);
}

// One bounce of every queued path: surface, material, scatter, vertex record
// (parity) or throughput update (unbiased), append continuing rays to queue
// q^1 with a ballot + one atomic per wave.
template <bool kDouble, int kEst, bool kBatch>
__global__ __launch_bounds__(256) void k_shade(TraceParams p, int q)
{
    const uint32_t n = p.paths.counters[q];
    // k_extend has finished with the work counter: zero it for the next one
    if (blockIdx.x == 0 && threadIdx.x == 0) *p.work = 0u;
    const uint32_t cap = p.paths.capacity;
    const int nq = q ^ 1;
    uint32_t hits = 0, misses = 0;
    const uint32_t stride = gridDim.x * 256;
    for (uint32_t i0 = xcd_block() * 256; i0 < n; i0 += stride)
    {
        uint32_t i = i0 + threadIdx.x;
        bool push = false, shadow = false;
        v3 no{0, 0, 0}, nd{0, 0, 0}, so{0, 0, 0}, sd{0, 0, 0};
        uint32_t path = 0;
        if (i < n)
        {
            path = p.paths.qpath[q][i];
            uint32_t slot = p.paths.pathSlot[path];
            uint32_t pixel = p.slotPixel[slot];
            const DevScene &S = kBatch ? p.scenes[p.slotItem[slot]] : p.scene;
            Rng rng = load_rng(p, path, pixel);
            v3 o{p.paths.qo[q][i], p.paths.qo[q][cap + i], p.paths.qo[q][2 * cap + i]};
            v3 d{p.paths.qd[q][i], p.paths.qd[q][cap + i], p.paths.qd[q][2 * cap + i]};
            float t = p.paths.ht[i];
            uint32_t inst = p.paths.hinst[i];
            uint32_t vc = p.paths.vcount[path];
            v3 V = -d;
            bool hit = inst != SPG_MISS && t > 0.0f;
            if (kEst == SPG_EST_PARITY)
            {
                v3 E, X;
                float cosine;
                if (hit)
                {
                    hits++;
                    v3 N;
                    v2 uv;
                    surface_at(S, inst, p.paths.hprim[i], p.paths.hu[i],
                               p.paths.hv[i], N, uv);
                    v3 P = o + d * t;
                    // RandomDirectionOnHemisphere, src/math_lib.h:932-946
                    float theta = kPi * rng.uni();
                    float phi = kPi * rng.bi();
                    v3 dir;
                    float st, ct, sp, cp;
                    Tx<kDouble>::sincos_(theta, &st, &ct);
                    Tx<kDouble>::sincos_(phi, &sp, &cp);
                    dir.x = st * cp;
                    dir.z = st * sp;
                    dir.y = ct;
                    if (dot(dir, N) < 0.0f) dir = -dir;
                    MatOut m = eval_material<kDouble>(
                        S, S.inst[inst].material, uv, V);
                    E = m.emission;
                    X = brdf_bracket<kDouble>(N, V, dir, m, false);
                    cosine = fmax_sel(0.0f, dot(N, dir));
                    if (vc + 1 < p.bounces)
                    {
                        push = true;
                        no = P + N * 0.0001f;
                        nd = dir;
                    }
                }
                else
                {
                    misses++;
                    v2 uv0{0.0f, 0.0f};
                    MatOut m = eval_material<kDouble>(S, S.background,
                                                      uv0, V);
                    v3 Z{0.0f, 0.0f, 0.0f};
                    E = m.emission;
                    X = brdf_bracket<kDouble>(Z, V, Z, m, false);
                    cosine = fmax_sel(0.0f, dot(Z, Z));
                }
                float *vt = p.paths.vtx + (size_t)vc * 7 * cap + path;
                vt[0] = E.x;
                vt[cap] = E.y;
                vt[2 * cap] = E.z;
                vt[3 * cap] = X.x;
                vt[4 * cap] = X.y;
                vt[5 * cap] = X.z;
                vt[6 * cap] = cosine;
                p.paths.vcount[path] = vc + 1;
            }
            else if (kEst == SPG_EST_UNBIASED || kEst == SPG_EST_NEE)
            {
                v3 beta{p.paths.beta[path], p.paths.beta[cap + path],
                        p.paths.beta[2 * cap + path]};
                v3 L{p.paths.radiance[path], p.paths.radiance[cap + path],
                     p.paths.radiance[2 * cap + path]};
                if (hit)
                {
                    hits++;
                    v3 N;
                    v2 uv;
                    surface_at(S, inst, p.paths.hprim[i], p.paths.hu[i],
                               p.paths.hv[i], N, uv);
                    MatOut m = eval_material<kDouble>(
                        S, S.inst[inst].material, uv, V);
                    if (kEst == SPG_EST_NEE && vc > 0 && S.emitCount > 0 &&
                        S.emitBase[inst] >= 0)
                    {
                        // a BSDF-sampled ray that reached an emitter: its
                        // share against the emitter sampler (power heuristic)
                        float w = nee_hit_weight(p, S, path, inst, p.paths.hprim[i], t, d, o);
                        L = L + had(beta, m.emission * w);
                    }
                    else
                        L = L + had(beta, m.emission);
                    // cosine lobe, or GGX with probability ps on rough
                    // materials; weight X cos / pdf (X pi for ps = 0)
                    const float ps = m.roughness > 0.0f ? kSpecProb : 0.0f;
                    float u0 = ps > 0.0f ? rng.u24() : 1.0f;
                    float u1 = rng.u24(), u2 = rng.u24();
                    v3 wi = lobe_direction<kDouble>(N, V, m.roughness, u0 < ps, u1, u2);
                    float cosine = dot(N, wi);
                    if (kEst == SPG_EST_NEE && vc + 1 < p.bounces)
                    {
                        // next-event estimation against the BSDF sampler's
                        // density; the contribution waits for the shadow
                        // ray's answer
                        const v3 x = (o + d * t) + N * 0.0001f;
                        shadow = nee_light_sample<kDouble>(
                            p, S, path, rng, x, N, V, m, beta,
                            [&](v3 wl, float cl) {
                                return bsdf_pdf(N, V, wl, cl, m.roughness, ps);
                            },
                            so, sd);
                    }
                    if (cosine > 0.0f && vc + 1 < p.bounces)
                    {
                        v3 X = brdf_bracket<kDouble>(N, V, wi, m, true);
                        float pdfc = ps > 0.0f
                                         ? bsdf_pdf(N, V, wi, cosine, m.roughness, ps)
                                         : cosine * (1.0f / kPi);
                        float w = ps > 0.0f ? cosine / pdfc : kPi;
                        beta = had(beta, X * w);
                        if (kEst == SPG_EST_NEE)
                        {
                            p.paths.lastPdf[path] = pdfc;
                            p.paths.lastN[path] = N.x;
                            p.paths.lastN[cap + path] = N.y;
                            p.paths.lastN[2 * cap + path] = N.z;
                        }
                        v3 P = o + d * t;
                        push = true;
                        no = P + N * 0.0001f;
                        nd = wi;
                    }
                }
                else
                {
                    misses++;
                    v2 uv0{0.0f, 0.0f};
                    MatOut m = eval_material<kDouble>(S, S.background,
                                                      uv0, V);
                    if (kEst == SPG_EST_NEE && vc > 0)
                    {
                        // the BSDF-sampled direction's share of the
                        // environment (power heuristic)
                        float w = nee_hit_weight(p, S, path, SPG_MISS, 0, t, d, o);
                        L = L + had(beta, m.emission * w);
                    }
                    else
                        L = L + had(beta, m.emission);
                }
                p.paths.beta[path] = beta.x;
                p.paths.beta[cap + path] = beta.y;
                p.paths.beta[2 * cap + path] = beta.z;
                p.paths.radiance[path] = L.x;
                p.paths.radiance[cap + path] = L.y;
                p.paths.radiance[2 * cap + path] = L.z;
                p.paths.vcount[path] = vc + 1;
            }
            else
            {
                // Learned guiding: one-sample MIS of cosine sampling
                // (probability fb) and the guide's vMF mixture, pdf
                // fb cos/pi + (1 - fb) q(wi); the vertex record is
                // (E, X, cos/pdf) for the reverse recurrence of k_finalize.
                // With next-event estimation (SPG_EST_GUIDED_NEE) emission
                // reached by a scattered ray carries its power-heuristic
                // share against the light sampler, and each scattering
                // vertex takes one light sample weighted against that
                // scattering density; k_resolve_shadow adds it to the
                // vertex's E once its shadow ray reaches the light.
                constexpr bool kNee = kEst == SPG_EST_GUIDED_NEE;
                v3 E{0.0f, 0.0f, 0.0f}, X{0.0f, 0.0f, 0.0f};
                float c = 0.0f;
                if (hit)
                {
                    hits++;
                    v3 N;
                    v2 uv;
                    surface_at(S, inst, p.paths.hprim[i], p.paths.hu[i],
                               p.paths.hv[i], N, uv);
                    MatOut m = eval_material<kDouble>(
                        S, S.inst[inst].material, uv, V);
                    E = m.emission;
                    if (kNee && vc > 0 && S.emitCount > 0 && S.emitBase[inst] >= 0)
                        E = E * nee_hit_weight(p, S, path, inst, p.paths.hprim[i], t, d, o);
                    if (vc + 1 < p.bounces)
                    {
                        const LobeRef Lg{p.guide.lobes, cap, i};
                        float u0 = rng.u24(), u1 = rng.u24(), u2 = rng.u24();
                        const float fb = p.guide.bsdfFraction;
                        const float ps = m.roughness > 0.0f ? kSpecProb : 0.0f;
                        v3 wi;
                        if (u0 < fb)
                            wi = lobe_direction<kDouble>(N, V, m.roughness, u0 < fb * ps,
                                                         u1, u2);
                        else
                            wi = mixture_sample<kDouble>(Lg, (u0 - fb) / (1.0f - fb), u1,
                                                         u2);
                        float cosv = dot(N, wi);
                        float pdf = cosv > 0.0f
                                        ? fb * bsdf_pdf(N, V, wi, cosv, m.roughness, ps) +
                                              (1.0f - fb) * mixture_pdf<kDouble>(Lg, wi)
                                        : 0.0f;
                        if (p.guide.rec)
                            record_scatter(p.guide, guide_pn(p.guide, o + d * t), N, V,
                                           m.roughness, u0, u1, u2, Lg, wi, pdf, o + d * t);
                        if (kNee)
                        {
                            const v3 x = (o + d * t) + N * 0.0001f;
                            const v3 one{1.0f, 1.0f, 1.0f};
                            shadow = nee_light_sample<kDouble>(
                                p, S, path, rng, x, N, V, m, one,
                                [&](v3 wl, float cl) {
                                    return fb * bsdf_pdf(N, V, wl, cl, m.roughness, ps) +
                                           (1.0f - fb) * mixture_pdf<kDouble>(Lg, wl);
                                },
                                so, sd);
                        }
                        if (cosv > 0.0f && pdf > 0.0f)
                        {
                            X = brdf_bracket<kDouble>(N, V, wi, m, true);
                            c = cosv / pdf;
                            v3 P = o + d * t;
                            push = true;
                            no = P + N * 0.0001f;
                            nd = wi;
                            if (kNee)
                            {
                                p.paths.lastPdf[path] = pdf;
                                p.paths.lastN[path] = N.x;
                                p.paths.lastN[cap + path] = N.y;
                                p.paths.lastN[2 * cap + path] = N.z;
                            }
                            if (p.guide.train)
                            {
                                v3 pn = guide_pn(p.guide, P);
                                float *gv = p.guide.gvtx + (size_t)vc * cap + path;
                                const size_t st = (size_t)p.bounces * cap;
                                gv[0] = pn.x;
                                gv[st] = pn.y;
                                gv[2 * st] = pn.z;
                                gv[3 * st] = N.x;
                                gv[4 * st] = N.y;
                                gv[5 * st] = N.z;
                                gv[6 * st] = V.x;
                                gv[7 * st] = V.y;
                                gv[8 * st] = V.z;
                                gv[9 * st] = wi.x;
                                gv[10 * st] = wi.y;
                                gv[11 * st] = wi.z;
                            }
                        }
                    }
                }
                else
                {
                    misses++;
                    v2 uv0{0.0f, 0.0f};
                    MatOut m = eval_material<kDouble>(S, S.background,
                                                      uv0, V);
                    E = m.emission;
                    if (kNee && vc > 0) E = E * nee_hit_weight(p, S, path, SPG_MISS, 0, t, d, o);
                }
                float *vt = p.paths.vtx + (size_t)vc * 7 * cap + path;
                vt[0] = E.x;
                vt[cap] = E.y;
                vt[2 * cap] = E.z;
                vt[3 * cap] = X.x;
                vt[4 * cap] = X.y;
                vt[5 * cap] = X.z;
                vt[6 * cap] = c;
                p.paths.vcount[path] = vc + 1;
            }
            store_rng(p, path, rng);
        }
        // compaction of continuing rays into queue nq (and of shadow rays
        // into queue 2): one atomic per block and queue
        const uint32_t lane = lane_id();
        const unsigned long long mask = __ballot(push);
        constexpr bool kShadow = kEst == SPG_EST_NEE || kEst == SPG_EST_GUIDED_NEE;
        const unsigned long long smask = kShadow ? __ballot(shadow) : 0ull;
        uint32_t base, sbase;
        block_queue_bases(&p.paths.counters[nq], (uint32_t)__popcll(mask),
                          kShadow ? &p.paths.counters[2] : nullptr,
                          (uint32_t)__popcll(smask), base, sbase);
        if (push)
        {
            uint32_t slotq = base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
            p.paths.qo[nq][slotq] = no.x;
            p.paths.qo[nq][cap + slotq] = no.y;
            p.paths.qo[nq][2 * cap + slotq] = no.z;
            p.paths.qd[nq][slotq] = nd.x;
            p.paths.qd[nq][cap + slotq] = nd.y;
            p.paths.qd[nq][2 * cap + slotq] = nd.z;
            p.paths.qpath[nq][slotq] = path;
        }
        if (kShadow)
        {
            if (shadow)
            {
                uint32_t k = sbase + (uint32_t)__popcll(smask & ((1ull << lane) - 1ull));
                p.paths.qo[2][k] = so.x;
                p.paths.qo[2][cap + k] = so.y;
                p.paths.qo[2][2 * cap + k] = so.z;
                p.paths.qd[2][k] = sd.x;
                p.paths.qd[2][cap + k] = sd.y;
                p.paths.qd[2][2 * cap + k] = sd.z;
                p.paths.qpath[2][k] = path;
            }
        }
    }
    block_add2(&p.paths.metrics[1], hits, misses);
}

// Shadow rays of next-event estimation (queue 2, traced by k_extend): one
// that reaches its light (the sky: no hit; an emissive triangle: that
// triangle is the closest hit) adds its pending contribution to the path's
// radiance (SPG_EST_NEE) or to the emission term of the vertex that took
// the sample (SPG_EST_GUIDED_NEE).
__global__ __launch_bounds__(256) void k_resolve_shadow(TraceParams p)
{
    const uint32_t n = p.paths.counters[2];
    if (blockIdx.x == 0 && threadIdx.x == 0) *p.work = 0u;  // for the next k_extend
    const uint32_t cap = p.paths.capacity;
    unsigned long long lum[2] = {0ull, 0ull};  // learned light split: sky, emitters
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
    {
        const uint32_t path = p.paths.qpath[2][i];
        const uint32_t ti = p.paths.neeTarget[path], tp = p.paths.neeTarget[cap + path];
        const bool hitAny = p.paths.hinst[i] != SPG_MISS && p.paths.ht[i] > 0.0f;
        // the sky is seen through a miss, a light triangle by hitting it first
        const bool visible = ti == SPG_MISS
                                 ? !hitAny
                                 : hitAny && p.paths.hinst[i] == ti && p.paths.hprim[i] == tp;
        if (!visible) continue;
        const v3 C{p.paths.nee[path], p.paths.nee[cap + path], p.paths.nee[2 * cap + path]};
        // the unbiased estimator's radiance, or the guided one's record of
        // the vertex that took the light sample (the last one written)
        float *acc = p.paths.radiance + path;
        if (p.estimator == SPG_EST_GUIDED_NEE)
            acc = p.paths.vtx + (size_t)(p.paths.vcount[path] - 1u) * 7 * cap + path;
        acc[0] = acc[0] + C.x;
        acc[cap] = acc[cap] + C.y;
        acc[2 * cap] = acc[2 * cap] + C.z;
        if (p.learnLight)
        {
            // luminance in 2^-16 units, capped at 2^24 per sample; integer
            // sums do not depend on the order of the additions
            const float y = 0.2126f * C.x + 0.7152f * C.y + 0.0722f * C.z;
            const float f = fminf(fmaxf(y, 0.0f) * 65536.0f, 1.6777216e12f);
            lum[ti == SPG_MISS ? 0 : 1] += (unsigned long long)f;
        }
    }
    if (p.learnLight)  // uniform: every thread of the block reaches this
    {
        __shared__ unsigned long long sLum[2][4];
        for (int k = 0; k < 2; ++k)
        {
            unsigned long long v = lum[k];
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
            if ((threadIdx.x & 63) == 0) sLum[k][threadIdx.x >> 6] = v;
        }
        __syncthreads();
        if (threadIdx.x < 2)
        {
            const unsigned long long v = sLum[threadIdx.x][0] + sLum[threadIdx.x][1] +
                                         sLum[threadIdx.x][2] + sLum[threadIdx.x][3];
            if (v) atomicAdd(&p.paths.metrics[8 + threadIdx.x], v);
        }
    }
}

// The guided estimator's reverse recurrence of one path (cos/pdf in place
// of the cosine, no clamp): returns its radiance and, when the path is
// selected for training, calls sample(j, w) for each scattering vertex j
// (in reverse order) whose training weight w = Y(f cos Li / pdf) = Y(T_j)
// is positive.
template <class F>
__device__ __forceinline__ v3 guided_recurrence(const TraceParams &p, uint32_t path,
                                                F sample)
{
    const uint32_t cap = p.paths.capacity;
    const int32_t vc = (int32_t)p.paths.vcount[path];
    const bool sel = p.guide.train && guide_hash(path ^ p.guide.salt) < p.guide.selectBelow;
    v3 radiance{0.0f, 0.0f, 0.0f};
    for (int32_t j = vc - 1; j >= 0; --j)
    {
        const float *vt = p.paths.vtx + (size_t)j * 7 * cap + path;
        v3 E{vt[0], vt[cap], vt[2 * cap]};
        v3 X{vt[3 * cap], vt[4 * cap], vt[5 * cap]};
        float c = vt[6 * cap];
        v3 T = had(X, radiance) * c;
        if (sel && c > 0.0f)
        {
            float w = 0.2126f * T.x + 0.7152f * T.y + 0.0722f * T.z;
            if (w > 0.0f) sample(j, w);
        }
        radiance = E + T;
    }
    return radiance;
}

// Per pixel slot: the radiance of each of its paths in sample order
// (parity: ComputeRadianceForPath's reverse recurrence with the clamp on the
// incoming radiance, src/simd_path_tracer.cpp:113-172), then
// total += radiance * (1/spp) (src/simd_path_tracer.cpp:323).
__global__ __launch_bounds__(256) void k_finalize(TraceParams p)
{
    const uint32_t cap = p.paths.capacity;
    const float invSpp = 1.0f / (float)p.spp;
    for (uint32_t slot = xcd_block() * 256 + threadIdx.x; slot < p.slots;
         slot += gridDim.x * 256)
    {
        v3 total{p.accum[slot], p.accum[p.slots + slot],
                 p.accum[2 * p.slots + slot]};
        uint32_t lastXs = 0;
        for (uint32_t k = 0; k < p.sppPass; ++k)
        {
            uint32_t path = slot * p.sppPass + k;
            v3 radiance{0.0f, 0.0f, 0.0f};
            if (p.estimator == SPG_EST_PARITY)
            {
                int32_t vc = (int32_t)p.paths.vcount[path];
                for (int32_t i = vc - 1; i >= 0; --i)
                {
                    const float *vt = p.paths.vtx + (size_t)i * 7 * cap + path;
                    v3 E{vt[0], vt[cap], vt[2 * cap]};
                    v3 X{vt[3 * cap], vt[4 * cap], vt[5 * cap]};
                    float cosine = vt[6 * cap];
                    v3 incoming = radiance;
                    if (p.clampValue > 0.0f)
                        incoming = vclamp(incoming, 0.0f, p.clampValue);
                    radiance = E + had(X, incoming) * cosine;
                }
            }
            else if (p.estimator == SPG_EST_UNBIASED || p.estimator == SPG_EST_NEE)
                radiance = v3{p.paths.radiance[path], p.paths.radiance[cap + path],
                              p.paths.radiance[2 * cap + path]};
            else
            {
                // guided: the same recurrence with cos/pdf in place of the
                // cosine and no clamp; a selected path's training samples
                // are counted here and written by k_guide_emit
                uint32_t emitted = 0;
                radiance = guided_recurrence(p, path, [&](int32_t, float) { emitted++; });
                if (p.guide.train)
                    p.guide.gcnt[(uint32_t)((uint64_t)path * p.guide.permInv %
                                            ((uint64_t)p.slots * p.sppPass))] = emitted;
            }
            total = total + radiance * invSpp;
            lastXs = p.paths.rngXs[path];
        }
        p.accum[slot] = total.x;
        p.accum[p.slots + slot] = total.y;
        p.accum[2 * p.slots + slot] = total.z;
        if (p.rngMode != SPG_RNG_PCG) p.paths.slotRng[slot] = lastXs;
    }
}

// Training samples of a pass in a fixed order.  gcnt[v] holds the count of
// the path p = v * permMul mod paths (k_finalize); the samples are stored
// in v order: k_gscan_sums adds each 4096-entry segment, k_gscan_top turns
// the sums into segment bases (and the total into the sample count),
// k_guide_emit scans its segment and writes each path's samples at its
// offset (those past sampleCap are dropped).  Consecutive v are spread over
// the image, so every training batch -- a contiguous range -- covers all
// of it, and the result does not depend on scheduling.
__device__ __forceinline__ uint32_t block_exclusive_scan_1024(uint32_t x, uint32_t *sh,
                                                              uint32_t &total)
{
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t inc = x;
    for (int o = 1; o < 64; o <<= 1)
    {
        uint32_t y = __shfl_up(inc, o);
        if ((int)lane >= o) inc += y;
    }
    if (lane == 63) sh[wave] = inc;
    __syncthreads();
    if (wave == 0)
    {
        uint32_t w = lane < 16 ? sh[lane] : 0u, wi = w;
        for (int o = 1; o < 16; o <<= 1)
        {
            uint32_t y = __shfl_up(wi, o);
            if ((int)lane >= o) wi += y;
        }
        if (lane < 16) sh[16 + lane] = wi - w;
        if (lane == 15) sh[32] = wi;
    }
    __syncthreads();
    total = sh[32];
    return sh[16 + wave] + inc - x;
}

__global__ __launch_bounds__(1024) void k_gscan_sums(TraceParams p)
{
    __shared__ uint32_t sh[40];
    const uint32_t N = p.slots * p.sppPass;
    const uint32_t v0 = blockIdx.x * kGScanSeg + threadIdx.x * 4;
    uint32_t c = 0;
    for (int k = 0; k < 4; ++k)
        if (v0 + k < N) c += p.guide.gcnt[v0 + k];
    uint32_t total;
    block_exclusive_scan_1024(c, sh, total);
    if (threadIdx.x == 0) p.guide.scanBase[blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void k_gscan_top(TraceParams p, uint32_t segs)
{
    __shared__ uint32_t sh[40];
    const uint32_t x = threadIdx.x < segs ? p.guide.scanBase[threadIdx.x] : 0u;
    uint32_t total;
    const uint32_t e = block_exclusive_scan_1024(x, sh, total);
    if (threadIdx.x < segs) p.guide.scanBase[threadIdx.x] = e;
    if (threadIdx.x == 0) *p.guide.sampleCount = total;
}

__global__ __launch_bounds__(1024) void k_guide_emit(TraceParams p)
{
    __shared__ uint32_t sh[40];
    const uint64_t N = (uint64_t)p.slots * p.sppPass;
    const uint32_t v0 = blockIdx.x * kGScanSeg + threadIdx.x * 4;
    uint32_t c[4], sum = 0;
    for (int k = 0; k < 4; ++k)
    {
        c[k] = v0 + k < N ? p.guide.gcnt[v0 + k] : 0u;
        sum += c[k];
    }
    uint32_t total;
    uint32_t at = p.guide.scanBase[blockIdx.x] + block_exclusive_scan_1024(sum, sh, total);
    const uint32_t cap = p.paths.capacity;
    const size_t stride = (size_t)p.bounces * cap;
    const size_t C = p.guide.sampleCap;
    for (int k = 0; k < 4; ++k)
    {
        if (!c[k]) continue;
        const uint32_t path = (uint32_t)((uint64_t)(v0 + k) * p.guide.permMul % N);
        guided_recurrence(p, path, [&](int32_t j, float w) {
            if (at < C)
            {
                const float *gv = p.guide.gvtx + (size_t)j * cap + path;
                float *dst = p.guide.samples + at;
                for (int f = 0; f < kGVertexFloats; ++f) dst[f * C] = gv[f * stride];
                dst[12 * C] = w;
            }
            ++at;
        });
    }
}

// Guide inference for every queued ray that will scatter: hit surface ->
// encoding -> MLP (FP16 MFMA, FP32 accumulation) -> vMF mixture, stored
// per queue entry for k_shade.  One block of kInferWaves waves per CU
// shares the weights and biases in LDS (41,344 + 1,216 bytes); each wave
// evaluates 64 rays per step through 4,608 bytes of scratch, so the 16
// waves (four per SIMD) fit in one CU's LDS.
constexpr int kInferWaves = 16;
template <bool kBatch>
__global__ __launch_bounds__(64 * kInferWaves) void k_guide_infer(TraceParams p,
                                                                  int q)
{
    __shared__ _Float16 sW[kWLdsHalves];
    __shared__ float sBias[kGRows];
    __shared__ float waveOut[kInferWaves][kTileScratchFloats];
    stage_rows(sW, p.guide.w16, kGRows);
    for (int k = threadIdx.x; k < kGRows; k += blockDim.x) sBias[k] = p.guide.bias[k];
    __syncthreads();
    const uint32_t n = p.paths.counters[q];
    const uint32_t cap = p.paths.capacity;
    const float *qo = p.paths.qo[q], *qd = p.paths.qd[q];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    float *scr = waveOut[wave];
    unsigned long long evals = 0;
    for (uint32_t tile = xcd_block() * kInferWaves + wave; tile * 64 < n;
         tile += gridDim.x * kInferWaves)
    {
        const uint32_t i = tile * 64 + lane;
        bool valid = false;
        v3 pn{0, 0, 0}, N{0, 0, 0}, V{0, 0, 0};
        if (i < n)
        {
            uint32_t inst = p.paths.hinst[i];
            float t = p.paths.ht[i];
            uint32_t path = p.paths.qpath[q][i];
            if (inst != SPG_MISS && t > 0.0f && p.paths.vcount[path] + 1 < p.bounces)
            {
                valid = true;
                v3 o{qo[i], qo[cap + i], qo[2 * cap + i]};
                v3 d{qd[i], qd[cap + i], qd[2 * cap + i]};
                v2 uv;
                const DevScene &S =
                    kBatch ? p.scenes[p.slotItem[p.paths.pathSlot[path]]] : p.scene;
                surface_at(S, inst, p.paths.hprim[i], p.paths.hu[i], p.paths.hv[i], N, uv);
                pn = guide_pn(p.guide, o + d * t);
                V = -d;
                if (p.guide.analytic)
                {
                    // verification mode: the analytic mixture, no network
                    Lobes L;
                    guide_analytic_lobes(o + d * t, N, V, L);
                    store_lobes(p.guide.lobes, cap, i, L);
                }
            }
        }
        const unsigned long long vmask = __ballot(valid);  // wave-uniform
        evals += (unsigned long long)__popcll(vmask);
        if (p.guide.analytic || vmask == 0ull) continue;  // wave-uniform
        Lobes L;
        guide_forward_tile(sW, sBias, scr, valid, pn, N, V, L, [](int, const float *) {});
        if (valid) store_lobes(p.guide.lobes, cap, i, L);
    }
    // one counter update per wave (every lane reaches this point)
    if (lane == 0 && evals) atomicAdd(&p.paths.metrics[7], evals);
}

template <bool kBatch>
__global__ void k_write_pixels(TraceParams p, float *image)
{
    for (uint32_t slot = blockIdx.x * 256 + threadIdx.x; slot < p.slots;
         slot += gridDim.x * 256)
    {
        uint32_t pixel = p.slotPixel[slot];
        float4 c = make_float4(p.accum[slot], p.accum[p.slots + slot],
                               p.accum[2 * p.slots + slot], 1.0f);
        float *dst = kBatch ? p.images[p.slotItem[slot]] : image;
        reinterpret_cast<float4 *>(dst)[pixel] = c;
    }
}

// Host-pixel calls: the chunk's pixels in slot order (tile order), so one
// contiguous copy brings every tile to the host.
__global__ void k_write_slots(TraceParams p, float *out)
{
    for (uint32_t slot = blockIdx.x * 256 + threadIdx.x; slot < p.slots;
         slot += gridDim.x * 256)
        reinterpret_cast<float4 *>(out)[slot] =
            make_float4(p.accum[slot], p.accum[p.slots + slot],
                        p.accum[2 * p.slots + slot], 1.0f);
}

__global__ void k_clear_u32(uint32_t *p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

__global__ void k_clear_f32(float *p, size_t n)
{
    for (size_t i = blockIdx.x * 256 + threadIdx.x; i < n;
         i += (size_t)gridDim.x * 256)
        p[i] = 0.0f;
}

// ------------------------------------------------------------ launchers
static uint32_t grid_for(uint64_t n, uint32_t block, uint32_t cap)
{
    uint64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (uint32_t)g;
}

static uint32_t device_cus()
{
    static uint32_t cus = 0;
    if (!cus)
    {
        int dev = 0, n = 0;
        hipGetDevice(&dev);
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev);
        cus = n > 0 ? (uint32_t)n : 1u;
    }
    return cus;
}

// Waves a persistent launch keeps resident: CUs x resident one-wave blocks
// per CU, capped at kMaxPersistWavesPerCU per CU (the spill buffer size).
template <class K> static uint32_t persistent_grid(K kernel)
{
    int blocks = 0;
    hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kernel, kTraceBlock, 0);
    if (blocks < 1) blocks = 1;
    if (blocks > (int)kMaxPersistWavesPerCU) blocks = kMaxPersistWavesPerCU;
    return device_cus() * (uint32_t)blocks;
}

size_t spill_ints()
{
    return (size_t)device_cus() * kMaxPersistWavesPerCU * kTraceBlock *
           (size_t)(kStackDepth - kMinShortStack);
}

static uint32_t persist_blocks(uint32_t grid, uint64_t n, uint32_t chunk = kSceneChunk)
{
    uint64_t need = (n + chunk - 1) / chunk;
    if (need < 1) need = 1;
    return need < grid ? (uint32_t)need : grid;
}

__global__ void k_set_u32(uint32_t *p, uint32_t v) { *p = v; }

hipError_t launch_closest_mesh(hipStream_t s, const DevMesh *mesh,
                               const float *ox, const float *oy,
                               const float *oz, const float *dx,
                               const float *dy, const float *dz, uint32_t n,
                               uint32_t *prim, float *t, float *u, float *v,
                               unsigned long long *stats, const Persist &ps)
{
    if (n == 0) return hipSuccess;
    static uint32_t gCount = persistent_grid(k_closest_mesh<true>);
    static uint32_t gPlain = persistent_grid(k_closest_mesh<false>);
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, ps.counter, 0u);
    if (stats)
        hipLaunchKernelGGL(k_closest_mesh<true>, dim3(persist_blocks(gCount, n, kMeshChunk)),
                           dim3(kTraceBlock), 0, s, *mesh, ox, oy, oz, dx, dy, dz,
                           n, prim, t, u, v, stats, ps.counter, ps.spill);
    else
        hipLaunchKernelGGL(k_closest_mesh<false>, dim3(persist_blocks(gPlain, n, kMeshChunk)),
                           dim3(kTraceBlock), 0, s, *mesh, ox, oy, oz, dx, dy, dz,
                           n, prim, t, u, v, stats, ps.counter, ps.spill);
    return hipGetLastError();
}

// ---- allocation-free closest hit (spg_intersect_closest_device_noalloc)
// Its scratch lives in this file-scope array, one copy per GPU for the
// life of the process: the persistent launch's work counter (first 256
// bytes), then the traversal-stack spill of at most kNoAllocWaves one-wave
// blocks, kNoAllocSpill entries per lane beyond the LDS stack (entry i of
// grid lane gl at spill[(i - kMeshShortStack) * gridLanes + gl]).  Host code
// never uses the array's name as an address: noalloc_scratch_address
// returns the device address (hipGetSymbolAddress), which the launcher
// passes to the kernels.
constexpr uint32_t kNoAllocSpill = kNoAllocStack - kMeshShortStack;
static_assert(kNoAllocStack > (uint32_t)kMeshShortStack, "spill entries");
constexpr size_t kNoAllocScratchBytes =
    256 + (size_t)kNoAllocWaves * kTraceBlock * kNoAllocSpill * sizeof(int);
static_assert(kNoAllocScratchBytes == SPG_NOALLOC_SCRATCH_BYTES &&
                  kNoAllocStack == SPG_NOALLOC_STACK,
              "the header's limits");
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

hipError_t noalloc_scratch_address(void **ptr)
{
    return hipGetSymbolAddress(ptr, HIP_SYMBOL(scratch));
}

hipError_t launch_closest_mesh_noalloc(hipStream_t s, const DevMesh *mesh,
                                       const float *ox, const float *oy,
                                       const float *oz, const float *dx,
                                       const float *dy, const float *dz, uint32_t n,
                                       uint32_t *prim, float *t, float *u, float *v,
                                       void *scratchPtr)
{
    if (n == 0) return hipSuccess;
    static uint32_t g = std::min(persistent_grid(k_closest_mesh<false>), kNoAllocWaves);
    uint32_t *counter = (uint32_t *)scratchPtr;
    int *spill = (int *)((char *)scratchPtr + 256);
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, counter, 0u);
    hipLaunchKernelGGL(k_closest_mesh<false>, dim3(persist_blocks(g, n, kMeshChunk)),
                       dim3(kTraceBlock), 0, s, *mesh, ox, oy, oz, dx, dy, dz, n, prim,
                       t, u, v, nullptr, counter, spill);
    return hipGetLastError();
}

hipError_t launch_intersect_scene(hipStream_t s, const DevScene &scene,
                                  const float *rays6, uint32_t n, float *out7,
                                  uint32_t *object, uint32_t *prim,
                                  const Persist &ps)
{
    if (n == 0) return hipSuccess;
    static uint32_t g = persistent_grid(k_intersect_scene);
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, ps.counter, 0u);
    hipLaunchKernelGGL(k_intersect_scene, dim3(persist_blocks(g, n)),
                       dim3(kTraceBlock), 0, s, scene, rays6, n, out7, object,
                       prim, ps.counter, ps.spill);
    return hipGetLastError();
}

hipError_t launch_raygen(hipStream_t s, const TraceParams &p)
{
    uint64_t n = (uint64_t)p.slots * p.sppPass;
    if (p.slotItem)
        hipLaunchKernelGGL(k_raygen<true>, dim3(grid_for(n, 256, 8192)), dim3(256), 0,
                           s, p);
    else
        hipLaunchKernelGGL(k_raygen<false>, dim3(grid_for(n, 256, 8192)), dim3(256), 0,
                           s, p);
    return hipGetLastError();
}

hipError_t launch_extend(hipStream_t s, const TraceParams &p, int q,
                         const Persist &ps)
{
    static uint32_t gCount = persistent_grid(k_extend<true, false, false>);
    static uint32_t gPlain = persistent_grid(k_extend<false, false, false>);
    static uint32_t gCountB = persistent_grid(k_extend<true, true, false>);
    static uint32_t gPlainB = persistent_grid(k_extend<false, true, false>);
    static uint32_t gCountS = persistent_grid(k_extend<true, false, true>);
    static uint32_t gPlainS = persistent_grid(k_extend<false, false, true>);
    static uint32_t gCountBS = persistent_grid(k_extend<true, true, true>);
    static uint32_t gPlainBS = persistent_grid(k_extend<false, true, true>);
    uint64_t n = (uint64_t)p.slots * p.sppPass;  // bound on the queue length
    const bool batch = p.slotItem != nullptr;
    const bool shadow = q == 2;
    uint32_t g = batch ? (shadow ? (p.collect ? gCountBS : gPlainBS)
                                 : (p.collect ? gCountB : gPlainB))
                 : shadow ? (p.collect ? gCountS : gPlainS)
                          : (p.collect ? gCount : gPlain);
    if (ps.maxBlocks && g > ps.maxBlocks) g = ps.maxBlocks;
    dim3 grid(persist_blocks(g, n));
    // ps.counter was zeroed by the kernel before (TraceParams::work)
#define SPG_EXT(C, B, S)                                                          \
    hipLaunchKernelGGL((k_extend<C, B, S>), grid, dim3(kTraceBlock), 0, s, p, q, \
                       ps.counter, ps.spill)
    if (batch && shadow)
    {
        if (p.collect) SPG_EXT(true, true, true);
        else SPG_EXT(false, true, true);
    }
    else if (batch)
    {
        if (p.collect) SPG_EXT(true, true, false);
        else SPG_EXT(false, true, false);
    }
    else if (shadow)
    {
        if (p.collect) SPG_EXT(true, false, true);
        else SPG_EXT(false, false, true);
    }
    else if (p.collect)
        SPG_EXT(true, false, false);
    else
        SPG_EXT(false, false, false);
#undef SPG_EXT
    return hipGetLastError();
}

hipError_t launch_shade(hipStream_t s, const TraceParams &p, int q)
{
    uint32_t g = grid_for((uint64_t)p.slots * p.sppPass, 256, 4096);
    g = (g + 7) / 8 * 8;
    const bool dbl = p.tx == SPG_TX_DOUBLE;
#define SPG_SHADE(D, E) \
    hipLaunchKernelGGL((k_shade<D, E, false>), dim3(g), dim3(256), 0, s, p, q)
#define SPG_SHADE_B(D, E) \
    hipLaunchKernelGGL((k_shade<D, E, true>), dim3(g), dim3(256), 0, s, p, q)
    if (p.slotItem)
    {
        switch (p.estimator)
        {
        case SPG_EST_PARITY:
            if (dbl) SPG_SHADE_B(true, SPG_EST_PARITY);
            else SPG_SHADE_B(false, SPG_EST_PARITY);
            break;
        case SPG_EST_UNBIASED:
            if (dbl) SPG_SHADE_B(true, SPG_EST_UNBIASED);
            else SPG_SHADE_B(false, SPG_EST_UNBIASED);
            break;
        case SPG_EST_NEE:
            if (dbl) SPG_SHADE_B(true, SPG_EST_NEE);
            else SPG_SHADE_B(false, SPG_EST_NEE);
            break;
        case SPG_EST_GUIDED_NEE:
            if (dbl) SPG_SHADE_B(true, SPG_EST_GUIDED_NEE);
            else SPG_SHADE_B(false, SPG_EST_GUIDED_NEE);
            break;
        default:
            if (dbl) SPG_SHADE_B(true, SPG_EST_GUIDED);
            else SPG_SHADE_B(false, SPG_EST_GUIDED);
            break;
        }
        return hipGetLastError();
    }
    switch (p.estimator)
    {
    case SPG_EST_PARITY:
        if (dbl) SPG_SHADE(true, SPG_EST_PARITY);
        else SPG_SHADE(false, SPG_EST_PARITY);
        break;
    case SPG_EST_UNBIASED:
        if (dbl) SPG_SHADE(true, SPG_EST_UNBIASED);
        else SPG_SHADE(false, SPG_EST_UNBIASED);
        break;
    case SPG_EST_NEE:
        if (dbl) SPG_SHADE(true, SPG_EST_NEE);
        else SPG_SHADE(false, SPG_EST_NEE);
        break;
    case SPG_EST_GUIDED_NEE:
        if (dbl) SPG_SHADE(true, SPG_EST_GUIDED_NEE);
        else SPG_SHADE(false, SPG_EST_GUIDED_NEE);
        break;
    default:
        if (dbl) SPG_SHADE(true, SPG_EST_GUIDED);
        else SPG_SHADE(false, SPG_EST_GUIDED);
        break;
    }
#undef SPG_SHADE
#undef SPG_SHADE_B
    return hipGetLastError();
}

hipError_t launch_resolve_shadow(hipStream_t s, const TraceParams &p)
{
    // at most 1024 blocks: with the learned light split each block adds its
    // sums with one atomic per light kind
    uint32_t g = grid_for((uint64_t)p.slots * p.sppPass, 256, 1024);
    hipLaunchKernelGGL(k_resolve_shadow, dim3(g), dim3(256), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_guide_infer(hipStream_t s, const TraceParams &p, int q)
{
    uint64_t n = (uint64_t)p.slots * p.sppPass;  // bound on the queue length
    uint64_t blocks = (n + 64 * kInferWaves - 1) / (64 * kInferWaves);
    uint32_t cap = device_cus();
    if (blocks > cap) blocks = cap;
    if (blocks < 1) blocks = 1;
    if (p.slotItem)
        hipLaunchKernelGGL(k_guide_infer<true>, dim3((uint32_t)blocks),
                           dim3(64 * kInferWaves), 0, s, p, q);
    else
        hipLaunchKernelGGL(k_guide_infer<false>, dim3((uint32_t)blocks),
                           dim3(64 * kInferWaves), 0, s, p, q);
    return hipGetLastError();
}

hipError_t launch_finalize(hipStream_t s, const TraceParams &p)
{
    hipLaunchKernelGGL(k_finalize, dim3(grid_for(p.slots, 256, 4096)), dim3(256),
                       0, s, p);
    return hipGetLastError();
}

hipError_t launch_guide_collect(hipStream_t s, const TraceParams &p)
{
    const uint64_t n = (uint64_t)p.slots * p.sppPass;
    const uint32_t segs = (uint32_t)((n + kGScanSeg - 1) / kGScanSeg);
    if (segs == 0) return hipSuccess;
    if (segs > (uint32_t)kGScanBlocks) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_gscan_sums, dim3(segs), dim3(1024), 0, s, p);
    hipLaunchKernelGGL(k_gscan_top, dim3(1), dim3(1024), 0, s, p, segs);
    hipLaunchKernelGGL(k_guide_emit, dim3(segs), dim3(1024), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_write_pixels(hipStream_t s, const TraceParams &p,
                               float *image)
{
    if (p.slotItem)
        hipLaunchKernelGGL(k_write_pixels<true>, dim3(grid_for(p.slots, 256, 4096)),
                           dim3(256), 0, s, p, image);
    else
        hipLaunchKernelGGL(k_write_pixels<false>, dim3(grid_for(p.slots, 256, 4096)),
                           dim3(256), 0, s, p, image);
    return hipGetLastError();
}

hipError_t launch_write_slots(hipStream_t s, const TraceParams &p, float *out)
{
    hipLaunchKernelGGL(k_write_slots, dim3(grid_for(p.slots, 256, 4096)), dim3(256), 0, s,
                       p, out);
    return hipGetLastError();
}

hipError_t launch_clear_counters(hipStream_t s, uint32_t *counters, int n)
{
    hipLaunchKernelGGL(k_clear_u32, dim3((n + 255) / 256), dim3(256), 0, s,
                       counters, n);
    return hipGetLastError();
}

hipError_t launch_set_counter(hipStream_t s, uint32_t *counter, uint32_t v)
{
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, counter, v);
    return hipGetLastError();
}

hipError_t launch_clear_floats(hipStream_t s, float *p, size_t n)
{
    hipLaunchKernelGGL(k_clear_f32, dim3(grid_for(n, 256, 4096)), dim3(256), 0,
                       s, p, n);
    return hipGetLastError();
}

}  // namespace spg
