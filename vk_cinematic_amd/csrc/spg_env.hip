// gfx950 kernels for environment maps.
//
//   irradiance cube map   k_irradiance   CreateIrradianceCubeMap
//                         (src/cubemap.cpp:108-233, uniform-sampling build)
//
// The reference sums about a thousand samples per texel in one loop.  Here
// one wave owns a texel: its 64 lanes evaluate 64 consecutive samples of
// that loop at once (direction, bilinear lookup, clamp, L cos sin), and
// three lanes then add the x, y and z contributions in loop order, so every
// texel's sum is the reference's sequential FP32 sum bit for bit.
#include "spg_internal.h"
#include "spg_tx.h"

namespace spg {

constexpr int kIrrWaves = 4;

// SampleImageBilinear, src/image.h:34-74 (ClampToEdge; Max is compare-and-
// select, so a NaN coordinate reads texel 0).
__device__ __forceinline__ v3 sample_bilinear3(const float4 *__restrict__ img,
                                               uint32_t w, uint32_t h, float u,
                                               float v)
{
    float px = (u * (float)w) - 0.5f;
    float py = (v * (float)h) - 0.5f;
    px = fmax_sel(px, 0.0f);
    py = fmax_sel(py, 0.0f);
    uint32_t x0 = (uint32_t)floorf(px), x1 = x0 + 1;
    uint32_t y0 = (uint32_t)floorf(py), y1 = y0 + 1;
    float fx = px - (float)x0;
    float fy = py - (float)y0;
    x1 = x1 < w - 1 ? x1 : w - 1;
    y1 = y1 < h - 1 ? y1 : h - 1;
    // u, v <= 1 keep x0 < w and y0 < h; clamp anyway so no input reads
    // outside the image
    x0 = x0 < w - 1 ? x0 : w - 1;
    y0 = y0 < h - 1 ? y0 : h - 1;
    float4 s0 = img[(size_t)y0 * w + x0], s1 = img[(size_t)y0 * w + x1];
    float4 s2 = img[(size_t)y1 * w + x0], s3 = img[(size_t)y1 * w + x1];
    // Lerp(vec4), src/math_lib.h:432-436: a (1 - t) + b t
    auto lerp = [](float a, float b, float t) { return a * (1.0f - t) + b * t; };
    v3 t0{lerp(s0.x, s1.x, fx), lerp(s0.y, s1.y, fx), lerp(s0.z, s1.z, fx)};
    v3 t1{lerp(s2.x, s3.x, fx), lerp(s2.y, s3.y, fx), lerp(s2.z, s3.z, fx)};
    return v3{lerp(t0.x, t1.x, fy), lerp(t0.y, t1.y, fy), lerp(t0.z, t1.z, fy)};
}

// MapCubeMapLayerIndexToBasisVectors, src/cubemap.cpp:55-103
__device__ __forceinline__ void cube_basis(uint32_t face, v3 &fwd, v3 &up, v3 &right)
{
    const float B[6][9] = {
        {1, 0, 0, 0, 1, 0, 0, 0, -1}, {-1, 0, 0, 0, 1, 0, 0, 0, 1},
        {0, 1, 0, 0, 0, -1, 1, 0, 0}, {0, -1, 0, 0, 0, 1, 1, 0, 0},
        {0, 0, 1, 0, 1, 0, 1, 0, 0},  {0, 0, -1, 0, 1, 0, -1, 0, 0},
    };
    const float *b = B[face];
    fwd = v3{b[0], b[1], b[2]};
    up = v3{b[3], b[4], b[5]};
    right = v3{b[6], b[7], b[8]};
}

// One wave per texel of 6 faces of w x h.  phis / thetas hold the values
// the reference's float loop counters take (phi += 0.1f below 2 pi, theta
// += 0.1f below pi/2); sample k is (phis[k / nTheta], thetas[k % nTheta]).
template <bool kDouble>
__global__ __launch_bounds__(64 * kIrrWaves) void k_irradiance(
    const float4 *__restrict__ src, uint32_t ew, uint32_t eh, uint32_t w,
    uint32_t h, const float *__restrict__ phis, uint32_t nPhi,
    const float *__restrict__ thetas, uint32_t nTheta, float clampValue,
    float4 *__restrict__ out)
{
    __shared__ float part[kIrrWaves][3][64];
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t texel = blockIdx.x * kIrrWaves + wave;
    if (texel >= 6u * w * h) return;  // whole waves; no block barrier below
    const uint32_t face = texel / (w * h), rem = texel % (w * h);
    const uint32_t y = rem / w, x = rem % w;
    v3 fwd, up, right;
    cube_basis(face, fwd, up, right);
    float fx = (float)x / (float)w;
    float fy = (float)y / (float)h;
    fy = 1.0f - fy;
    fx = fx * 2.0f - 1.0f;
    fy = fy * 2.0f - 1.0f;
    const v3 n = normalize((fwd + right * fx) + up * fy);
    const v3 t = normalize(cross(up, n));
    const v3 b = normalize(cross(n, t));
    const uint32_t count = nPhi * nTheta;
    float acc = 0.0f;  // lanes 0, 1, 2: the x, y, z sums
    float *mine = part[wave][0];
    for (uint32_t base = 0; base < count; base += 64)
    {
        const uint32_t k = base + lane;
        v3 c{0.0f, 0.0f, 0.0f};
        if (k < count)
        {
            const float phi = phis[k / nTheta], theta = thetas[k % nTheta];
            // MapSphericalToCartesianCoordinates, src/math_lib.h:864-871
            const float st = Tx<kDouble>::sin_(theta), ct = Tx<kDouble>::cos_(theta);
            const v3 td{st * Tx<kDouble>::cos_(phi), ct, st * Tx<kDouble>::sin_(phi)};
            const v3 wd = (n * td.y + t * td.x) + b * td.z;
            // ToSphericalCoordinates + MapToEquirectangular,
            // src/math_lib.h:849-884, v flipped
            float inc = Tx<kDouble>::atan2_(sqrtf(wd.x * wd.x + wd.z * wd.z), wd.y);
            float az = Tx<kDouble>::atan2_(wd.z, wd.x);
            if (az < 0.0f) az += 2.0f * kPi;
            const float u = az / (2.0f * kPi);
            const float v = 1.0f - (Tx<kDouble>::cos_(inc) * 0.5f + 0.5f);
            v3 L = sample_bilinear3(src, ew, eh, u, v);
            if (clampValue > 0.0f) L = vclamp(L, 0.0f, clampValue);
            c = (L * ct) * st;
        }
        part[wave][0][lane] = c.x;
        part[wave][1][lane] = c.y;
        part[wave][2][lane] = c.z;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (lane < 3)
        {
            const uint32_t m = count - base < 64u ? count - base : 64u;
            const float *row = mine + lane * 64;
            for (uint32_t j = 0; j < m; ++j) acc += row[j];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    // irradiance = PI * irradiance * (1 / count)
    const float r = (kPi * acc) * (1.0f / (float)count);
    float *o = reinterpret_cast<float *>(out + texel);
    if (lane < 3) o[lane] = r;
    else if (lane == 3) o[3] = 1.0f;
}

hipError_t launch_irradiance(hipStream_t s, const float4 *src, uint32_t ew,
                             uint32_t eh, uint32_t w, uint32_t h, const float *phis,
                             uint32_t nPhi, const float *thetas, uint32_t nTheta,
                             float clampValue, bool dbl, float4 *out)
{
    const uint32_t texels = 6u * w * h;
    const uint32_t blocks = (texels + kIrrWaves - 1) / kIrrWaves;
    if (dbl)
        k_irradiance<true><<<blocks, 64 * kIrrWaves, 0, s>>>(
            src, ew, eh, w, h, phis, nPhi, thetas, nTheta, clampValue, out);
    else
        k_irradiance<false><<<blocks, 64 * kIrrWaves, 0, s>>>(
            src, ew, eh, w, h, phis, nPhi, thetas, nTheta, clampValue, out);
    return hipGetLastError();
}

}  // namespace spg
