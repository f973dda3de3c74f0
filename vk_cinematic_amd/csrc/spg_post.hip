// gfx950 kernels for the step after the path: progressive accumulation of
// images across trace calls and the viewer's tone mapping.
#include "spg_internal.h"
#include "spg_tx.h"

namespace spg {

// dst = (dst * before + src * added) / (before + added), RGB; alpha 1.
__global__ __launch_bounds__(256) void k_accumulate(float4 *__restrict__ dst,
                                                    const float4 *__restrict__ src,
                                                    uint32_t pixels, float before,
                                                    float added, float total)
{
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < pixels;
         i += gridDim.x * 256)
    {
        float4 a = dst[i], b = src[i];
        a.x = (a.x * before + b.x * added) / total;
        a.y = (a.y * before + b.y * added) / total;
        a.z = (a.z * before + b.z * added) / total;
        a.w = 1.0f;
        dst[i] = a;
    }
}

// PerformToneMapping (src/shaders/post_processing.frag.glsl:19-26):
// c *= exposure; c = c / (1 + c); c = pow(c, 1/2.2); then the UNORM8 write
// of the swapchain: round to nearest of 255 c.  Alpha 255.
template <bool kDouble>
__global__ __launch_bounds__(256) void k_tonemap(const float4 *__restrict__ src,
                                                 uint32_t pixels, float exposure,
                                                 uchar4 *__restrict__ out)
{
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < pixels;
         i += gridDim.x * 256)
    {
        const float4 c = src[i];
        const float ch[3] = {c.x, c.y, c.z};
        uint8_t q[3];
        for (int k = 0; k < 3; ++k)
        {
            float x = ch[k] * exposure;
            x = x / (1.0f + x);
            float g = kDouble ? tpow(x, 1.0f / 2.2f) : __powf(x, 1.0f / 2.2f);
            g = clamp_sel(g, 0.0f, 1.0f);  // NaN and negative inputs give 0
            q[k] = (uint8_t)(uint32_t)rintf(g * 255.0f);
        }
        out[i] = make_uchar4(q[0], q[1], q[2], 255);
    }
}

hipError_t launch_accumulate(hipStream_t s, float4 *dst, const float4 *src,
                             uint32_t pixels, uint32_t before, uint32_t added)
{
    if (pixels == 0) return hipSuccess;
    const uint32_t blocks = (pixels + 255) / 256 < 65535u ? (pixels + 255) / 256 : 65535u;
    k_accumulate<<<blocks, 256, 0, s>>>(dst, src, pixels, (float)before, (float)added,
                                        (float)before + (float)added);
    return hipGetLastError();
}

hipError_t launch_tonemap(hipStream_t s, const float4 *src, uint32_t pixels,
                          float exposure, bool dbl, uchar4 *out)
{
    if (pixels == 0) return hipSuccess;
    const uint32_t blocks = (pixels + 255) / 256 < 65535u ? (pixels + 255) / 256 : 65535u;
    if (dbl) k_tonemap<true><<<blocks, 256, 0, s>>>(src, pixels, exposure, out);
    else k_tonemap<false><<<blocks, 256, 0, s>>>(src, pixels, exposure, out);
    return hipGetLastError();
}

}  // namespace spg
