// Wave-level FP16 MFMA building blocks of the guide's MLP (gfx950).
//
// One wave evaluates 64 rays (or training samples) at a time in the
// transposed form H' = W H (features x rays), with
// v_mfma_f32_16x16x32_f16 (FP16 operands, FP32 accumulation).  Lane
// layouts (lane = 16 g + r):
//   A (16 x 32):  lane holds A[m = r][k = 8 g + i], i = 0..7
//   B (32 x 16):  lane holds B[k = 8 g + i][n = r]
//   D (16 x 16):  lane holds D[m = 4 g + i][n = r], i = 0..3
// Output tiles of one layer feed the next layer's B operand without leaving
// registers: the K slots (c, g, i) of K-step c are mapped to features
//   f(c, g, i) = 32 c + 16 (i / 4) + 4 g + (i % 4),
// so B(rb, c) = (D(rb, 2c), D(rb, 2c+1)) converted to FP16, and the weight
// fragments are read with the same feature order (two 8-byte LDS reads).
#pragma once

#include "spg_guide.h"

namespace spg {

typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

// LDS row pitch in halves for weight rows and per-wave encodings: 136 bytes
// keep the 16 rows of a fragment read on distinct banks.
constexpr int kLdsPitch = 68;
constexpr int kWLdsHalves = kGRows * kLdsPitch;          // forward weights
constexpr int kWtLdsHalves = kGLayers * 64 * kLdsPitch;  // transposed weights
// Per-wave scratch: FP32 outputs [48][64] (aliases the FP16 encodings
// [64][kLdsPitch]).
constexpr int kWaveScratchFloats = kGOut * 64;

__device__ __forceinline__ f4 mfma16x16x32(h8 a, h8 b, f4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// LDS exchange between lanes of one wave: every LDS access issued before
// this point has completed (s_waitcnt lgkmcnt(0); outstanding global loads
// and stores are not waited for) and the compiler may not move memory
// accesses across it.
constexpr int kWaitLgkm0 = 0xC07F;  // vmcnt 63, expcnt 7, lgkmcnt 0
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_s_waitcnt(kWaitLgkm0);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Fragment in the permuted feature order: 4 halves at col, 4 at col + 16.
__device__ __forceinline__ h8 lds_frag(const _Float16 *row, int col)
{
    h4 lo = *(const h4 *)(row + col);
    h4 hi = *(const h4 *)(row + col + 16);
    return __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
}

__device__ __forceinline__ h8 pack_tiles(f4 a, f4 b)
{
    h8 v;
    v[0] = (_Float16)a[0];
    v[1] = (_Float16)a[1];
    v[2] = (_Float16)a[2];
    v[3] = (_Float16)a[3];
    v[4] = (_Float16)b[0];
    v[5] = (_Float16)b[1];
    v[6] = (_Float16)b[2];
    v[7] = (_Float16)b[3];
    return v;
}

// Copy fp16 rows of 64 (global) into LDS rows of kLdsPitch, whole block.
__device__ __forceinline__ void stage_rows(_Float16 *dst, const _Float16 *src,
                                           int rows)
{
    for (int k = threadIdx.x; k < rows * 16; k += blockDim.x)
    {
        int r = k >> 4, c = (k & 15) * 4;
        *(h4 *)(dst + r * kLdsPitch + c) = *(const h4 *)(src + r * 64 + c);
    }
}

// sin and cos of 2 pi x with the hardware's v_sin_f32 / v_cos_f32 (x in
// revolutions, |x| <= 256; absolute error near 2^-20, far below the FP16
// the features are stored in).
__device__ __forceinline__ void sincos_rev(float x, float *s, float *c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    *s = __builtin_amdgcn_sinf(x);
    *c = __builtin_amdgcn_cosf(x);
#else
    *s = sinf(6.28318531f * x);
    *c = cosf(6.28318531f * x);
#endif
}

// Encoding of one hit into 64 features (see spg_guide.h), written as one
// LDS row of the wave's encoding tile.  Invalid rays get zeros.  sin/cos(2^l
// pi p) for p in [0, 1] is 2^(l-1) p revolutions (at most 64).
__device__ __forceinline__ void encode_row(_Float16 *row, bool valid, v3 pn, v3 n,
                                           v3 v)
{
    float e[kGIn];
    float p[3] = {pn.x, pn.y, pn.z};
    for (int l = 0; l < 8; ++l)
        for (int a = 0; a < 3; ++a)
        {
            float s, c;
            sincos_rev(ldexpf(p[a], l - 1), &s, &c);
            e[6 * l + 2 * a] = s;
            e[6 * l + 2 * a + 1] = c;
        }
    e[48] = pn.x;
    e[49] = pn.y;
    e[50] = pn.z;
    e[51] = n.x;
    e[52] = n.y;
    e[53] = n.z;
    e[54] = v.x;
    e[55] = v.y;
    e[56] = v.z;
    e[57] = 1.0f;
    for (int k = 58; k < kGIn; ++k) e[k] = 0.0f;
    for (int k = 0; k < kGIn; k += 4)
    {
        h4 q;
        for (int j = 0; j < 4; ++j) q[j] = valid ? (_Float16)e[k + j] : (_Float16)0.0f;
        *(h4 *)(row + k) = q;
    }
}

// B operands of the input layer from the wave's encoding tile.
__device__ __forceinline__ void load_input(const _Float16 *enc, h8 (&B)[4][2])
{
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    for (int rb = 0; rb < 4; ++rb)
        for (int c = 0; c < 2; ++c)
            B[rb][c] = lds_frag(enc + (rb * 16 + r) * kLdsPitch, 32 * c + 4 * g);
}

// z = W_l h + b_l for 64 rows, D layout, l < 4.
__device__ __forceinline__ void layer64(const _Float16 *sW, const float *bias,
                                        int l, const h8 (&B)[4][2],
                                        f4 (&acc)[4][4])
{
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    for (int mb = 0; mb < 4; ++mb)
    {
        f4 b = *(const f4 *)(bias + l * 64 + mb * 16 + 4 * g);
        for (int rb = 0; rb < 4; ++rb) acc[rb][mb] = b;
    }
    for (int c = 0; c < 2; ++c)
        for (int mb = 0; mb < 4; ++mb)
        {
            h8 a = lds_frag(sW + (l * 64 + mb * 16 + r) * kLdsPitch, 32 * c + 4 * g);
            for (int rb = 0; rb < 4; ++rb)
                acc[rb][mb] = mfma16x16x32(a, B[rb][c], acc[rb][mb]);
        }
}

// ReLU, ReLU mask (bit (rb*4 + mb)*4 + i) and the next layer's B operands.
__device__ __forceinline__ uint64_t relu_to_b(f4 (&acc)[4][4], h8 (&B)[4][2])
{
    uint64_t mask = 0;
    for (int rb = 0; rb < 4; ++rb)
        for (int mb = 0; mb < 4; ++mb)
            for (int i = 0; i < 4; ++i)
            {
                float x = acc[rb][mb][i];
                bool on = x > 0.0f;
                acc[rb][mb][i] = on ? x : 0.0f;
                if (on) mask |= 1ull << ((rb * 4 + mb) * 4 + i);
            }
    for (int rb = 0; rb < 4; ++rb)
        for (int c = 0; c < 2; ++c)
            B[rb][c] = pack_tiles(acc[rb][2 * c], acc[rb][2 * c + 1]);
    return mask;
}

// Output layer (48 rows), D layout.
__device__ __forceinline__ void layer_out(const _Float16 *sW, const float *bias,
                                          const h8 (&B)[4][2], f4 (&acc)[4][3])
{
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    for (int mb = 0; mb < 3; ++mb)
    {
        f4 b = *(const f4 *)(bias + 256 + mb * 16 + 4 * g);
        for (int rb = 0; rb < 4; ++rb) acc[rb][mb] = b;
    }
    for (int c = 0; c < 2; ++c)
        for (int mb = 0; mb < 3; ++mb)
        {
            h8 a = lds_frag(sW + (256 + mb * 16 + r) * kLdsPitch, 32 * c + 4 * g);
            for (int rb = 0; rb < 4; ++rb)
                acc[rb][mb] = mfma16x16x32(a, B[rb][c], acc[rb][mb]);
        }
}

// Outputs (D layout) -> wave scratch [48][64], out-major.
__device__ __forceinline__ void store_outputs(float *scr, const f4 (&acc)[4][3])
{
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    for (int rb = 0; rb < 4; ++rb)
        for (int mb = 0; mb < 3; ++mb)
            for (int i = 0; i < 4; ++i)
                scr[(mb * 16 + 4 * g + i) * 64 + rb * 16 + r] = acc[rb][mb][i];
}

// exp, 1/x and 1/sqrt(x) with the hardware's v_exp_f32 / v_rcp_f32 /
// v_rsq_f32 (within a few units in the last place): the mixture is a
// proposal, sampled and
// evaluated from the same stored numbers, so its rounding does not bias the
// estimator.
__device__ __forceinline__ float fast_exp(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __expf(x);
#else
    return expf(x);
#endif
}
__device__ __forceinline__ float fast_rcp(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}
__device__ __forceinline__ float fast_rsqrt(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rsqf(x);
#else
    return 1.0f / sqrtf(x);
#endif
}

// The mixture of one ray from its 48 outputs (o[k] at stride 64).
__device__ __forceinline__ void outputs_to_lobes(const float *o, Lobes &L)
{
    float mx = o[0];
    for (int k = 1; k < kGLobes; ++k) mx = fmaxf(mx, o[k * 64]);
    float s = 0.0f;
    for (int k = 0; k < kGLobes; ++k) s += (L.pi[k] = fast_exp(o[k * 64] - mx));
    float inv = fast_rcp(s);
    for (int k = 0; k < kGLobes; ++k)
    {
        L.pi[k] *= inv;
        float sg = fast_rcp(1.0f + fast_exp(-o[(8 + k) * 64]));
        L.kappa[k] = fast_exp(kLogKappa0 + kLogKappa1 * sg);
        float mx3 = o[(16 + 3 * k) * 64], my = o[(17 + 3 * k) * 64],
              mz = o[(18 + 3 * k) * 64];
        float len2 = mx3 * mx3 + my * my + mz * mz;
        if (len2 < 1e-38f)
        {
            mx3 = 0.0f;
            my = 0.0f;
            mz = 1.0f;
            len2 = 1.0f;
        }
        float il = fast_rsqrt(len2);
        L.mu[k][0] = mx3 * il;
        L.mu[k][1] = my * il;
        L.mu[k][2] = mz * il;
    }
}

__device__ __forceinline__ void store_lobes(float *dst, uint32_t cap, uint32_t i,
                                            const Lobes &L)
{
    for (int k = 0; k < kGLobes; ++k)
    {
        dst[(size_t)k * cap + i] = L.pi[k];
        dst[(size_t)(8 + k) * cap + i] = L.kappa[k];
        dst[(size_t)(16 + 3 * k) * cap + i] = L.mu[k][0];
        dst[(size_t)(17 + 3 * k) * cap + i] = L.mu[k][1];
        dst[(size_t)(18 + 3 * k) * cap + i] = L.mu[k][2];
    }
}

// Normalized position of a hit in the scene bounds.
__device__ __forceinline__ v3 guide_pn(const DevGuide &G, v3 P)
{
    return v3{fminf(fmaxf((P.x - G.lo[0]) * G.invExt[0], 0.0f), 1.0f),
              fminf(fmaxf((P.y - G.lo[1]) * G.invExt[1], 0.0f), 1.0f),
              fminf(fmaxf((P.z - G.lo[2]) * G.invExt[2], 0.0f), 1.0f)};
}

}  // namespace spg
