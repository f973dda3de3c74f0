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
// Per-wave scratch of the inference kernels, which pass the encodings in
// two halves of 32 features ([64][kEncPitch] halves; 72-byte rows keep a
// fragment read's 16 rows on distinct bank pairs) and the outputs in tiles
// of 16 rows ([16][64] floats).
constexpr int kEncPitch = 36;
constexpr int kTileScratchFloats =
    64 * kEncPitch / 2 > 16 * 64 ? 64 * kEncPitch / 2 : 16 * 64;

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

// Encoding of one hit into 64 features (see spg_guide.h), half kHalf of it
// (features 32 kHalf .. 32 kHalf + 31) written to row[0..31].  Invalid rays
// get zeros.  Feature 2k, 2k + 1 (k < 24) are sin, cos of 2^l pi p_a with
// l = k / 3, a = k % 3: 2^(l-1) p_a revolutions (at most 64) for p in
// [0, 1]; then the raw position, normal, view direction, 1, zeros.
template <int kHalf>
__device__ __forceinline__ void encode_half(_Float16 *row, bool valid, v3 pn, v3 n,
                                            v3 v)
{
    float e[32];
    const float p[3] = {pn.x, pn.y, pn.z};
    for (int k = 16 * kHalf; k < (kHalf ? 24 : 16); ++k)
    {
        float s, c;
        sincos_rev(ldexpf(p[k % 3], k / 3 - 1), &s, &c);
        e[2 * k - 32 * kHalf] = s;
        e[2 * k + 1 - 32 * kHalf] = c;
    }
    if (kHalf)
    {
        e[16] = pn.x;
        e[17] = pn.y;
        e[18] = pn.z;
        e[19] = n.x;
        e[20] = n.y;
        e[21] = n.z;
        e[22] = v.x;
        e[23] = v.y;
        e[24] = v.z;
        e[25] = 1.0f;
        for (int k = 26; k < 32; ++k) e[k] = 0.0f;
    }
    for (int k = 0; k < 32; k += 4)
    {
        h4 q;
        for (int j = 0; j < 4; ++j) q[j] = valid ? (_Float16)e[k + j] : (_Float16)0.0f;
        *(h4 *)(row + k) = q;
    }
}

// All 64 features as one LDS row of the wave's encoding tile.
__device__ __forceinline__ void encode_row(_Float16 *row, bool valid, v3 pn, v3 n,
                                           v3 v)
{
    encode_half<0>(row, valid, pn, n, v);
    encode_half<1>(row + 32, valid, pn, n, v);
}

// B operands of the input layer from the wave's encoding tile.
__device__ __forceinline__ void load_input(const _Float16 *enc, h8 (&B)[4][2])
{
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    for (int rb = 0; rb < 4; ++rb)
        for (int c = 0; c < 2; ++c)
            B[rb][c] = lds_frag(enc + (rb * 16 + r) * kLdsPitch, 32 * c + 4 * g);
}

// K step c of the input layer's B operands from a tile of encoding halves
// (rows of kEncPitch halves holding features 32 c .. 32 c + 31).
__device__ __forceinline__ void load_input_half(const _Float16 *enc, h8 (&B)[4][2],
                                                int c)
{
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    for (int rb = 0; rb < 4; ++rb)
        B[rb][c] = lds_frag(enc + (rb * 16 + r) * kEncPitch, 4 * g);
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

// Output rows 16 mb .. 16 mb + 15 (D layout) -> wave scratch [16][64].
__device__ __forceinline__ void store_output_tile(float *scr, const f4 (&acc)[4][3],
                                                  int mb)
{
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    for (int rb = 0; rb < 4; ++rb)
        for (int i = 0; i < 4; ++i)
            scr[(4 * g + i) * 64 + rb * 16 + r] = acc[rb][mb][i];
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

// Mixture weights and concentrations of one ray from outputs 0..15 (o[k]
// at stride 64): pi = softmax(o[0..7]), log kappa = kLogKappa0 +
// kLogKappa1 sigmoid(o[8 + k]).
__device__ __forceinline__ void lobes_weights(const float *o, Lobes &L)
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
    }
}

// Unnormalized mean directions from outputs first .. first + count - 1 (o[k]
// at stride 64 is output first + k): output 16 + m is component m % 3 of
// lobe m / 3 (m < 24); outputs 40..47 are unused.
__device__ __forceinline__ void lobes_mean_raw(const float *o, int first, int count,
                                               Lobes &L)
{
    for (int k = 0; k < count; ++k)
    {
        const int m = first + k - 16;
        if (m >= 0 && m < 3 * kGLobes) L.mu[m / 3][m % 3] = o[k * 64];
    }
}

// Mean directions normalized (a zero vector becomes +z).
__device__ __forceinline__ void lobes_normalize(Lobes &L)
{
    for (int k = 0; k < kGLobes; ++k)
    {
        float mx3 = L.mu[k][0], my = L.mu[k][1], mz = L.mu[k][2];
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

// The mixture of one ray from its 48 outputs (o[k] at stride 64).
__device__ __forceinline__ void outputs_to_lobes(const float *o, Lobes &L)
{
    lobes_weights(o, L);
    lobes_mean_raw(o + 16 * 64, 16, 32, L);
    lobes_normalize(L);
}

// One wave's 64 rays through the guide, for the inference kernels: the
// encodings in two halves through the wave's scratch (kTileScratchFloats),
// the MLP on MFMA (weights sW in LDS, biases bias), the outputs in three
// tiles of 16 rows through the same scratch, decoded into the lane's
// mixture.  raw(mb, o) sees each tile first (o[k * 64] = output 16 mb + k
// of the lane's ray).  Every lane of the wave takes part.
template <class Raw>
__device__ __forceinline__ void guide_forward_tile(const _Float16 *sW, const float *bias,
                                                   float *scr, bool valid, v3 pn, v3 n,
                                                   v3 v, Lobes &L, Raw raw)
{
    const int lane = threadIdx.x & 63;
    _Float16 *enc = (_Float16 *)scr;
    h8 B[4][2];
    encode_half<0>(enc + lane * kEncPitch, valid, pn, n, v);
    wave_lds_sync();
    load_input_half(enc, B, 0);
    wave_lds_sync();
    encode_half<1>(enc + lane * kEncPitch, valid, pn, n, v);
    wave_lds_sync();
    load_input_half(enc, B, 1);
    wave_lds_sync();
    f4 acc[4][4];
    for (int l = 0; l < 4; ++l)
    {
        layer64(sW, bias, l, B, acc);
        relu_to_b(acc, B);
    }
    f4 out[4][3];
    layer_out(sW, bias, B, out);
    for (int mb = 0; mb < 3; ++mb)
    {
        store_output_tile(scr, out, mb);
        wave_lds_sync();
        const float *o = scr + lane;
        raw(mb, o);
        if (mb == 0) lobes_weights(o, L);
        else lobes_mean_raw(o, 16 * mb, 16, L);
        wave_lds_sync();
    }
    lobes_normalize(L);
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
