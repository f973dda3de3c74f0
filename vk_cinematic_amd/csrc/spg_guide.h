// Learned path guide: data layout shared by the tracer kernels, the guide's
// MLP kernels and the host.  The reference has no guiding; the methods are
// Jakob 2012 (vMF sampling), Dong et al. 2023 (neural parametric mixtures
// trained online by minimizing KL(f cos Li || q)) and Mildenhall et al.
// 2020 (frequency encoding).
//
// Network: 64 inputs -> 4 hidden layers of 64 (ReLU) -> 48 outputs.
// Parameters (float, flat): W0[64][64] b0[64] W1 b1 W2 b2 W3 b3 W4[48][64]
// b4[48], W[out][in] row-major.  Outputs: o[0..7] lobe logits, o[8..15]
// kappa parameters (log kappa = kLogKappa0 + kLogKappa1 * sigmoid(o)),
// o[16+3k..18+3k] lobe k's unnormalized mean direction, o[40..47] unused.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "spg_math.h"
#include "spg_tx.h"

namespace spg {

constexpr int kGIn = 64;
constexpr int kGWidth = 64;
constexpr int kGOut = 48;
constexpr int kGLobes = 8;
constexpr int kGLayers = 5;                      // 4 hidden + output
constexpr int kGRows = 4 * 64 + 48;              // weight rows over all layers
constexpr int kGParams = 4 * (64 * 64 + 64) + 48 * 64 + 48;
constexpr int kGLobeFloats = 40;                 // pi[8] kappa[8] mu[8][3]
constexpr int kGSampleFloats = 13;               // pn n v wi weight
constexpr int kGVertexFloats = 12;               // pn n v wi
constexpr int kGradChunk = 512;                  // samples per partial weight gradient
constexpr int kGRecordFloats = 64;               // one recorded scatter event
constexpr int kGScanSeg = 4096;                  // paths per segment of the sample scan
constexpr int kGScanBlocks = 1024;               // segments (paths per pass <= 2^22)
constexpr float kLogKappa0 = -2.0f;
constexpr float kLogKappa1 = 8.0f;

__host__ __device__ constexpr int g_rows(int l) { return l < 4 ? 64 : 48; }
__host__ __device__ constexpr int g_w_off(int l) { return l * (64 * 64 + 64); }
__host__ __device__ constexpr int g_b_off(int l) { return g_w_off(l) + g_rows(l) * 64; }

// What the tracer kernels see of a guide.
struct DevGuide
{
    const _Float16 *w16;   // [kGRows][64] forward weights, fp16
    const float *bias;     // [kGRows] biases, fp32 (layer l at l*64)
    float *lobes;          // [kGLobeFloats][capacity] per queued ray
    float *gvtx;           // [kGVertexFloats][bounces][capacity]
    float *samples;        // [kGSampleFloats][sampleCap]
    uint32_t *sampleCount;
    uint32_t sampleCap;
    // deterministic collection: per pass, path p's sample count goes to
    // gcnt[v] with v = p * permInv mod paths (a bijection spreading
    // consecutive v over the image: p = v * permMul mod paths), and the
    // samples are stored in v order after an exclusive scan of gcnt
    uint32_t *gcnt;        // [capacity]
    uint32_t *scanBase;    // [kGScanBlocks] per 4096-entry segment: sum, then base
    uint32_t permMul, permInv;
    uint32_t selectBelow;  // a path feeds training if hash(path, salt) < this
    uint32_t salt;
    uint32_t train;        // record vertices and append training samples
    float bsdfFraction;    // one-sample MIS probability of BSDF sampling
    // verification hooks (off in normal use): analytic != 0 replaces the
    // network by guide_analytic_lobes; rec != nullptr records up to recCap
    // scatter events (kGRecordFloats each) of the guided estimator
    uint32_t analytic;
    float *rec;
    uint32_t *recCount;
    uint32_t recCap;
    float lo[3], invExt[3];// scene bounds -> [0,1]^3
};

// ---- the mixture, float -------------------------------------------------
struct Lobes
{
    float pi[kGLobes], kappa[kGLobes], mu[kGLobes][3];
};

// A ray's mixture read in place from the SoA lobe buffer (keeps the 40
// floats out of registers in the shading kernel).
struct LobeRef
{
    const float *L;
    uint32_t cap, i;
    __device__ __forceinline__ float at(int f) const { return L[(size_t)f * cap + i]; }
    __device__ __forceinline__ float pi(int k) const { return at(k); }
    __device__ __forceinline__ float kappa(int k) const { return at(8 + k); }
    __device__ __forceinline__ void mu(int k, float *m) const
    {
        m[0] = at(16 + 3 * k);
        m[1] = at(17 + 3 * k);
        m[2] = at(18 + 3 * k);
    }
};

// kappa / (2 pi (1 - e^-2kappa)) * e^(kappa (mu.w - 1)); kDouble: exp and
// expm1 evaluated in double and rounded once (bit-reproducible on the host)
template <bool kDouble>
__device__ __forceinline__ float vmf_pdf(float kappa, const float *mu, v3 w)
{
    float d = mu[0] * w.x + mu[1] * w.y + mu[2] * w.z;
    return kappa / (6.28318548f * -Tx<kDouble>::expm1_(-2.0f * kappa)) *
           Tx<kDouble>::exp_(kappa * (d - 1.0f));
}

template <bool kDouble>
__device__ __forceinline__ float mixture_pdf(const LobeRef &L, v3 w)
{
    float q = 0.0f;
    // not unrolled: eight inlined copies of the double-evaluated lobe density
    // would hold the shading kernel at one wave per SIMD
#pragma unroll 1
    for (int k = 0; k < kGLobes; ++k)
    {
        float m[3];
        L.mu(k, m);
        q += L.pi(k) * vmf_pdf<kDouble>(L.kappa(k), m, w);
    }
    return q;
}

// Orthonormal basis about n (Duff et al., JCGT 2017).
__device__ __forceinline__ void onb(v3 n, v3 &t1, v3 &t2)
{
    float sign = copysignf(1.0f, n.z);
    float a = -1.0f / (sign + n.z);
    float b = n.x * n.y * a;
    t1 = v3{1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x};
    t2 = v3{b, sign + n.y * n.y * a, -n.y};
}

// Jakob 2012: cos = 1 + log(u1 + (1 - u1) e^-2kappa) / kappa, uniform
// azimuth 2 pi u2 about mu.  kDouble: log, exp and the azimuth's sin/cos in
// double, each rounded once.
template <bool kDouble>
__device__ __forceinline__ v3 vmf_sample(float kappa, const float *mu, float u1,
                                         float u2)
{
    float c = 1.0f + Tx<kDouble>::log_(u1 + (1.0f - u1) * Tx<kDouble>::exp_(-2.0f * kappa)) /
                         kappa;
    c = fminf(fmaxf(c, -1.0f), 1.0f);
    float s = sqrtf(fmaxf(0.0f, 1.0f - c * c));
    float sp, cp;
    if (kDouble) tsincos(6.28318548f * u2, &sp, &cp);
    else sincospif(2.0f * u2, &sp, &cp);
    v3 m{mu[0], mu[1], mu[2]}, t1, t2;
    onb(m, t1, t2);
    return (t1 * (s * cp) + t2 * (s * sp)) + m * c;
}

// u0 in [0,1) picks the lobe by the cumulative weights.
template <bool kDouble>
__device__ __forceinline__ v3 mixture_sample(const LobeRef &L, float u0, float u1,
                                             float u2)
{
    int k = kGLobes - 1;
    float c = 0.0f;
    for (int j = 0; j < kGLobes - 1; ++j)
    {
        c += L.pi(j);
        if (u0 < c)
        {
            k = j;
            break;
        }
    }
    float m[3];
    L.mu(k, m);
    return vmf_sample<kDouble>(L.kappa(k), m, u1, u2);
}

// The verification mixture (no network): lobe 0 toward the fixed point
// (0.3, 2.5, -0.4) from the hit, kappa 40, weight 0.5; lobe 1 about the
// normal, kappa 3, weight 0.3; lobe 2 about the mirror direction
// 2 (N.V) N - V, kappa 120, weight 0.2; lobes 3..7 weight 0.  Unit vectors
// as (x, y, z) * (1 / sqrt(x^2 + y^2 + z^2)).
__device__ __forceinline__ v3 unit_rcp(v3 a)
{
    float inv = 1.0f / sqrtf(a.x * a.x + a.y * a.y + a.z * a.z);
    return v3{a.x * inv, a.y * inv, a.z * inv};
}
__device__ __forceinline__ void guide_analytic_lobes(v3 P, v3 N, v3 V, Lobes &L)
{
    const v3 c0{0.3f, 2.5f, -0.4f};
    const float nv = N.x * V.x + N.y * V.y + N.z * V.z;
    const v3 mus[3] = {unit_rcp(v3{c0.x - P.x, c0.y - P.y, c0.z - P.z}), N,
                       unit_rcp(v3{2.0f * nv * N.x - V.x, 2.0f * nv * N.y - V.y,
                                   2.0f * nv * N.z - V.z})};
    const float pis[3] = {0.5f, 0.3f, 0.2f}, kaps[3] = {40.0f, 3.0f, 120.0f};
    for (int k = 0; k < kGLobes; ++k)
    {
        L.pi[k] = k < 3 ? pis[k] : 0.0f;
        L.kappa[k] = k < 3 ? kaps[k] : 1.0f;
        L.mu[k][0] = k < 3 ? mus[k].x : 0.0f;
        L.mu[k][1] = k < 3 ? mus[k].y : 0.0f;
        L.mu[k][2] = k < 3 ? mus[k].z : 1.0f;
    }
}

__device__ __forceinline__ uint32_t guide_hash(uint32_t x)
{
    // lowbias32 (C. Wellons)
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

}  // namespace spg
