// gfx950 kernels of the guide's network: evaluation on given inputs, the
// training step (forward + backward over a batch, weight gradients, Adam)
// and the weight conversions.  Inference on the tracer's hit queues is
// k_guide_infer in spg_trace.hip; both use the MFMA blocks of spg_mlp.h.
//
// Training step over a batch of B samples (pn, n, v, wi, weight):
//   k_guide_wsum    sum of the batch's weights (normalization to mean 1)
//   k_guide_fwdbwd  per 64-sample tile: forward (keeping the ReLU masks),
//                   d(-w log q)/d(outputs) per sample, backward through the
//                   layers with the transposed weights; writes every
//                   layer's input H_l and output gradient G_l (FP16,
//                   feature-major [64][B])
//   k_guide_wgrad   dW_l = G_l H_l^T and db_l = G_l 1 as MFMA products with
//                   the samples on K, one FP32 partial gradient per chunk
//   k_guide_greduce the partials added in chunk order (the step's result
//                   does not depend on scheduling)
//   k_guide_adam    Adam on the FP32 master parameters, refreshes the FP16
//                   forward and transposed copies, clears the gradient
// Gradients carry a fixed loss scale (kLossScale) while in FP16 and are
// clamped to the FP16 range; the FP32 accumulation is unscaled in Adam.
#include "spg_internal.h"
#include "spg_mlp.h"

namespace spg {

constexpr int kEvalWaves = 4;
constexpr int kTrainWaves = 4;
constexpr float kFp16Max = 60000.0f;

__device__ __forceinline__ float clamp16(float x)
{
    return fminf(fmaxf(x, -kFp16Max), kFp16Max);
}

// Parameter k (flat layout) into the FP16 forward copy, the FP16 transposed
// copy (layer l: [in][out], out padded to 64) or the FP32 bias copy.
__device__ __forceinline__ void write_param_copies(const GuideTrainArgs &a,
                                                   uint32_t k, float p)
{
    int l = 0;
    while (l < 4 && k >= (uint32_t)g_w_off(l + 1)) ++l;
    uint32_t rel = k - (uint32_t)g_w_off(l);
    uint32_t nw = (uint32_t)g_rows(l) * 64u;
    if (rel < nw)
    {
        uint32_t o = rel / 64u, i = rel % 64u;
        a.w16[(l * 64 + o) * 64 + i] = (_Float16)p;
        a.wt16[(l * 64 + i) * 64 + o] = (_Float16)p;
    }
    else
        a.bias[l * 64 + (rel - nw)] = p;
}

// Samples in the step's batch: [first, first + n).
__device__ __forceinline__ uint32_t batch_count(const GuideTrainArgs &a)
{
    uint32_t have = *a.count;
    if (have > a.sampleCap) have = a.sampleCap;
    if (have <= a.first) return 0;
    uint32_t n = have - a.first;
    return n < a.batch ? n : a.batch;
}

// ---------------------------------------------------------------- eval
__global__ __launch_bounds__(64 * kEvalWaves) void k_guide_eval(
    const _Float16 *w16, const float *bias, const float *in9, uint32_t n,
    float *out48, float *lobes40)
{
    __shared__ _Float16 sW[kWLdsHalves];
    __shared__ float sBias[kGRows];
    __shared__ float scratch[kEvalWaves][kTileScratchFloats];
    stage_rows(sW, w16, kGRows);
    for (int k = threadIdx.x; k < kGRows; k += blockDim.x) sBias[k] = bias[k];
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    float *scr = scratch[wave];
    for (uint32_t tile = blockIdx.x * kEvalWaves + wave; tile * 64 < n;
         tile += gridDim.x * kEvalWaves)
    {
        uint32_t s = tile * 64 + lane;
        bool valid = s < n;
        v3 pn{0, 0, 0}, nn{0, 0, 0}, vv{0, 0, 0};
        if (valid)
        {
            const float *x = in9 + (size_t)s * 9;
            pn = v3{x[0], x[1], x[2]};
            nn = v3{x[3], x[4], x[5]};
            vv = v3{x[6], x[7], x[8]};
        }
        Lobes L;
        // the same per-wave path as the tracer's inference kernel
        guide_forward_tile(sW, sBias, scr, valid, pn, nn, vv, L,
                           [&](int mb, const float *o) {
                               if (valid)
                                   for (int k = 0; k < 16; ++k)
                                       out48[(size_t)s * kGOut + 16 * mb + k] = o[k * 64];
                           });
        if (valid)
        {
            float *d = lobes40 + (size_t)s * kGLobeFloats;
            for (int k = 0; k < kGLobes; ++k)
            {
                d[k] = L.pi[k];
                d[8 + k] = L.kappa[k];
                d[16 + 3 * k] = L.mu[k][0];
                d[17 + 3 * k] = L.mu[k][1];
                d[18 + 3 * k] = L.mu[k][2];
            }
        }
    }
}

// ------------------------------------------------------------ training
// Sum of the batch's weights in a fixed order (one block: each thread sums
// a fixed stride of the batch, then a fixed tree), so a training step's
// result does not depend on scheduling.
__global__ __launch_bounds__(256) void k_guide_wsum(GuideTrainArgs a)
{
    __shared__ float part[256];
    const uint32_t n = batch_count(a);
    float s = 0.0f;
    for (uint32_t j = threadIdx.x; j < n; j += 256)
        s += a.samples[(size_t)12 * a.sampleCap + a.first + j];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1)
    {
        if ((int)threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) *a.wsum = part[0];
}

// d(-w log q)/d(outputs) of one sample (o[k] at stride 64), FP32; returns
// -w log q.  Same formulas as the FP64 oracle.
__device__ __forceinline__ float output_grads(const float *o, v3 wi, float w,
                                              float *g /* stride 64 */)
{
    Lobes L;
    outputs_to_lobes(o, L);
    float la[kGLobes], mx = -INFINITY;
    for (int k = 0; k < kGLobes; ++k)
    {
        float d = L.mu[k][0] * wi.x + L.mu[k][1] * wi.y + L.mu[k][2] * wi.z;
        float kap = L.kappa[k];
        la[k] = logf(L.pi[k]) + logf(kap) - 1.83787706641f /* log 2pi */ -
                logf(-expm1f(-2.0f * kap)) + kap * (d - 1.0f);
        mx = fmaxf(mx, la[k]);
    }
    float s = 0.0f;
    for (int k = 0; k < kGLobes; ++k) s += expf(la[k] - mx);
    float lq = mx + logf(s);
    for (int k = 0; k < kGLobes; ++k)
    {
        float gam = expf(la[k] - lq);
        float kap = L.kappa[k];
        float d = L.mu[k][0] * wi.x + L.mu[k][1] * wi.y + L.mu[k][2] * wi.z;
        g[k * 64] = w * (L.pi[k] - gam);
        float sg = 1.0f / (1.0f + expf(-o[(8 + k) * 64]));
        float dlogv = 1.0f - 2.0f * kap / expm1f(2.0f * kap) + kap * (d - 1.0f);
        g[(8 + k) * 64] = -w * gam * dlogv * kLogKappa1 * sg * (1.0f - sg);
        float mx3 = o[(16 + 3 * k) * 64], my = o[(17 + 3 * k) * 64],
              mz = o[(18 + 3 * k) * 64];
        float len = sqrtf(mx3 * mx3 + my * my + mz * mz);
        float f = len > 1e-20f ? -w * gam * kap / len : 0.0f;
        g[(16 + 3 * k) * 64] = f * (wi.x - L.mu[k][0] * d);
        g[(17 + 3 * k) * 64] = f * (wi.y - L.mu[k][1] * d);
        g[(18 + 3 * k) * 64] = f * (wi.z - L.mu[k][2] * d);
    }
    for (int k = kGLobeFloats; k < kGOut; ++k) g[k * 64] = 0.0f;
    return -w * lq;
}

// Stores a D-layout tile set (rows 16 mb + 4 g + i, samples rb 16 + r) of
// FP16 values, feature-major [rows][batch].
__device__ __forceinline__ void store_tiles(_Float16 *dst, uint32_t batch,
                                            uint32_t s0, const f4 (&t)[4][4],
                                            int mbs)
{
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    for (int rb = 0; rb < 4; ++rb)
        for (int mb = 0; mb < mbs; ++mb)
            for (int i = 0; i < 4; ++i)
                dst[(size_t)(mb * 16 + 4 * g + i) * batch + s0 + rb * 16 + r] =
                    (_Float16)t[rb][mb][i];
}

__global__ __launch_bounds__(64 * kTrainWaves) void k_guide_fwdbwd(GuideTrainArgs a)
{
    __shared__ _Float16 sW[kWLdsHalves];
    __shared__ _Float16 sWt[kWtLdsHalves];
    __shared__ float scratch[kTrainWaves][kWaveScratchFloats];
    stage_rows(sW, a.w16, kGRows);
    stage_rows(sWt, a.wt16, kGLayers * 64);
    __syncthreads();
    const uint32_t n = batch_count(a);
    const uint32_t tiles = (n + 63) / 64;
    const float wsum = a.wsum ? *a.wsum : 0.0f;
    const float norm = a.wsum ? (wsum > 0.0f ? (float)n / wsum : 0.0f) : 1.0f;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = lane & 15, g = lane >> 4;
    float *scr = scratch[wave];
    _Float16 *enc = (_Float16 *)scr;
    const size_t layerStride = (size_t)64 * a.batch;
    float lossSum = 0.0f;
    for (uint32_t tile = blockIdx.x * kTrainWaves + wave; tile < tiles;
         tile += gridDim.x * kTrainWaves)
    {
        const uint32_t s0 = tile * 64;
        const uint32_t s = s0 + lane;
        const bool valid = s < n;
        v3 pn{0, 0, 0}, nn{0, 0, 0}, vv{0, 0, 0}, wi{0, 0, 1};
        float w = 0.0f;
        if (valid)
        {
            const float *x = a.samples + a.first + s;
            const size_t C = a.sampleCap;
            pn = v3{x[0], x[C], x[2 * C]};
            nn = v3{x[3 * C], x[4 * C], x[5 * C]};
            vv = v3{x[6 * C], x[7 * C], x[8 * C]};
            wi = v3{x[9 * C], x[10 * C], x[11 * C]};
            w = x[12 * C] * norm;
            if (a.wmax > 0.0f) w = fminf(w, a.wmax);
        }
        // input layer: encodings, kept for the weight gradient of layer 0
        _Float16 *row = enc + lane * kLdsPitch;
        encode_row(row, valid, pn, nn, vv);
        for (int f = 0; f < kGIn; ++f) a.H[(size_t)f * a.batch + s] = row[f];
        wave_lds_sync();
        h8 B[4][2];
        load_input(enc, B);
        wave_lds_sync();
        // forward
        uint64_t mask[4];
        f4 acc[4][4];
        for (int l = 0; l < 4; ++l)
        {
            layer64(sW, a.bias, l, B, acc);
            mask[l] = relu_to_b(acc, B);
            store_tiles(a.H + (l + 1) * layerStride, a.batch, s0, acc, 4);
        }
        f4 out[4][3];
        layer_out(sW, a.bias, B, out);
        store_outputs(scr, out);
        wave_lds_sync();
        // output gradients, one sample per lane, in place, loss-scaled
        if (valid)
            lossSum += output_grads(scr + lane, wi, w, scr + lane);
        for (int k = 0; k < kGOut; ++k)
            scr[k * 64 + lane] = valid ? clamp16(scr[k * 64 + lane] * kLossScale) : 0.0f;
        wave_lds_sync();
        f4 gt[4][4];
        for (int rb = 0; rb < 4; ++rb)
        {
            for (int mb = 0; mb < 3; ++mb)
                for (int i = 0; i < 4; ++i)
                    gt[rb][mb][i] = scr[(mb * 16 + 4 * g + i) * 64 + rb * 16 + r];
            gt[rb][3] = f4{0.0f, 0.0f, 0.0f, 0.0f};
        }
        wave_lds_sync();
        store_tiles(a.G + 4 * layerStride, a.batch, s0, gt, 3);
        h8 Bg[4][2];
        for (int rb = 0; rb < 4; ++rb)
            for (int c = 0; c < 2; ++c)
                Bg[rb][c] = pack_tiles(gt[rb][2 * c], gt[rb][2 * c + 1]);
        // backward: G_{l-1} = (W_l^T G_l) o relu'(z_{l-1})
        for (int l = 4; l >= 1; --l)
        {
            for (int rb = 0; rb < 4; ++rb)
                for (int mb = 0; mb < 4; ++mb) acc[rb][mb] = f4{0.0f, 0.0f, 0.0f, 0.0f};
            for (int c = 0; c < 2; ++c)
                for (int mb = 0; mb < 4; ++mb)
                {
                    h8 at = lds_frag(sWt + (l * 64 + mb * 16 + r) * kLdsPitch,
                                     32 * c + 4 * g);
                    for (int rb = 0; rb < 4; ++rb)
                        acc[rb][mb] = mfma16x16x32(at, Bg[rb][c], acc[rb][mb]);
                }
            const uint64_t m = mask[l - 1];
            for (int rb = 0; rb < 4; ++rb)
                for (int mb = 0; mb < 4; ++mb)
                    for (int i = 0; i < 4; ++i)
                    {
                        bool on = (m >> ((rb * 4 + mb) * 4 + i)) & 1ull;
                        acc[rb][mb][i] = on ? clamp16(acc[rb][mb][i]) : 0.0f;
                    }
            store_tiles(a.G + (l - 1) * layerStride, a.batch, s0, acc, 4);
            for (int rb = 0; rb < 4; ++rb)
                for (int c = 0; c < 2; ++c)
                    Bg[rb][c] = pack_tiles(acc[rb][2 * c], acc[rb][2 * c + 1]);
        }
    }
    for (int off = 32; off > 0; off >>= 1) lossSum += __shfl_xor(lossSum, off);
    if (lane == 0 && lossSum != 0.0f) atomicAdd(a.loss, lossSum);
}

// dW_l = G_l H_l^T, db_l = G_l 1 over one chunk of the batch, stored as
// that chunk's partial gradient (every element of the layer written once);
// k_guide_greduce adds the chunks in order.  Block = one wave, blockIdx.y =
// layer, kMb = output rows / 16.
template <int kMb>
__device__ __forceinline__ void wgrad_chunk(const GuideTrainArgs &a, int l,
                                            uint32_t k0, uint32_t k1)
{
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    const size_t layerStride = (size_t)64 * a.batch;
    const _Float16 *G = a.G + l * layerStride;
    const _Float16 *H = a.H + l * layerStride;
    f4 acc[kMb][4];
    float db[kMb];
    for (int mb = 0; mb < kMb; ++mb)
    {
        db[mb] = 0.0f;
        for (int nb = 0; nb < 4; ++nb) acc[mb][nb] = f4{0.0f, 0.0f, 0.0f, 0.0f};
    }
    for (uint32_t ks = k0; ks < k1; ks += 32)
    {
        h8 av[kMb], bv[4];
        for (int mb = 0; mb < kMb; ++mb)
            av[mb] = *(const h8 *)(G + (size_t)(mb * 16 + r) * a.batch + ks + 8 * g);
        for (int nb = 0; nb < 4; ++nb)
            bv[nb] = *(const h8 *)(H + (size_t)(nb * 16 + r) * a.batch + ks + 8 * g);
        for (int mb = 0; mb < kMb; ++mb)
        {
            for (int nb = 0; nb < 4; ++nb)
                acc[mb][nb] = mfma16x16x32(av[mb], bv[nb], acc[mb][nb]);
            for (int i = 0; i < 8; ++i) db[mb] += (float)av[mb][i];
        }
    }
    float *part = a.part + (size_t)(k0 / kGradChunk) * kGParams;
    float *dW = part + g_w_off(l);
    float *dB = part + g_b_off(l);
    for (int mb = 0; mb < kMb; ++mb)
    {
        float v = db[mb];
        v += __shfl_xor(v, 16);
        v += __shfl_xor(v, 32);
        if (g == 0) dB[mb * 16 + r] = v;
        for (int nb = 0; nb < 4; ++nb)
            for (int i = 0; i < 4; ++i)
                dW[(mb * 16 + 4 * g + i) * 64 + nb * 16 + r] = acc[mb][nb][i];
    }
}

__global__ __launch_bounds__(64) void k_guide_wgrad(GuideTrainArgs a)
{
    const uint32_t n = batch_count(a);
    const uint32_t nPad = (n + 63) / 64 * 64;
    const uint32_t k0 = blockIdx.x * kGradChunk;
    if (k0 >= nPad) return;
    const uint32_t k1 = min(k0 + (uint32_t)kGradChunk, nPad);
    const int l = blockIdx.y;
    if (l < 4) wgrad_chunk<4>(a, l, k0, k1);
    else wgrad_chunk<3>(a, l, k0, k1);
}

// The gradient: the chunks' partial gradients added in chunk order.
__global__ __launch_bounds__(256) void k_guide_greduce(GuideTrainArgs a)
{
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= (uint32_t)kGParams) return;
    const uint32_t nPad = (batch_count(a) + 63) / 64 * 64;
    const uint32_t chunks = (nPad + kGradChunk - 1) / kGradChunk;
    float g = 0.0f;
    for (uint32_t c = 0; c < chunks; ++c) g += a.part[(size_t)c * kGParams + k];
    a.grad[k] = g;
}

// Adam on the master parameters, then the FP16 forward / transposed copies
// and the FP32 biases; clears the gradient.  gradScale = 1 / (n * loss
// scale) is applied here (n read from the device).
__global__ __launch_bounds__(256) void k_guide_adam(GuideTrainArgs a, GuideAdam h)
{
    const uint32_t n = batch_count(a);
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= (uint32_t)kGParams) return;
    float p = a.params[k];
    if (n > 0 && h.apply)
    {
        // bias corrections 1 - beta^t in double, rounded once
        const double t = (double)*h.step;
        const float c1 = (float)(1.0 - pow((double)h.beta1, t));
        const float c2 = (float)(1.0 - pow((double)h.beta2, t));
        float gr = a.grad[k] * (1.0f / ((float)n * kLossScale));
        float m = h.beta1 * a.m[k] + (1.0f - h.beta1) * gr;
        float v = h.beta2 * a.v[k] + (1.0f - h.beta2) * gr * gr;
        a.m[k] = m;
        a.v[k] = v;
        p -= h.lr * (m / c1) / (sqrtf(v / c2) + h.eps);
        a.params[k] = p;
    }
    a.grad[k] = 0.0f;
    write_param_copies(a, k, p);
}

// Refresh the FP16 / bias copies from the master parameters (after set).
__global__ __launch_bounds__(256) void k_guide_convert(GuideTrainArgs a)
{
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= (uint32_t)kGParams) return;
    write_param_copies(a, k, a.params[k]);
}

// ------------------------------------------------------------ launchers
hipError_t launch_guide_eval(hipStream_t s, const _Float16 *w16, const float *bias,
                             const float *in9, uint32_t n, float *out48,
                             float *lobes40)
{
    if (n == 0) return hipSuccess;
    uint32_t tiles = (n + 63) / 64;
    uint32_t blocks = (tiles + kEvalWaves - 1) / kEvalWaves;
    hipLaunchKernelGGL(k_guide_eval, dim3(blocks < 1024 ? blocks : 1024),
                       dim3(64 * kEvalWaves), 0, s, w16, bias, in9, n, out48, lobes40);
    return hipGetLastError();
}

hipError_t launch_guide_convert(hipStream_t s, const GuideTrainArgs &a)
{
    hipLaunchKernelGGL(k_guide_convert, dim3((kGParams + 255) / 256), dim3(256), 0,
                       s, a);
    return hipGetLastError();
}

hipError_t launch_guide_gradient(hipStream_t s, const GuideTrainArgs &a,
                                 uint32_t cus)
{
    if (a.wsum) hipLaunchKernelGGL(k_guide_wsum, dim3(1), dim3(256), 0, s, a);
    uint32_t tiles = (a.batch + 63) / 64;
    uint32_t blocks = (tiles + kTrainWaves - 1) / kTrainWaves;
    if (blocks > cus) blocks = cus;
    hipLaunchKernelGGL(k_guide_fwdbwd, dim3(blocks), dim3(64 * kTrainWaves), 0, s, a);
    uint32_t chunks = (a.batch + kGradChunk - 1) / kGradChunk;
    hipLaunchKernelGGL(k_guide_wgrad, dim3(chunks, kGLayers), dim3(64), 0, s, a);
    hipLaunchKernelGGL(k_guide_greduce, dim3((kGParams + 255) / 256), dim3(256), 0, s,
                       a);
    return hipGetLastError();
}

__global__ void k_guide_step(uint32_t *step) { *step += 1u; }

hipError_t launch_guide_step(hipStream_t s, uint32_t *step)
{
    hipLaunchKernelGGL(k_guide_step, dim3(1), dim3(1), 0, s, step);
    return hipGetLastError();
}

hipError_t launch_guide_adam(hipStream_t s, const GuideTrainArgs &a,
                             const GuideAdam &h)
{
    hipLaunchKernelGGL(k_guide_adam, dim3((kGParams + 255) / 256), dim3(256), 0, s,
                       a, h);
    return hipGetLastError();
}

}  // namespace spg
