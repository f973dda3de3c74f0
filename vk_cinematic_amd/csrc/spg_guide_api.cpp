// Host side of the learned path guide: the spg_guide_* entry points and the
// tracer's hooks (buffers per trace, per-pass sample collection, training
// steps after each pass).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "spg_internal.h"

using namespace spg;

namespace {

constexpr float kBeta1 = 0.9f, kBeta2 = 0.99f, kAdamEps = 1e-8f;
constexpr uint32_t kMaxGradientSamples = 65536;

size_t take(size_t &off, size_t bytes)
{
    size_t o = off;
    off += (bytes + 255) / 256 * 256;
    return o;
}

void free_dev(void *&p)
{
    if (p) hipFree(p);
    p = nullptr;
}

uint64_t splitmix64(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

float uniform_pm(uint64_t &s, float a)
{
    double u = (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
    return (float)((2.0 * u - 1.0) * a);
}

// He-uniform hidden layers, a small output layer, lobe means spread on a
// Fibonacci lattice through the output biases (so |mean| starts near 1).
void init_params(std::vector<float> &p, uint32_t seed)
{
    p.assign(kGParams, 0.0f);
    uint64_t s = 0x5EEDull * 0x100000001ull + seed;
    const float he = std::sqrt(6.0f / 64.0f);
    for (int l = 0; l < kGLayers; ++l)
    {
        float a = l < 4 ? he : 0.1f * he;
        for (int k = 0; k < g_rows(l) * 64; ++k) p[g_w_off(l) + k] = uniform_pm(s, a);
    }
    float *b4 = p.data() + g_b_off(4);
    for (int k = 0; k < kGLobes; ++k)
    {
        double z = 1.0 - 2.0 * (k + 0.5) / kGLobes;
        double r = std::sqrt(std::max(0.0, 1.0 - z * z));
        double phi = 3.14159265358979 * (1.0 + std::sqrt(5.0)) * (k + 0.5);
        b4[16 + 3 * k] = (float)(r * std::cos(phi));
        b4[17 + 3 * k] = (float)(r * std::sin(phi));
        b4[18 + 3 * k] = (float)z;
    }
}

GuideTrainArgs make_args(spg_Guide *g, uint32_t first, uint32_t batch,
                         bool normalize, float wmax)
{
    GuideTrainArgs a{};
    a.w16 = g->w16;
    a.wt16 = g->wt16;
    a.bias = g->bias;
    a.params = g->params;
    a.m = g->m;
    a.v = g->v;
    a.grad = g->grad;
    a.samples = g->samples;
    a.count = g->count;
    a.sampleCap = g->sampleCap;
    a.first = first;
    a.batch = batch;
    a.wsum = normalize ? g->wsum : nullptr;
    a.wmax = wmax;
    a.H = g->H;
    a.G = g->G;
    a.part = g->part;
    a.loss = g->loss;
    return a;
}

int ensure_samples(spg_Guide *g, uint32_t cap)
{
    if (g->samplesBlock && g->sampleCap >= cap) return SPG_OK;
    if (g->pins)
        return fail(SPG_E_STATE, "guide: a batch has captured its training buffers, "
                                 "which cannot grow while it lives");
    free_dev(g->samplesBlock);
    g->sampleCap = 0;
    HIPCHK(hipMalloc(&g->samplesBlock, (size_t)kGSampleFloats * cap * sizeof(float)));
    g->samples = (float *)g->samplesBlock;
    g->sampleCap = cap;
    return SPG_OK;
}

int ensure_hg(spg_Guide *g, uint32_t batch)
{
    if (g->hgBlock && g->batchCap >= batch) return SPG_OK;
    if (g->pins)
        return fail(SPG_E_STATE, "guide: a batch has captured its training buffers, "
                                 "which cannot grow while it lives");
    free_dev(g->hgBlock);
    g->batchCap = 0;
    size_t half = (size_t)kGLayers * 64 * batch;
    size_t hgBytes = (2 * half * sizeof(_Float16) + 255) / 256 * 256;
    size_t chunks = (batch + kGradChunk - 1) / kGradChunk;
    HIPCHK(hipMalloc(&g->hgBlock, hgBytes + chunks * kGParams * sizeof(float)));
    g->H = (_Float16 *)g->hgBlock;
    g->G = g->H + half;
    g->part = (float *)((char *)g->hgBlock + hgBytes);
    g->batchCap = batch;
    return SPG_OK;
}

int upload_samples(spg_Guide *g, const float *rows, uint32_t n)
{
    CHK(ensure_samples(g, n));
    std::vector<float> soa((size_t)kGSampleFloats * g->sampleCap, 0.0f);
    for (uint32_t j = 0; j < n; ++j)
        for (int k = 0; k < kGSampleFloats; ++k)
            soa[(size_t)k * g->sampleCap + j] = rows[(size_t)j * kGSampleFloats + k];
    HIPCHK(hipMemcpyAsync(g->samples, soa.data(), soa.size() * sizeof(float),
                          hipMemcpyHostToDevice, g->dev->stream));
    HIPCHK(hipMemcpyAsync(g->count, &n, 4, hipMemcpyHostToDevice, g->dev->stream));
    HIPCHK(hipStreamSynchronize(g->dev->stream));
    return SPG_OK;
}

int adam_step(spg_Guide *g, const GuideTrainArgs &a, hipStream_t st)
{
    GuideAdam h{};
    h.lr = g->cfg.learningRate;
    h.beta1 = kBeta1;
    h.beta2 = kBeta2;
    h.eps = kAdamEps;
    h.step = g->stepDev;
    h.apply = 1;
    HIPCHK(launch_guide_step(st, g->stepDev));
    HIPCHK(launch_guide_adam(st, a, h));
    return SPG_OK;
}

uint32_t read_steps(spg_Guide *g)
{
    uint32_t n = 0;
    hipMemcpyAsync(&n, g->stepDev, 4, hipMemcpyDeviceToHost, g->dev->stream);
    hipStreamSynchronize(g->dev->stream);
    return n;
}

uint32_t cus_of(spg_Device *d)
{
    int n = 0;
    hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, d->ordinal);
    return n > 0 ? (uint32_t)n : 1u;
}

int set_params_locked(spg_Guide *g, const float *p)
{
    spg_Device *d = g->dev;
    HIPCHK(hipMemcpyAsync(g->params, p, sizeof(float) * kGParams,
                          hipMemcpyHostToDevice, d->stream));
    HIPCHK(hipMemsetAsync(g->m, 0, sizeof(float) * kGParams, d->stream));
    HIPCHK(hipMemsetAsync(g->v, 0, sizeof(float) * kGParams, d->stream));
    HIPCHK(hipMemsetAsync(g->grad, 0, sizeof(float) * kGParams, d->stream));
    HIPCHK(launch_set_counter(d->stream, g->stepDev, 0u));
    HIPCHK(launch_guide_convert(d->stream, make_args(g, 0, 64, false, 0.0f)));
    HIPCHK(hipStreamSynchronize(d->stream));
    return SPG_OK;
}

}  // namespace

// ------------------------------------------------------------ tracer hooks
namespace spg {

// Per-path buffers of the guided estimator (lobes per queued ray, vertex
// records, the training-sample scan) for `capacity` paths in one
// allocation *block; the pointers into it go to dg.
static int alloc_path_buffers(void *&block, uint32_t capacity, uint32_t bounces,
                              DevGuide &dg)
{
    size_t off = 0;
    size_t oL = take(off, (size_t)kGLobeFloats * capacity * 4);
    size_t oV = take(off, (size_t)kGVertexFloats * bounces * capacity * 4);
    size_t oC = take(off, (size_t)capacity * 4);
    size_t oS = take(off, (size_t)kGScanBlocks * 4);
    HIPCHK(hipMalloc(&block, off));
    dg.lobes = (float *)((char *)block + oL);
    dg.gvtx = (float *)((char *)block + oV);
    dg.gcnt = (uint32_t *)((char *)block + oC);
    dg.scanBase = (uint32_t *)((char *)block + oS);
    return SPG_OK;
}

float light_split(const uint64_t stat[2], float fallback)
{
    const double sky = (double)stat[0], em = (double)stat[1];
    if (!(sky + em > 0.0)) return fallback;
    return (float)std::min(0.95, std::max(0.05, sky / (sky + em)));
}

float guide_light_split(const spg_Guide *g, float fallback)
{
    return light_split(g->lightStat, fallback);
}

// ownBlock != nullptr: the per-path buffers are allocated there for the
// caller (a batch, whose captured graph must not see them move) instead of
// in the guide's own block.
int guide_prepare_trace(spg_Guide *g, const spg_Scene *s, uint32_t capacity,
                        uint32_t bounces, bool train, DevGuide &out, void **ownBlock)
{
    DevGuide pb{};
    if (ownBlock)
        CHK(alloc_path_buffers(*ownBlock, capacity, bounces, pb));
    else
    {
        if (!g->pathBlock || g->pathCap < capacity || g->pathBounces < bounces)
        {
            free_dev(g->pathBlock);
            g->pathCap = g->pathBounces = 0;
            DevGuide t{};
            CHK(alloc_path_buffers(g->pathBlock, capacity, bounces, t));
            g->lobes = t.lobes;
            g->gvtx = t.gvtx;
            g->gcnt = t.gcnt;
            g->scanBase = t.scanBase;
            g->pathCap = capacity;
            g->pathBounces = bounces;
        }
        pb.lobes = g->lobes;
        pb.gvtx = g->gvtx;
        pb.gcnt = g->gcnt;
        pb.scanBase = g->scanBase;
    }
    const uint32_t B = g->cfg.batchSize, steps = g->cfg.stepsPerPass;
    if (train)
    {
        CHK(ensure_samples(g, std::max<uint32_t>(4u * B * steps, 64u)));
        CHK(ensure_hg(g, B));
    }
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (const HostObject &o : s->objects)
        for (int a = 0; a < 3; ++a)
        {
            lo[a] = std::min(lo[a], o.aabbLo[a]);
            hi[a] = std::max(hi[a], o.aabbHi[a]);
        }
    DevGuide dg{};
    dg.w16 = g->w16;
    dg.bias = g->bias;
    dg.lobes = pb.lobes;
    dg.gvtx = pb.gvtx;
    dg.gcnt = pb.gcnt;
    dg.scanBase = pb.scanBase;
    dg.permMul = dg.permInv = 1;
    dg.samples = g->samples;
    dg.sampleCount = g->count;
    dg.sampleCap = g->sampleCap;
    dg.selectBelow = 0xFFFFFFFFu;
    dg.train = train ? 1u : 0u;
    dg.bsdfFraction = g->cfg.bsdfFraction;
    dg.analytic = g->analytic;
    dg.rec = g->recCap ? (float *)((char *)g->recBlock + 256) : nullptr;
    dg.recCount = g->recCap ? (uint32_t *)g->recBlock : nullptr;
    dg.recCap = g->recCap;
    for (int a = 0; a < 3; ++a)
    {
        if (!(lo[a] <= hi[a]))
        {
            lo[a] = 0.0f;
            hi[a] = 1.0f;
        }
        dg.lo[a] = lo[a];
        dg.invExt[a] = 1.0f / std::max(hi[a] - lo[a], 1e-6f);
    }
    out = dg;
    return SPG_OK;
}

// A multiplier m coprime to n near n / golden ratio and its inverse mod n:
// v -> v * m mod n visits 0..n-1 once, consecutive v far apart (a Weyl
// sequence), so the training samples stored in v order are spread over the
// image in every prefix.
static void weyl_permutation(uint32_t n, uint32_t &mul, uint32_t &inv)
{
    mul = inv = n > 1 ? 1u : 0u;
    if (n < 3) return;
    auto gcd = [](uint64_t a, uint64_t b) {
        while (b) { uint64_t t = a % b; a = b; b = t; }
        return a;
    };
    uint64_t m = (uint64_t)((double)n * 0.6180339887498949) | 1u;
    while (gcd(m % n, n) != 1) m += 2;
    m %= n;
    // inverse by the extended Euclidean algorithm
    int64_t r0 = (int64_t)n, r1 = (int64_t)m, t0 = 0, t1 = 1;
    while (r1)
    {
        int64_t q = r0 / r1, r2 = r0 - q * r1, t2 = t0 - q * t1;
        r0 = r1; r1 = r2; t0 = t1; t1 = t2;
    }
    if (t0 < 0) t0 += (int64_t)n;
    mul = (uint32_t)m;
    inv = (uint32_t)t0;
}

// Each pass feeds training from a hashed subset of its paths sized to
// about batchSize * stepsPerPass paths (a path yields a sample per
// scattering vertex that saw light).
int guide_begin_pass(spg_Guide *g, hipStream_t st, DevGuide &dg, uint32_t salt,
                     uint32_t paths)
{
    dg.salt = salt;
    weyl_permutation(paths, dg.permMul, dg.permInv);
    double want = (double)g->cfg.batchSize * g->cfg.stepsPerPass;
    double psel = paths ? std::min(1.0, want / (double)paths) : 1.0;
    dg.selectBelow = (uint32_t)std::min(4294967295.0, psel * 4294967296.0);
    if (dg.train) HIPCHK(launch_set_counter(st, g->count, 0u));
    return SPG_OK;
}

int guide_train_pass(spg_Guide *g, hipStream_t st)
{
    const uint32_t B = g->cfg.batchSize;
    const uint32_t cus = cus_of(g->dev);
    for (uint32_t k = 0; k < g->cfg.stepsPerPass; ++k)
    {
        GuideTrainArgs a = make_args(g, k * B, B, true, g->cfg.maxWeight);
        HIPCHK(launch_clear_floats(st, g->loss, 1));
        HIPCHK(launch_guide_gradient(st, a, cus));
        CHK(adam_step(g, a, st));
    }
    return SPG_OK;
}

}  // namespace spg

// ------------------------------------------------------------ C ABI
extern "C" {

void spg_guide_default_config(spg_GuideConfig *c)
{
    if (!c) return;
    c->seed = 1;
    c->learningRate = 3e-3f;
    c->batchSize = 16384;
    c->stepsPerPass = 4;
    c->maxWeight = 50.0f;
    c->bsdfFraction = 0.5f;
}

int spg_guide_create(spg_Device *dev, const spg_GuideConfig *cfg, spg_Guide **out)
{
    if (!dev || !out) return fail(SPG_E_INVALID, "spg_guide_create: NULL");
    spg_GuideConfig c;
    spg_guide_default_config(&c);
    if (cfg) c = *cfg;
    if (c.batchSize == 0 || c.batchSize % 64 || c.batchSize > (1u << 20))
        return fail(SPG_E_INVALID, "spg_guide_create: batchSize must be a multiple of 64 in 64..2^20");
    if (!(c.bsdfFraction >= 0.0f && c.bsdfFraction <= 1.0f) || c.learningRate < 0.0f)
        return fail(SPG_E_INVALID, "spg_guide_create: bsdfFraction in [0,1], learningRate >= 0");
    std::lock_guard<std::mutex> lk(dev->lock);
    if (dev->destroyRequested)
        return fail(SPG_E_STATE, "spg_guide_create: device is being destroyed");
    HIPCHK(hipSetDevice(dev->ordinal));
    spg_Guide *g = new spg_Guide();
    g->dev = dev;
    g->cfg = c;
    size_t off = 0;
    const size_t P = (size_t)kGParams * 4;
    size_t oP = take(off, P), oM = take(off, P), oV = take(off, P), oG = take(off, P),
           oB = take(off, (size_t)kGRows * 4), oS = take(off, 16), oL = take(off, 16),
           oC = take(off, 16), oT0 = take(off, 16), oW = take(off, (size_t)kGRows * 64 * 2),
           oT = take(off, (size_t)kGLayers * 64 * 64 * 2);
    hipError_t e = hipMalloc(&g->block, off);
    if (e != hipSuccess)
    {
        delete g;
        return hip_check(e, "hipMalloc(guide)");
    }
    char *b = (char *)g->block;
    g->params = (float *)(b + oP);
    g->m = (float *)(b + oM);
    g->v = (float *)(b + oV);
    g->grad = (float *)(b + oG);
    g->bias = (float *)(b + oB);
    g->wsum = (float *)(b + oS);
    g->loss = (float *)(b + oL);
    g->count = (uint32_t *)(b + oC);
    g->stepDev = (uint32_t *)(b + oT0);
    g->w16 = (_Float16 *)(b + oW);
    g->wt16 = (_Float16 *)(b + oT);
    std::vector<float> p;
    init_params(p, c.seed);
    int r = hip_check(hipMemsetAsync(g->block, 0, off, dev->stream), "hipMemset(guide)");
    if (!r) r = set_params_locked(g, p.data());
    if (r)
    {
        free_dev(g->block);
        delete g;
        return r;
    }
    dev->liveScenes++;
    *out = g;
    return SPG_OK;
}

int spg_guide_destroy(spg_Guide *g)
{
    if (!g) return SPG_OK;
    spg_Device *d = g->dev;
    bool last;
    {
        std::lock_guard<std::mutex> lk(d->lock);
        if (g->pins)
            return fail(SPG_E_STATE, "spg_guide_destroy: a batch still uses the guide");
        hipSetDevice(d->ordinal);
        hipStreamSynchronize(d->stream);
        free_dev(g->block);
        free_dev(g->samplesBlock);
        free_dev(g->hgBlock);
        free_dev(g->pathBlock);
        free_dev(g->recBlock);
        delete g;
        d->liveScenes--;
        last = d->destroyRequested && d->liveScenes == 0;
    }
    if (last) spg_device_destroy(d);
    return SPG_OK;
}

int spg_guide_get_params(spg_Guide *g, float *params, uint32_t count)
{
    if (!g || !params || count != (uint32_t)kGParams)
        return fail(SPG_E_INVALID, "spg_guide_get_params: need SPG_GUIDE_PARAMS floats");
    std::lock_guard<std::mutex> lk(g->dev->lock);
    HIPCHK(hipSetDevice(g->dev->ordinal));
    HIPCHK(hipMemcpyAsync(params, g->params, sizeof(float) * kGParams,
                          hipMemcpyDeviceToHost, g->dev->stream));
    HIPCHK(hipStreamSynchronize(g->dev->stream));
    return SPG_OK;
}

int spg_guide_set_params(spg_Guide *g, const float *params, uint32_t count)
{
    if (!g || !params || count != (uint32_t)kGParams)
        return fail(SPG_E_INVALID, "spg_guide_set_params: need SPG_GUIDE_PARAMS floats");
    std::lock_guard<std::mutex> lk(g->dev->lock);
    HIPCHK(hipSetDevice(g->dev->ordinal));
    return set_params_locked(g, params);
}

int spg_guide_evaluate(spg_Guide *g, const float *in9, uint32_t n, float *out48,
                       float *lobes40)
{
    if (!g || (!in9 && n)) return fail(SPG_E_INVALID, "spg_guide_evaluate: NULL");
    if (n == 0) return SPG_OK;
    spg_Device *d = g->dev;
    std::lock_guard<std::mutex> lk(d->lock);
    HIPCHK(hipSetDevice(d->ordinal));
    void *buf = nullptr;
    size_t bytes = (size_t)n * (9 + kGOut + kGLobeFloats) * 4;
    HIPCHK(hipMalloc(&buf, bytes));
    float *dIn = (float *)buf, *dOut = dIn + (size_t)n * 9,
          *dLobes = dOut + (size_t)n * kGOut;
    int r = hip_check(hipMemcpyAsync(dIn, in9, (size_t)n * 36, hipMemcpyHostToDevice,
                                     d->stream), "hipMemcpy");
    if (!r)
        r = hip_check(launch_guide_eval(d->stream, g->w16, g->bias, dIn, n, dOut, dLobes),
                      "k_guide_eval");
    if (!r && out48)
        r = hip_check(hipMemcpyAsync(out48, dOut, (size_t)n * kGOut * 4,
                                     hipMemcpyDeviceToHost, d->stream), "hipMemcpy");
    if (!r && lobes40)
        r = hip_check(hipMemcpyAsync(lobes40, dLobes, (size_t)n * kGLobeFloats * 4,
                                     hipMemcpyDeviceToHost, d->stream), "hipMemcpy");
    if (!r) r = hip_check(hipStreamSynchronize(d->stream), "hipStreamSynchronize");
    hipFree(buf);
    return r;
}

int spg_guide_gradient(spg_Guide *g, const float *samples, uint32_t n, int normalize,
                       float maxWeight, float *grad, float *loss)
{
    if (!g || !samples || !grad) return fail(SPG_E_INVALID, "spg_guide_gradient: NULL");
    if (n == 0 || n > kMaxGradientSamples)
        return fail(SPG_E_CAPACITY, "spg_guide_gradient: n must be in 1..65536");
    spg_Device *d = g->dev;
    std::lock_guard<std::mutex> lk(d->lock);
    HIPCHK(hipSetDevice(d->ordinal));
    CHK(upload_samples(g, samples, n));
    const uint32_t batch = (n + 63) / 64 * 64;
    CHK(ensure_hg(g, batch));
    HIPCHK(launch_clear_floats(d->stream, g->loss, 1));
    GuideTrainArgs a = make_args(g, 0, batch, normalize != 0, maxWeight);
    HIPCHK(launch_guide_gradient(d->stream, a, cus_of(d)));
    float l = 0.0f;
    HIPCHK(hipMemcpyAsync(grad, g->grad, sizeof(float) * kGParams,
                          hipMemcpyDeviceToHost, d->stream));
    HIPCHK(hipMemcpyAsync(&l, g->loss, 4, hipMemcpyDeviceToHost, d->stream));
    HIPCHK(launch_clear_floats(d->stream, g->grad, kGParams));
    HIPCHK(hipStreamSynchronize(d->stream));
    const float scale = 1.0f / ((float)n * kLossScale);
    for (int k = 0; k < kGParams; ++k) grad[k] *= scale;
    if (loss) *loss = l / (float)n;
    return SPG_OK;
}

int spg_guide_train(spg_Guide *g, const float *samples, uint32_t n, uint32_t steps,
                    float *loss)
{
    if (!g || !samples || n == 0) return fail(SPG_E_INVALID, "spg_guide_train: no samples");
    spg_Device *d = g->dev;
    std::lock_guard<std::mutex> lk(d->lock);
    HIPCHK(hipSetDevice(d->ordinal));
    CHK(upload_samples(g, samples, n));
    const uint32_t B = g->cfg.batchSize;
    CHK(ensure_hg(g, B));
    const uint32_t batches = std::max(1u, n / B);
    const uint32_t cus = cus_of(d);
    uint32_t lastN = 0;
    for (uint32_t k = 0; k < steps; ++k)
    {
        uint32_t first = (k % batches) * B;
        lastN = std::min(B, n - first);
        GuideTrainArgs a = make_args(g, first, B, true, g->cfg.maxWeight);
        HIPCHK(launch_clear_floats(d->stream, g->loss, 1));
        HIPCHK(launch_guide_gradient(d->stream, a, cus));
        CHK(adam_step(g, a, d->stream));
    }
    float l = 0.0f;
    HIPCHK(hipMemcpyAsync(&l, g->loss, 4, hipMemcpyDeviceToHost, d->stream));
    HIPCHK(hipStreamSynchronize(d->stream));
    if (loss) *loss = lastN ? l / (float)lastN : 0.0f;
    return SPG_OK;
}

int spg_guide_stats(spg_Guide *g, uint32_t *collected, uint32_t *steps, float *loss)
{
    if (!g) return fail(SPG_E_INVALID, "spg_guide_stats: NULL");
    spg_Device *d = g->dev;
    std::lock_guard<std::mutex> lk(d->lock);
    HIPCHK(hipSetDevice(d->ordinal));
    uint32_t c = 0;
    float l = 0.0f;
    HIPCHK(hipMemcpyAsync(&c, g->count, 4, hipMemcpyDeviceToHost, d->stream));
    HIPCHK(hipMemcpyAsync(&l, g->loss, 4, hipMemcpyDeviceToHost, d->stream));
    HIPCHK(hipStreamSynchronize(d->stream));
    if (collected) *collected = c;
    if (steps) *steps = read_steps(g);
    if (loss)
    {
        uint32_t have = std::min(c, g->sampleCap);
        uint32_t k = g->cfg.stepsPerPass ? g->cfg.stepsPerPass - 1 : 0;
        uint32_t first = k * g->cfg.batchSize;
        uint32_t n = have > first ? std::min(have - first, g->cfg.batchSize) : 0;
        *loss = n ? l / (float)n : 0.0f;
    }
    return SPG_OK;
}

int spg_guide_debug(spg_Guide *g, uint32_t flags, uint32_t recordCap)
{
    if (!g || (flags & ~SPG_GUIDE_ANALYTIC_MIXTURE))
        return fail(SPG_E_INVALID, "spg_guide_debug: NULL guide or unknown flag");
    if (recordCap > (1u << 24))
        return fail(SPG_E_CAPACITY, "spg_guide_debug: at most 2^24 records");
    spg_Device *d = g->dev;
    std::lock_guard<std::mutex> lk(d->lock);
    HIPCHK(hipSetDevice(d->ordinal));
    g->analytic = flags & SPG_GUIDE_ANALYTIC_MIXTURE;
    if (recordCap != g->recCap)
    {
        HIPCHK(hipStreamSynchronize(d->stream));
        free_dev(g->recBlock);
        g->recCap = 0;
        if (recordCap)
            HIPCHK(hipMalloc(&g->recBlock, 256 + (size_t)recordCap * kGRecordFloats * 4));
        g->recCap = recordCap;
    }
    if (g->recCap) HIPCHK(launch_set_counter(d->stream, (uint32_t *)g->recBlock, 0u));
    HIPCHK(hipStreamSynchronize(d->stream));
    return SPG_OK;
}

int spg_guide_debug_records(spg_Guide *g, float *records, uint32_t cap, uint32_t *count)
{
    if (!g || !count) return fail(SPG_E_INVALID, "spg_guide_debug_records: NULL");
    spg_Device *d = g->dev;
    std::lock_guard<std::mutex> lk(d->lock);
    *count = 0;
    if (!g->recCap) return SPG_OK;
    HIPCHK(hipSetDevice(d->ordinal));
    uint32_t n = 0;
    HIPCHK(hipMemcpyAsync(&n, g->recBlock, 4, hipMemcpyDeviceToHost, d->stream));
    HIPCHK(hipStreamSynchronize(d->stream));
    n = std::min(n, g->recCap);
    *count = n;
    if (records && cap)
    {
        HIPCHK(hipMemcpyAsync(records, (char *)g->recBlock + 256,
                              (size_t)std::min(n, cap) * kGRecordFloats * 4,
                              hipMemcpyDeviceToHost, d->stream));
        HIPCHK(hipStreamSynchronize(d->stream));
    }
    return SPG_OK;
}

int spg_guide_samples(spg_Guide *g, float *rows13, uint32_t cap, uint32_t *count)
{
    if (!g || !count) return fail(SPG_E_INVALID, "spg_guide_samples: NULL");
    spg_Device *d = g->dev;
    std::lock_guard<std::mutex> lk(d->lock);
    *count = 0;
    if (!g->samplesBlock) return SPG_OK;
    HIPCHK(hipSetDevice(d->ordinal));
    uint32_t n = 0;
    HIPCHK(hipMemcpyAsync(&n, g->count, 4, hipMemcpyDeviceToHost, d->stream));
    HIPCHK(hipStreamSynchronize(d->stream));
    n = std::min(n, g->sampleCap);
    *count = n;
    if (!rows13 || !cap || !n) return SPG_OK;
    const uint32_t m = std::min(n, cap);
    std::vector<float> soa((size_t)kGSampleFloats * m);
    for (int k = 0; k < kGSampleFloats; ++k)
        HIPCHK(hipMemcpyAsync(soa.data() + (size_t)k * m, g->samples + (size_t)k * g->sampleCap,
                              (size_t)m * 4, hipMemcpyDeviceToHost, d->stream));
    HIPCHK(hipStreamSynchronize(d->stream));
    for (uint32_t j = 0; j < m; ++j)
        for (int k = 0; k < kGSampleFloats; ++k)
            rows13[(size_t)j * kGSampleFloats + k] = soa[(size_t)k * m + j];
    return SPG_OK;
}

// Checkpoint layout: magic, version, parameter count, Adam steps, the
// configuration, then params, m, v (FP32, kGParams each).
static const char kGuideMagic[8] = {'S', 'P', 'G', 'G', 'U', 'I', 'D', 'E'};
constexpr uint32_t kGuideVersion = 2;  // 2: + the learned light split's sums

int spg_guide_save(spg_Guide *g, const char *path)
{
    if (!g || !path) return fail(SPG_E_INVALID, "spg_guide_save: NULL");
    std::vector<float> st((size_t)3 * kGParams);
    uint32_t steps;
    spg_GuideConfig cfg;
    {
        std::lock_guard<std::mutex> lk(g->dev->lock);
        HIPCHK(hipSetDevice(g->dev->ordinal));
        const float *src[3] = {g->params, g->m, g->v};
        for (int k = 0; k < 3; ++k)
            HIPCHK(hipMemcpyAsync(st.data() + (size_t)k * kGParams, src[k],
                                  sizeof(float) * kGParams, hipMemcpyDeviceToHost,
                                  g->dev->stream));
        HIPCHK(hipStreamSynchronize(g->dev->stream));
        steps = read_steps(g);
        cfg = g->cfg;
    }
    FILE *f = fopen(path, "wb");
    if (!f) return fail(SPG_E_INVALID, std::string("spg_guide_save: cannot open ") + path);
    const uint32_t head[3] = {kGuideVersion, (uint32_t)kGParams, steps};
    bool ok = fwrite(kGuideMagic, 8, 1, f) == 1 && fwrite(head, sizeof head, 1, f) == 1 &&
              fwrite(&cfg, sizeof cfg, 1, f) == 1 &&
              fwrite(st.data(), sizeof(float), st.size(), f) == st.size() &&
              fwrite(g->lightStat, sizeof g->lightStat, 1, f) == 1;
    ok = (fclose(f) == 0) && ok;
    return ok ? SPG_OK : fail(SPG_E_INVALID, std::string("spg_guide_save: write failed: ") + path);
}

int spg_guide_load(spg_Guide *g, const char *path)
{
    if (!g || !path) return fail(SPG_E_INVALID, "spg_guide_load: NULL");
    FILE *f = fopen(path, "rb");
    if (!f) return fail(SPG_E_INVALID, std::string("spg_guide_load: cannot open ") + path);
    char magic[8];
    uint32_t head[3];
    spg_GuideConfig cfg;
    std::vector<float> st((size_t)3 * kGParams);
    uint64_t light[2] = {0, 0};
    bool ok = fread(magic, 8, 1, f) == 1 && fread(head, sizeof head, 1, f) == 1 &&
              fread(&cfg, sizeof cfg, 1, f) == 1 &&
              memcmp(magic, kGuideMagic, 8) == 0 &&
              (head[0] == 1 || head[0] == kGuideVersion) && head[1] == (uint32_t)kGParams &&
              fread(st.data(), sizeof(float), st.size(), f) == st.size();
    // version 2 adds the learned light split's sums after the moments
    if (ok && head[0] == kGuideVersion) ok = fread(light, sizeof light, 1, f) == 1;
    char extra;
    ok = ok && fread(&extra, 1, 1, f) == 0;  // nothing after them
    fclose(f);
    if (!ok) return fail(SPG_E_INVALID, std::string("spg_guide_load: not a guide checkpoint: ") + path);
    if (cfg.batchSize == 0 || cfg.batchSize % 64 || cfg.batchSize > (1u << 20) ||
        !(cfg.bsdfFraction >= 0.0f && cfg.bsdfFraction <= 1.0f) || !(cfg.learningRate >= 0.0f))
        return fail(SPG_E_INVALID, "spg_guide_load: checkpoint configuration out of range");
    spg_Device *d = g->dev;
    std::lock_guard<std::mutex> lk(d->lock);
    HIPCHK(hipSetDevice(d->ordinal));
    float *dst[3] = {g->params, g->m, g->v};
    for (int k = 0; k < 3; ++k)
        HIPCHK(hipMemcpyAsync(dst[k], st.data() + (size_t)k * kGParams,
                              sizeof(float) * kGParams, hipMemcpyHostToDevice, d->stream));
    HIPCHK(launch_clear_floats(d->stream, g->grad, kGParams));
    HIPCHK(launch_set_counter(d->stream, g->stepDev, head[2]));
    HIPCHK(launch_guide_convert(d->stream, make_args(g, 0, 64, false, 0.0f)));
    HIPCHK(hipStreamSynchronize(d->stream));
    g->cfg = cfg;
    g->lightStat[0] = light[0];
    g->lightStat[1] = light[1];
    return SPG_OK;
}

int spg_guide_light_split(spg_Guide *g, float *skyFraction)
{
    if (!g || !skyFraction) return fail(SPG_E_INVALID, "spg_guide_light_split: NULL");
    std::lock_guard<std::mutex> lk(g->dev->lock);
    *skyFraction = guide_light_split(g, -1.0f);
    return SPG_OK;
}

int spg_guide_set_light_split(spg_Guide *g, float skyFraction)
{
    if (!g) return fail(SPG_E_INVALID, "spg_guide_set_light_split: NULL");
    if (!(skyFraction >= 0.05f && skyFraction <= 0.95f) && skyFraction != -1.0f)
        return fail(SPG_E_INVALID,
                    "spg_guide_set_light_split: sky fraction in [0.05, 0.95], or -1");
    std::lock_guard<std::mutex> lk(g->dev->lock);
    if (skyFraction == -1.0f)
    {
        g->lightStat[0] = g->lightStat[1] = 0;
        return SPG_OK;
    }
    // sums in the learned form whose split is skyFraction (in double, then
    // rounded to float as the tracer rounds it)
    g->lightStat[0] = (uint64_t)llround((double)skyFraction * 4294967296.0);
    g->lightStat[1] = (uint64_t)4294967296ull - g->lightStat[0];
    return SPG_OK;
}

}  // extern "C"
