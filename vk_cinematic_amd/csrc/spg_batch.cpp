// Batched tracing: many (scene, camera, image) items rendered as one
// wavefront, captured once in a HIP graph and replayed.  The kernels are the
// single-scene ones in their batch variants (TraceParams::slotItem set):
// raygen reads the slot's camera, the persistent closest-hit launch walks
// each ray through its own scene's hierarchy, shading reads its scene's
// materials, and the pixel writer scatters into the item's image.
#include <algorithm>
#include <cstddef>
#include <chrono>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "spg_internal.h"

using namespace spg;

struct spg_Batch
{
    spg_Device *dev = nullptr;
    spg_TraceConfig cfg{};
    uint32_t n = 0;
    std::vector<spg_Scene *> scenes;
    void *tables = nullptr;
    const DevScene *dScenes = nullptr;
    const DevCamera *dCams = nullptr;
    float *const *dImages = nullptr;
    const uint32_t *dSlotItem = nullptr, *dSlotPixel = nullptr;
    uint64_t slots = 0;
    DevPaths paths{};
    void *pathsBlock = nullptr;
    size_t pathsBytes = 0;
    void *accum = nullptr;
    // guided estimator: the guide (pinned while the batch lives), the
    // batch's own per-path guide buffers and the guide view of the kernels
    spg_Guide *guide = nullptr;
    void *guideBlock = nullptr;
    DevGuide dg{};
    uint32_t maxBlocks = 0;
    uint32_t refill = 0;  // k_extend's refill threshold (0 = the default)
    // guided + NEE: the replays learn the light split (training) and draw
    // with it; the split the items' device scenes hold now (-1: default)
    bool learnLight = false;
    float splitOnDevice = -1.0f;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    spg_BatchTuning tuning{};
    double lastMs = 0.0;
};

namespace {

void free_dev(void *&p)
{
    if (p) hipFree(p);
    p = nullptr;
}

int alloc_state(spg_Batch *b, uint32_t capacity)
{
    free_dev(b->pathsBlock);
    free_dev(b->accum);
    free_dev(b->guideBlock);
    memset(&b->paths, 0, sizeof b->paths);
    CHK(alloc_path_state(b->paths, b->pathsBlock, b->pathsBytes, capacity,
                         b->cfg.bounces));
    HIPCHK(hipMalloc(&b->accum, (size_t)3 * b->paths.capacity * sizeof(float)));
    if (b->guide)
        CHK(guide_prepare_trace(b->guide, b->scenes[0], b->paths.capacity, b->cfg.bounces,
                                b->cfg.guideTrain != 0, b->dg, &b->guideBlock));
    return SPG_OK;
}

// Per-class kernel timing of an eager (not captured) render: events around
// each launch from a ring of pairs, drained when the ring fills.
struct ClassTimer
{
    spg_Device *d;
    double ms[6] = {0};
    uint32_t launches[6] = {0};
    int next = 0;
    int cls[(kEvents - 4) / 2];
    int drain()
    {
        if (next == 0) return SPG_OK;
        HIPCHK(hipEventSynchronize(d->ev[next - 1]));
        for (int k = 0; k < next; k += 2)
        {
            float t = 0.0f;
            HIPCHK(hipEventElapsedTime(&t, d->ev[k], d->ev[k + 1]));
            ms[cls[k / 2]] += t;
            launches[cls[k / 2]]++;
        }
        next = 0;
        return SPG_OK;
    }
    int begin(hipStream_t st, int c)
    {
        if (next >= kEvents - 4) CHK(drain());
        cls[next / 2] = c;
        HIPCHK(hipEventRecord(d->ev[next], st));
        return SPG_OK;
    }
    int end(hipStream_t st)
    {
        HIPCHK(hipEventRecord(d->ev[next + 1], st));
        next += 2;
        return SPG_OK;
    }
};

// The whole render as stream work: slot chunks of at most the path
// capacity, sample passes per chunk, bounces per pass.  No host copies or
// memsets, so the same sequence can be captured in a graph.  tm (eager runs
// only) times every launch by kernel class.
int record(spg_Batch *b, hipStream_t st, const spg_TraceConfig &cfg,
           uint32_t *kernels, ClassTimer *tm = nullptr)
{
    auto T0 = [&](int c) { return tm ? tm->begin(st, c) : SPG_OK; };
    auto T1 = [&]() { return tm ? tm->end(st) : SPG_OK; };
    const uint32_t cap = b->paths.capacity;
    TraceParams p{};
    p.paths = b->paths;
    p.accum = (float *)b->accum;
    p.spp = cfg.samplesPerPixel;
    p.bounces = cfg.bounces;
    p.clampValue = cfg.radianceClamp;
    p.estimator = cfg.estimator;
    p.rngMode = cfg.rngMode;
    p.seed = cfg.seed;
    p.tx = cfg.transcendentals;
    p.collect = cfg.collectMetrics;
    p.refill = b->refill;
    p.scenes = b->dScenes;
    p.cams = b->dCams;
    p.images = b->dImages;
    Persist ps = b->dev->persist;
    ps.maxBlocks = b->maxBlocks;
    p.work = ps.counter;
    const bool guided =
        cfg.estimator == SPG_EST_GUIDED || cfg.estimator == SPG_EST_GUIDED_NEE;
    const bool nee = cfg.estimator == SPG_EST_NEE || cfg.estimator == SPG_EST_GUIDED_NEE;
    const bool train = guided && cfg.guideTrain;
    if (guided)
    {
        p.guide = b->dg;
        p.guide.train = train ? 1u : 0u;
    }
    p.learnLight = cfg.estimator == SPG_EST_GUIDED_NEE && train ? 1u : 0u;
    b->learnLight = p.learnLight != 0;
    uint32_t k = 0;
    HIPCHK(launch_clear_counters(st, (uint32_t *)b->paths.metrics, 32));
    ++k;
    const bool pcg = cfg.rngMode == SPG_RNG_PCG;
    const uint32_t spp = cfg.samplesPerPixel;
    for (uint64_t s0 = 0; s0 < b->slots; s0 += cap)
    {
        uint32_t nslots = (uint32_t)std::min<uint64_t>(cap, b->slots - s0);
        p.slotPixel = b->dSlotPixel + s0;
        p.slotItem = b->dSlotItem + s0;
        p.slots = nslots;
        HIPCHK(launch_clear_floats(st, p.accum, 3 * (size_t)nslots));
        ++k;
        uint32_t sppPass = pcg ? std::max(1u, std::min(spp, cap / nslots)) : 1u;
        for (uint32_t done = 0; done < spp;)
        {
            uint32_t sp = std::min(sppPass, spp - done);
            p.sppPass = sp;
            p.sampleBase = cfg.sampleOffset + done;
            if (guided)
            {
                // the same pass salt as an eager trace of these slots
                CHK(guide_begin_pass(b->guide, st, p.guide,
                                     (p.sampleBase + 1u) * 0x9E3779B9u ^ cfg.seed ^
                                         (uint32_t)s0,
                                     nslots * sp));
                k += train ? 1 : 0;
            }
            CHK(T0(kRaygen));
            HIPCHK(launch_raygen(st, p));
            CHK(T1());
            ++k;
            int q = 0;
            for (uint32_t bb = 0; bb < cfg.bounces; ++bb)
            {
                CHK(T0(kExtend));
                HIPCHK(launch_extend(st, p, q, ps));
                CHK(T1());
                ++k;
                if (guided && bb + 1 < cfg.bounces)  // the last bounce scatters nothing
                {
                    CHK(T0(kGuide));
                    HIPCHK(launch_guide_infer(st, p, q));
                    CHK(T1());
                    ++k;
                }
                CHK(T0(kShade));
                HIPCHK(launch_shade(st, p, q));
                CHK(T1());
                ++k;
                if (nee && bb + 1 < cfg.bounces)
                {
                    // shadow rays of this bounce, then their contributions
                    CHK(T0(kExtend));
                    HIPCHK(launch_extend(st, p, 2, ps));
                    CHK(T1());
                    CHK(T0(kShade));
                    HIPCHK(launch_resolve_shadow(st, p));
                    CHK(T1());
                    k += 2;
                }
                q ^= 1;
            }
            CHK(T0(kFinal));
            HIPCHK(launch_finalize(st, p));
            CHK(T1());
            ++k;
            if (train)
            {
                CHK(T0(kGuide));
                HIPCHK(launch_guide_collect(st, p));
                CHK(guide_train_pass(b->guide, st));
                CHK(T1());
                k += 3 + 7 * b->guide->cfg.stepsPerPass;
            }
            done += sp;
        }
        HIPCHK(launch_write_pixels(st, p, nullptr));
        ++k;
    }
    if (kernels) *kernels = k;
    return SPG_OK;
}

int timed(spg_Batch *b, const spg_TraceConfig &cfg, float *ms)
{
    hipStream_t st = b->dev->stream;
    HIPCHK(hipEventRecord(b->ev[0], st));
    CHK(record(b, st, cfg, nullptr));
    HIPCHK(hipEventRecord(b->ev[1], st));
    HIPCHK(hipEventSynchronize(b->ev[1]));
    HIPCHK(hipEventElapsedTime(ms, b->ev[0], b->ev[1]));
    return SPG_OK;
}

// What the autotuner chose for one configuration: later batches of the
// same configuration on the same device reuse it without timing.
struct TunedShape
{
    uint32_t capacity, maxBlocks, refill;
    float ms;
};

// The configuration a tuned shape is kept for: the device, the items'
// image sizes and scene sizes, and the trace settings that change the
// amount of work per sample.
std::string tuning_key(const spg_Batch *b, const spg_BatchItem *items)
{
    std::string k = std::to_string(b->dev->ordinal) + "|" + std::to_string(b->n);
    for (uint32_t i = 0; i < b->n; ++i)
    {
        const spg_Scene *sc = items[i].scene;
        size_t tris = 0;
        for (const HostMesh &m : sc->meshes) tris += m.idx.size() / 3;
        k += "|" + std::to_string(items[i].camera->width) + "x" +
             std::to_string(items[i].camera->height) + ":" + std::to_string(tris) + ":" +
             std::to_string(sc->objects.size());
    }
    const spg_TraceConfig &c = b->cfg;
    k += "|" + std::to_string(c.samplesPerPixel) + "," + std::to_string(c.bounces) + "," +
         std::to_string(c.estimator) + "," + std::to_string(c.rngMode) + "," +
         std::to_string(c.transcendentals);
    return k;
}

std::mutex gTunedLock;
std::map<std::string, TunedShape> gTuned;

// Candidates: wavefront capacity x resident traversal blocks, each timed on
// one sample per pixel of the whole batch, then the traversal's refill
// threshold (idle lanes a wave waits for before taking new rays) at the
// fastest of those; the fastest is kept, and remembered for the key.
int autotune(spg_Batch *b, uint32_t cus, const std::string &key)
{
    spg_TraceConfig one = b->cfg;
    one.samplesPerPixel = 1;
    one.collectMetrics = 0;
    one.guideTrain = 0;  // timing runs leave the guide as it is
    std::vector<uint32_t> caps;
    for (uint32_t c : {1u << 20, 1u << 21, 1u << 22})
    {
        caps.push_back(c);
        if ((uint64_t)c >= b->slots) break;
    }
    const uint32_t blocks[] = {0u, cus * 16u, cus * 8u};
    float best = 0.0f, ms = 0.0f;
    uint32_t bestCap = caps[0], bestBlocks = 0, bestRefill = 0, tried = 0;
    // warm-up: first launches pay for code loading and cold caches
    CHK(alloc_state(b, caps[0]));
    b->maxBlocks = 0;
    b->refill = 0;
    CHK(timed(b, one, &ms));
    for (uint32_t c : caps)
    {
        CHK(alloc_state(b, c));
        for (uint32_t mb : blocks)
        {
            b->maxBlocks = mb;
            CHK(timed(b, one, &ms));
            ++tried;
            if (tried == 1 || ms < best)
            {
                best = ms;
                bestCap = c;
                bestBlocks = mb;
            }
        }
    }
    CHK(alloc_state(b, bestCap));
    b->maxBlocks = bestBlocks;
    for (uint32_t rf : {32u, 48u})
    {
        b->refill = rf;
        CHK(timed(b, one, &ms));
        ++tried;
        if (ms < best)
        {
            best = ms;
            bestRefill = rf;
        }
    }
    b->refill = bestRefill;
    b->tuning.candidates = tried;
    b->tuning.tunedMsPerSample = best;
    std::lock_guard<std::mutex> lk(gTunedLock);
    gTuned[key] = TunedShape{bestCap, bestBlocks, bestRefill, best};
    return SPG_OK;
}

void release(spg_Batch *b)
{
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    for (hipEvent_t &e : b->ev)
        if (e) hipEventDestroy(e);
    free_dev(b->tables);
    free_dev(b->pathsBlock);
    free_dev(b->accum);
    free_dev(b->guideBlock);
}

}  // namespace

extern "C" {

int spg_batch_create(spg_Device *dev, const spg_BatchItem *items, uint32_t n,
                     const spg_TraceConfig *cfg, int tune, spg_Batch **out)
{
    if (!dev || !items || !cfg || !out || n == 0)
        return fail(SPG_E_INVALID, "spg_batch_create: NULL or empty argument");
    const spg_TraceConfig &c = *cfg;
    if (c.samplesPerPixel == 0 || c.bounces == 0 || c.bounces > kMaxBounces)
        return fail(SPG_E_INVALID, "spg_batch_create: samplesPerPixel > 0, bounces 1..16");
    if (c.estimator != SPG_EST_PARITY && c.estimator != SPG_EST_UNBIASED &&
        c.estimator != SPG_EST_NEE && c.estimator != SPG_EST_GUIDED &&
        c.estimator != SPG_EST_GUIDED_NEE)
        return fail(SPG_E_INVALID, "spg_batch_create: unknown estimator");
    const bool guided = c.estimator == SPG_EST_GUIDED || c.estimator == SPG_EST_GUIDED_NEE;
    if (guided && (!c.guide || c.guide->dev != dev))
        return fail(SPG_E_INVALID,
                    "spg_batch_create: the guided estimator needs a guide on the device");
    if (guided)
        for (uint32_t k = 1; k < n; ++k)
            if (items[k].scene != items[0].scene)
                return fail(SPG_E_INVALID, "spg_batch_create: with the guided estimator "
                                           "every item traces the same scene (one "
                                           "guide learns one scene)");
    if (c.rngMode != SPG_RNG_XORSHIFT_PIXEL && c.rngMode != SPG_RNG_PCG)
        return fail(SPG_E_INVALID, "spg_batch_create: unknown rngMode");
    if (c.transcendentals != SPG_TX_DOUBLE && c.transcendentals != SPG_TX_FAST)
        return fail(SPG_E_INVALID, "spg_batch_create: unknown transcendentals mode");
    uint64_t slots = 0;
    for (uint32_t k = 0; k < n; ++k)
    {
        const spg_BatchItem &it = items[k];
        if (!it.camera || !it.scene || !it.deviceImage)
            return fail(SPG_E_INVALID, "spg_batch_create: item " + std::to_string(k) +
                                           " has a NULL camera, scene or image");
        if (it.scene->dev != dev)
            return fail(SPG_E_INVALID, "spg_batch_create: item " + std::to_string(k) +
                                           " is on another device");
        if (!it.scene->built)
            return fail(SPG_E_STATE, "spg_batch_create: item " + std::to_string(k) +
                                         " scene not built");
        if (it.camera->width == 0 || it.camera->height == 0)
            return fail(SPG_E_INVALID, "spg_batch_create: camera not configured");
        slots += (uint64_t)it.camera->width * it.camera->height;
    }
    if (slots >= 0xFFFFFFFFull)
        return fail(SPG_E_CAPACITY, "spg_batch_create: more than 2^32 pixels");

    std::lock_guard<std::mutex> lk(dev->lock);
    if (dev->destroyRequested)
        return fail(SPG_E_STATE, "spg_batch_create: device is being destroyed");
    HIPCHK(hipSetDevice(dev->ordinal));
    std::vector<DevScene> ds(n);
    std::vector<DevCamera> dc(n);
    std::vector<float *> im(n);
    for (uint32_t k = 0; k < n; ++k)
    {
        CHK(scene_sync(items[k].scene));
        ds[k] = items[k].scene->dev_scene;
        dc[k] = dev_camera(*items[k].camera);
        im[k] = items[k].deviceImage;
    }
    std::vector<uint32_t> slotItem(slots), slotPixel(slots);
    uint64_t at = 0;
    for (uint32_t k = 0; k < n; ++k)
    {
        uint32_t px = items[k].camera->width * items[k].camera->height;
        for (uint32_t i = 0; i < px; ++i, ++at)
        {
            slotItem[at] = k;
            slotPixel[at] = i;
        }
    }

    spg_Batch *b = new spg_Batch();
    b->dev = dev;
    b->cfg = c;
    b->guide = guided ? c.guide : nullptr;
    b->n = n;
    b->slots = slots;
    for (uint32_t k = 0; k < n; ++k) b->scenes.push_back(items[k].scene);
    int r = SPG_OK;
    auto step = [&](hipError_t e, const char *what) {
        if (!r) r = hip_check(e, what);
    };
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 255) / 256 * 256;
        return o;
    };
    size_t oS = take(sizeof(DevScene) * n), oC = take(sizeof(DevCamera) * n),
           oI = take(sizeof(float *) * n), oSI = take(4 * slots), oSP = take(4 * slots);
    step(hipMalloc(&b->tables, off), "hipMalloc(batch tables)");
    if (!r)
    {
        char *t = (char *)b->tables;
        b->dScenes = (const DevScene *)(t + oS);
        b->dCams = (const DevCamera *)(t + oC);
        b->dImages = (float *const *)(t + oI);
        b->dSlotItem = (const uint32_t *)(t + oSI);
        b->dSlotPixel = (const uint32_t *)(t + oSP);
        step(hipMemcpy(t + oS, ds.data(), sizeof(DevScene) * n, hipMemcpyHostToDevice),
             "hipMemcpy");
        step(hipMemcpy(t + oC, dc.data(), sizeof(DevCamera) * n, hipMemcpyHostToDevice),
             "hipMemcpy");
        step(hipMemcpy(t + oI, im.data(), sizeof(float *) * n, hipMemcpyHostToDevice),
             "hipMemcpy");
        step(hipMemcpy(t + oSI, slotItem.data(), 4 * slots, hipMemcpyHostToDevice),
             "hipMemcpy");
        step(hipMemcpy(t + oSP, slotPixel.data(), 4 * slots, hipMemcpyHostToDevice),
             "hipMemcpy");
    }
    for (hipEvent_t &e : b->ev) step(hipEventCreate(&e), "hipEventCreate");
    int cus = 0;
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev->ordinal);
    const std::string key = tune ? tuning_key(b, items) : std::string();
    bool cached = false;
    if (!r && tune)
    {
        TunedShape shape{};
        {
            std::lock_guard<std::mutex> tl(gTunedLock);
            auto it = gTuned.find(key);
            if (it != gTuned.end())
            {
                shape = it->second;
                cached = true;
            }
        }
        if (cached)
        {
            // the shape tuned for this configuration before; one eager
            // sample per pixel loads the kernels and sizes the persistent
            // grids outside the capture
            r = alloc_state(b, shape.capacity);
            b->maxBlocks = shape.maxBlocks;
            b->refill = shape.refill;
            spg_TraceConfig one = c;
            one.samplesPerPixel = 1;
            one.collectMetrics = 0;
            one.guideTrain = 0;
            float ms = 0.0f;
            if (!r) r = timed(b, one, &ms);
            b->tuning.candidates = 0;
            b->tuning.tunedMsPerSample = shape.ms;
        }
        else
            r = autotune(b, (uint32_t)std::max(cus, 1), key);
    }
    if (!r && !tune)
    {
        uint64_t want = slots * (c.rngMode == SPG_RNG_PCG ? c.samplesPerPixel : 1u);
        r = alloc_state(b, (uint32_t)std::min<uint64_t>(std::max<uint64_t>(want, 1),
                                                        kMaxPathsInFlight));
        // one eager sample per pixel: loads the kernels and sizes the
        // persistent grids outside the capture
        spg_TraceConfig one = c;
        one.samplesPerPixel = 1;
        one.guideTrain = 0;  // leaves the guide as it is
        float ms = 0.0f;
        if (!r) r = timed(b, one, &ms);
        b->tuning.candidates = 1;
        b->tuning.tunedMsPerSample = ms;
    }
    uint32_t kernels = 0;
    if (!r)
    {
        // capture the full render
        hipStream_t st = dev->stream;
        r = hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal),
                      "hipStreamBeginCapture");
        if (!r)
        {
            int rr = record(b, st, c, &kernels);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(st, &g);
            r = rr ? rr : hip_check(e, "hipStreamEndCapture");
            if (g) b->graph = g;
        }
        if (!r)
            r = hip_check(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0),
                          "hipGraphInstantiate");
    }
    if (r)
    {
        release(b);
        delete b;
        return r;
    }
    b->tuning.pathsInFlight = b->paths.capacity;
    b->tuning.extendBlocks = b->maxBlocks;
    b->tuning.refill = b->refill;
    b->tuning.cached = cached ? 1u : 0u;
    b->tuning.pixels = slots;
    b->tuning.graphNodes = kernels;
    if (b->guide) b->guide->pins++;
    dev->liveScenes++;
    *out = b;
    return SPG_OK;
}

int spg_batch_run(spg_Batch *b, spg_Metrics *metrics)
{
    if (!b) return fail(SPG_E_INVALID, "spg_batch_run: NULL");
    spg_Device *d = b->dev;
    std::lock_guard<std::mutex> lk(d->lock);
    HIPCHK(hipSetDevice(d->ordinal));
    for (spg_Scene *s : b->scenes)
        if (s->dirty)
            return fail(SPG_E_STATE, "spg_batch_run: a scene changed after the batch was made");
    auto t0 = std::chrono::steady_clock::now();
    hipStream_t st = d->stream;
    if (b->cfg.estimator == SPG_EST_GUIDED_NEE && b->guide)
    {
        // the learned light split, written into the items' device scenes
        // (every item traces the same scene) when it changed; the graph's
        // kernels read it from there
        const DevScene &S0 = b->scenes[0]->dev_scene;
        if (S0.envTex >= 0 && S0.emitCount > 0)
        {
            const float split = guide_light_split(b->guide, S0.envSelect);
            if (split != b->splitOnDevice)
            {
                char *base = (char *)const_cast<DevScene *>(b->dScenes);
                for (uint32_t k = 0; k < b->n; ++k)
                    HIPCHK(hipMemcpyAsync(base + k * sizeof(DevScene) +
                                              offsetof(DevScene, envSelect),
                                          &split, sizeof(float), hipMemcpyHostToDevice, st));
                HIPCHK(hipStreamSynchronize(st));  // `split` is on this stack
                b->splitOnDevice = split;
            }
        }
    }
    HIPCHK(hipEventRecord(b->ev[0], st));
    HIPCHK(hipGraphLaunch(b->exec, st));
    HIPCHK(hipEventRecord(b->ev[1], st));
    unsigned long long m[16] = {0};
    HIPCHK(hipMemcpyAsync(m, b->paths.metrics, 128, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    auto t1 = std::chrono::steady_clock::now();
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
    b->lastMs = ms;
    for (int k = 0; k < 8; ++k) d->lastCounters[k] = m[k];
    if (b->learnLight)
    {
        b->guide->lightStat[0] += m[8];
        b->guide->lightStat[1] += m[9];
    }
    if (metrics)
    {
        metrics->values[SPG_METRIC_ELAPSED_NS] += (uint64_t)
            std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0).count();
        metrics->values[SPG_METRIC_PATHS_TRACED] += b->slots * b->cfg.samplesPerPixel;
        metrics->values[SPG_METRIC_RAYS_TRACED] += m[0];
        metrics->values[SPG_METRIC_RAY_HIT_COUNT] += m[1];
        metrics->values[SPG_METRIC_RAY_MISS_COUNT] += m[2];
        metrics->values[SPG_METRIC_MIDPHASE_AABB_TESTS] += m[3];
        metrics->values[SPG_METRIC_MESH_TESTS] += m[4];
    }
    return SPG_OK;
}

int spg_batch_tuning(const spg_Batch *b, spg_BatchTuning *out)
{
    if (!b || !out) return fail(SPG_E_INVALID, "spg_batch_tuning: NULL");
    *out = b->tuning;
    return SPG_OK;
}

int spg_batch_last_ms(const spg_Batch *b, double *ms)
{
    if (!b || !ms) return fail(SPG_E_INVALID, "spg_batch_last_ms: NULL");
    *ms = b->lastMs;
    return SPG_OK;
}

int spg_batch_profile(spg_Batch *b, int collect, double *ms6, uint32_t *launches6,
                      uint64_t *counters8)
{
    if (!b) return fail(SPG_E_INVALID, "spg_batch_profile: NULL");
    spg_Device *d = b->dev;
    std::lock_guard<std::mutex> lk(d->lock);
    HIPCHK(hipSetDevice(d->ordinal));
    spg_TraceConfig c = b->cfg;
    c.collectMetrics = collect ? 1u : 0u;
    ClassTimer tm{d};
    auto t0 = std::chrono::steady_clock::now();
    CHK(record(b, d->stream, c, nullptr, &tm));
    CHK(tm.drain());
    unsigned long long m[16] = {0};
    HIPCHK(hipMemcpyAsync(m, b->paths.metrics, 128, hipMemcpyDeviceToHost, d->stream));
    HIPCHK(hipStreamSynchronize(d->stream));
    auto t1 = std::chrono::steady_clock::now();
    tm.ms[5] = std::chrono::duration<double, std::milli>(t1 - t0).count();
    for (int k = 0; k < 6; ++k)
    {
        if (ms6) ms6[k] = tm.ms[k];
        if (launches6) launches6[k] = tm.launches[k];
    }
    if (counters8)
    {
        for (int k = 0; k < 8; ++k) counters8[k] = m[k];
    }
    return SPG_OK;
}

int spg_batch_destroy(spg_Batch *b)
{
    if (!b) return SPG_OK;
    spg_Device *d = b->dev;
    bool last;
    {
        std::lock_guard<std::mutex> lk(d->lock);
        hipSetDevice(d->ordinal);
        hipStreamSynchronize(d->stream);
        if (b->guide) b->guide->pins--;
        release(b);
        delete b;
        d->liveScenes--;
        last = d->destroyRequested && d->liveScenes == 0;
    }
    if (last) spg_device_destroy(d);
    return SPG_OK;
}

}  // extern "C"
