// Internal structures shared by the kernels (spg_trace.hip) and the C-ABI
// implementation (spg_api.cpp).  Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <stddef.h>
#include <stdint.h>

#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/spgpu.h"
#include "spg_bvh.h"
#include "spg_error.h"
#include "spg_guide.h"
#include "spg_math.h"

namespace spg {

// Traversal stack depth per lane (the top entries in LDS, the rest in a
// global spill area).  The builder bounds the binary trees' depths; a
// four-wide node collapses at least one binary level on every path through
// it and leaves at most three siblings pending, so a walk holds at most
// 3 * (kMaxTlasDepth + 1 + kMaxBlasDepth + 1) references.  Scene builds
// also check the exact bound of their trees.
constexpr uint32_t kMaxTlasDepth = 20;
constexpr uint32_t kMaxBlasDepth = 42;
constexpr int kStackDepth = 3 * (kMaxTlasDepth + 1 + kMaxBlasDepth + 1);
constexpr int kTraceBlock = 64;  // traversal blocks are one wave
// Resident waves per CU a persistent traversal launch may use (sizes the
// global spill area of the traversal stacks).
constexpr uint32_t kMaxPersistWavesPerCU = 32;

// Per-launch work counter and stack spill area of persistent launches.
struct Persist
{
    uint32_t *counter;
    int *spill;
    uint32_t maxBlocks;  // cap on the resident one-wave blocks, 0 = occupancy
};

// Limits of the allocation-free closest-hit entry point: resident one-wave
// blocks of its persistent launch, and the traversal-stack entries per lane
// its scratch holds (LDS and spill together; a mesh whose walk needs more
// is rejected).
constexpr uint32_t kNoAllocWaves = 4096;
constexpr uint32_t kNoAllocStack = 56;

// Relative slack on the t-max cull: a box is skipped only if its entry
// exceeds the best t by more than 2^-12 of it, so rounding differences
// between the slab entry and the Moller-Trumbore t never cull a closer hit.
constexpr float kCullSlack = 1.0f + 1.0f / 4096.0f;

struct DevMesh
{
    const float4 *nodes;    // WideNode, 4 float4 each
    const float4 *tris;     // triangle records (4 float4 each) in leaf order,
                            // right after the nodes in the same allocation
    const float *verts;     // VertexPNT, 8 floats
    const uint32_t *idx;    // 3 per triangle
    const float4 *shade;    // per triangle (by index), 4 float4: its three
                            // vertex normals (smooth; flat: cross(b-a, c-a))
                            // and uvs, what shading reads at a hit
    uint32_t smooth;
    uint32_t triCount;
};

// 128 bytes.  The traversal reads it in 64-byte pieces: bytes 0..63 when
// leaving the instance (model rows), bytes 48..111 when entering it (inverse
// rows, cull scale, the mesh's records).
struct DevInstance
{
    m34 model;          // rows 0..2 of M = T*R*S       (src/sp_scene.cpp:91)
    m34 inv;            // rows 0..2 of S^-1*R(q*)*T(-p) (src/sp_scene.cpp:94-96)
    float cullScale;    // 1 / min |scale|: world t bound -> local t bound
    uint32_t nodeCount; // the mesh's wide node records (triangle k is nodeCount + k)
    const float4 *nodes;// the mesh's 64-byte records
    uint32_t mesh;
    int32_t material;   // dense material slot, -1 = not registered
    uint32_t materialId;
    int32_t orient;     // sign of det of M's 3x3 (0: degenerate): +1 keeps the
                        // mesh's winding in world space, -1 mirrors it
};
static_assert(sizeof(DevInstance) == 128, "DevInstance layout");
static_assert(offsetof(DevInstance, cullScale) == 96, "DevInstance layout");
static_assert(offsetof(DevInstance, nodes) == 104, "DevInstance layout");

struct DevMaterial
{
    float albedo[3];
    int32_t albedoTex;   // dense texture slot or -1
    float emission[3];
    int32_t emissionTex; // dense texture slot or -1
    float roughness;
    float pad[3];
};

struct DevTexture
{
    uint64_t offset;  // in float4 texels
    uint32_t width, height;
    uint32_t isCube;
    uint32_t pad;
};

struct DevScene
{
    const float4 *tlasNodes;    // WideNode; leaf references hold the instance index
    const DevInstance *inst;
    const DevMesh *meshes;
    const DevMaterial *mats;
    const DevTexture *texs;
    const float4 *texels;
    int32_t background;         // dense material slot, -1 = none
    uint32_t instCount;
    // importance sampling of the background's cube-map emission (next-event
    // estimation); envTex < 0: no table, the environment is not sampled
    const float *envPdf;        // per texel: P_i / area_i
    const float *envProb;       // alias table
    const uint32_t *envAlias;
    int32_t envTex;             // texture slot of the cube map
    uint32_t envCount;          // 6 * w * h
    // emissive triangles (next-event estimation toward emitting geometry);
    // emitCount 0: none
    const float *emitPdfA;      // per entry: P_g / world area
    const float *emitProb;      // alias table
    const uint32_t *emitAlias;
    const uint32_t *emitPair;   // [2 * g]: instance, triangle
    const int32_t *emitBase;    // per instance: first entry or -1
    uint32_t emitCount;
    // the spatial emitter choice (emitInstCount > 0): per emitter instance
    // (centre, radius) and (power, 1 / area, first entry, triangle count);
    // per instance its emitter index or -1; per entry an alias table over
    // its instance's triangles by area (local indices)
    const float4 *emitInst;
    const int32_t *emitSlot;
    const float *emitProbL;
    const uint32_t *emitAliasL;
    uint32_t emitInstCount;
    float envSelect;            // probability of sampling the environment
};

struct DevCamera
{
    v3 right, up, forward, position, filmCenter;
    uint32_t width, height;
    float halfPixelWidth, halfPixelHeight, halfFilmWidth, halfFilmHeight;
};

// Wavefront state for `capacity` paths in flight (structure of arrays).
struct DevPaths
{
    uint32_t capacity, bounces;
    uint32_t *pathSlot;          // pixel slot of each path
    uint32_t *pathSample;        // sample index of each path
    uint32_t *rngXs;             // xorshift state per path
    uint64_t *rngPcg;            // PCG state per path
    uint32_t *slotRng;           // xorshift state carried per pixel slot
    float *vtx;                  // [bounces*7][capacity]: E.xyz X.xyz cos
    uint32_t *vcount;            // vertices per path
    float *beta;                 // [3][capacity]
    float *radiance;             // [3][capacity]
    float *qo[3], *qd[3];        // [3][capacity] ray origins / directions
                                 // (queues 0, 1: path rays; 2: shadow rays)
    uint32_t *qpath[3];          // path index of each queued ray
    uint32_t *counters;          // [0..2] queue lengths
    float *lastPdf;              // density of each path's last sampled direction
    float *lastN;                // [3][capacity] shading normal where it was sampled
    float *lastW;                // the spatial emitter choice's weight sum there
    float *nee;                  // [3][capacity] pending shadow-ray contribution
    uint32_t *neeTarget;         // [2][capacity] instance, triangle the shadow
                                 // ray must reach (instance SPG_MISS: the sky)
    float *neeDist;              // its distance bound (emitter targets)
    unsigned long long *metrics; // rays, hits, misses, box, mesh, tri tests
    float *ht, *hu, *hv;         // hit records per queued ray
    uint32_t *hinst, *hprim;
};

enum KernelClass
{
    kRaygen = 0,
    kExtend = 1,
    kShade = 2,
    kFinal = 3,
    kGuide = 4,
    kTimingClasses = 5
};

struct TraceParams
{
    DevScene scene;
    DevCamera cam;
    DevPaths paths;
    const uint32_t *slotPixel;   // pixel index of each slot
    float *accum;                // [3][slots] running sums (sample order)
    uint32_t slots;
    uint32_t sppPass;            // samples of each slot in this pass
    uint32_t sampleBase;         // first sample index of this pass
    uint32_t spp;                // samples per pixel of the whole render
    uint32_t bounces;
    float clampValue;
    uint32_t estimator, rngMode, seed, tx, collect;
    uint32_t refill;             // k_extend's refill threshold, 0 = default
    // learned light split: k_resolve_shadow adds the luminance of the
    // reached light samples, sky and emitters apart, into metrics[8], [9]
    // (fixed point, 2^-16 units)
    uint32_t learnLight;
    // the persistent closest-hit launch's work counter: zeroed by the
    // kernel before each k_extend (k_raygen, k_shade, k_resolve_shadow),
    // so the launch sequence needs no separate clearing launches
    uint32_t *work;
    DevGuide guide;              // SPG_EST_GUIDED(_NEE) only
    // batched launches (slotItem != nullptr): each slot belongs to an item
    // with its own scene, camera and image
    const uint32_t *slotItem;
    const DevScene *scenes;
    const DevCamera *cams;
    float *const *images;
};

// Loss scale of the FP16 gradients in guide training.
constexpr float kLossScale = 1.0f / 32.0f;

// One guide training step's buffers (device pointers).
struct GuideTrainArgs
{
    _Float16 *w16, *wt16;        // [kGRows][64], [5][64][64] (in-major)
    float *bias;                 // [kGRows]
    float *params, *m, *v, *grad;// [kGParams] master, Adam moments, gradient
    const float *samples;        // [kGSampleFloats][sampleCap]
    const uint32_t *count;       // samples available
    uint32_t sampleCap;
    uint32_t first, batch;       // this step's samples [first, first+batch)
    float *wsum;                 // batch weight sum (normalize) or nullptr
    float wmax;                  // clamp of normalized weights, 0 = off
    _Float16 *H, *G;             // [5][64][batch] layer inputs / out grads
    float *part;                 // [batch / 512][kGParams] partial gradients
    float *loss;                 // += sum_j w_j (-log q_j)
};

// Adam hyperparameters; the step count t lives on the device (advanced by
// the step's first launch, read by the update), so a captured training
// sequence replays with the right bias corrections 1 - beta^t.
struct GuideAdam
{
    float lr, beta1, beta2, eps;
    uint32_t *step;
    uint32_t apply;
};

}  // namespace spg

// HIP events per device: timed intervals of a trace use [0, kEvents - 4),
// the trace total the last two, a closest-hit timing the two before; the
// post-step timing uses kEvents - 4 .. kEvents - 2 (each call reads its
// events before returning, under the device lock).
constexpr int kEvents = 256;

// One caller's host-pixel trace waiting to be combined with other callers'
// (spg_path_trace_tile / spg_path_trace_tiles): compatible requests -- the
// same scene, camera, trace options and host image -- are traced as one
// wavefront by whichever caller finds the device idle.
struct TileRequest
{
    const spg_Context *ctx;
    const spg_Tile *tiles;
    uint32_t count;
    spg_Metrics *metrics;
    float *hostPixels;
    std::chrono::steady_clock::time_point start;
    bool done = false;
    int status = 0;
    std::string error;
};

struct spg_Device
{
    int ordinal = 0;
    // host-pixel trace requests of concurrent callers (guarded by qlock,
    // never held while tracing)
    std::mutex qlock;
    std::condition_variable qcv;
    std::deque<TileRequest *> queue;
    bool leaderActive = false;
    size_t lastGroup = 0;        // requests the last combined trace took
    int liveScenes = 0;          // guarded by lock
    bool destroyRequested = false;
    hipStream_t stream = nullptr;
    std::mutex lock;
    spg::DevPaths paths{};
    size_t pathsBytes = 0;
    void *pathsBlock = nullptr;
    void *image = nullptr;       // host-pixel calls: the tiles' pixels in
    size_t imageCount = 0;       // slot order (RGBA32F per slot)
    void *hostStage = nullptr;   // pinned host copy of `image`
    size_t hostStageBytes = 0;
    void *slotPixel = nullptr;   // pixel index per slot
    size_t slotPixelCount = 0;
    // the slot list of the last trace call (pixels of its tiles in tile
    // order) and whether slotPixel already holds all of it
    std::vector<spg_Tile> slotTiles;
    uint32_t slotW = 0, slotH = 0;
    std::vector<uint32_t> slots;
    bool slotsOnDevice = false;
    void *accum = nullptr;       // per-slot radiance sums
    size_t accumCount = 0;
    void *persistBlock = nullptr; // launch counters + traversal stack spill
    void *noallocScratch = nullptr; // device address of the static scratch of
                                    // spg_intersect_closest_device_noalloc
    spg::Persist persist{};
    hipEvent_t ev[kEvents];
    bool eventsMade = false;
    double lastMs[6] = {0};
    uint32_t lastLaunches[6] = {0};
    uint64_t lastCounters[8] = {0};
};

struct spg_Guide
{
    spg_Device *dev = nullptr;
    spg_GuideConfig cfg{};
    // parameters and training state (one allocation)
    void *block = nullptr;
    float *params = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
    float *bias = nullptr, *wsum = nullptr, *loss = nullptr;
    _Float16 *w16 = nullptr, *wt16 = nullptr;
    uint32_t *count = nullptr;        // training samples collected
    uint32_t *stepDev = nullptr;      // Adam steps taken (device)
    // training samples, SoA [13][sampleCap]
    void *samplesBlock = nullptr;
    float *samples = nullptr;
    uint32_t sampleCap = 0;
    // per-layer activations / gradients of a batch, [5][64][batchCap] each
    void *hgBlock = nullptr;
    _Float16 *H = nullptr, *G = nullptr;
    float *part = nullptr;            // per-chunk partial gradients
    uint32_t batchCap = 0;
    // per-path buffers of the tracer (lobes per queued ray, vertex records)
    void *pathBlock = nullptr;
    float *lobes = nullptr, *gvtx = nullptr;
    uint32_t *gcnt = nullptr, *scanBase = nullptr;  // training-sample scan
    uint32_t pathCap = 0, pathBounces = 0;
    uint32_t pins = 0;                // live batches capturing this guide's buffers
    // learned light split of SPG_EST_GUIDED_NEE: luminance of the reached
    // light samples of its training passes, sky and emitters (2^-16 units)
    uint64_t lightStat[2] = {0, 0};
    // verification hooks (spg_guide_debug)
    uint32_t analytic = 0;
    void *recBlock = nullptr;         // count (256 B), then the records
    uint32_t recCap = 0;
};

struct HostMesh
{
    std::vector<float> verts;     // 8 per vertex
    std::vector<uint32_t> idx;
    int smooth = 0;
    spg::BvhBuildResult bvh;
    std::vector<spg::WideNode> wide;  // the traversal's four-wide form of bvh
    uint32_t widePending = 0;         // its walk's stack bound
    bool built = false;
    // device copies
    // one allocation of 64-byte records: nodes, then triangles in leaf
    // order (dTris points into it)
    void *dNodes = nullptr, *dTris = nullptr, *dVerts = nullptr, *dIdx = nullptr,
         *dShade = nullptr;
    uint32_t nodeCount = 0;
    float boundsLo[3], boundsHi[3];
};

struct HostObject
{
    uint32_t mesh, materialId;
    float pos[3], quat[4], scale[3];
    float model[16], inv[16];
    float aabbLo[3], aabbHi[3];   // TransformAabb of the mesh bounds
    float tightLo[3], tightHi[3]; // bounds of the transformed vertices (instance tree)
};

struct HostTexture
{
    uint32_t id, width, height, isCube;
    std::vector<float> rgba;
};

struct spg_Scene
{
    spg_Device *dev = nullptr;
    std::vector<HostMesh> meshes;
    std::vector<HostObject> objects;
    std::vector<uint32_t> materialKeys;
    std::vector<spg_Material> materials;
    std::vector<HostTexture> textures;
    uint32_t backgroundId = 0;
    bool built = false;
    bool dirty = true;
    // learned light split of SPG_EST_NEE (spg_scene_learn_light_split): the
    // luminance its calls' reached light samples delivered, sky and
    // emitters (2^-16 units)
    bool learnLight = false;
    uint64_t lightStat[2] = {0, 0};
    spg::BvhBuildResult tlas;
    uint32_t tlasPending = 0;       // stack bound of the wide instance tree
    // device-side tables
    void *dEnv = nullptr;           // environment table: pdf, prob, alias
    void *dEmit = nullptr;          // emitter table: pdfA, prob, alias, pair, base
    void *dTlasNodes = nullptr, *dInst = nullptr,
         *dMeshes = nullptr, *dMats = nullptr, *dTexs = nullptr,
         *dTexels = nullptr;
    spg::DevScene dev_scene{};
};

#define HIPCHK(x) CHK(spg::hip_check((x), #x))

namespace spg {
int alloc_path_state(DevPaths &p, void *&block, size_t &bytes, uint32_t capacity,
                     uint32_t bounces);
DevCamera dev_camera(const spg_Camera &c);
// Device tables of a built scene up to date (call with the device locked).
int scene_sync(spg_Scene *s);
constexpr uint32_t kMaxPathsInFlight = 1u << 21;
constexpr uint32_t kMaxBounces = 16;

// guide hooks of the tracer (spg_guide_api.cpp)
// A learned sky share of light samples from the delivered luminance sums
// (sky, emitters): the sky's part, clamped to [0.05, 0.95]; `fallback`
// before any was delivered.  guide_light_split: SPG_EST_GUIDED_NEE's, kept
// by the guide.
float light_split(const uint64_t stat[2], float fallback);
float guide_light_split(const spg_Guide *g, float fallback);
int guide_prepare_trace(spg_Guide *g, const spg_Scene *s, uint32_t capacity,
                        uint32_t bounces, bool train, DevGuide &out,
                        void **ownBlock = nullptr);
int guide_begin_pass(spg_Guide *g, hipStream_t st, DevGuide &dg, uint32_t salt,
                     uint32_t paths);
int guide_train_pass(spg_Guide *g, hipStream_t st);

// error reporting (thread-local message, spg_error.h); a HIP failure as one
int hip_check(hipError_t e, const char *what);

// kernel launchers (spg_trace.hip)
hipError_t launch_closest_mesh(hipStream_t s, const DevMesh *mesh,
                               const float *ox, const float *oy,
                               const float *oz, const float *dx,
                               const float *dy, const float *dz, uint32_t n,
                               uint32_t *prim, float *t, float *u, float *v,
                               unsigned long long *stats, const Persist &ps);
// device address of the allocation-free entry point's scratch (call with
// the device current)
hipError_t noalloc_scratch_address(void **ptr);
hipError_t launch_closest_mesh_noalloc(hipStream_t s, const DevMesh *mesh,
                                       const float *ox, const float *oy,
                                       const float *oz, const float *dx,
                                       const float *dy, const float *dz, uint32_t n,
                                       uint32_t *prim, float *t, float *u, float *v,
                                       void *scratchPtr);
hipError_t launch_intersect_scene(hipStream_t s, const DevScene &scene,
                                  const float *rays6, uint32_t n, float *out7,
                                  uint32_t *object, uint32_t *prim,
                                  const Persist &ps);
size_t spill_ints();
hipError_t launch_raygen(hipStream_t s, const TraceParams &p);
hipError_t launch_extend(hipStream_t s, const TraceParams &p, int q,
                         const Persist &ps);
hipError_t launch_shade(hipStream_t s, const TraceParams &p, int q);
hipError_t launch_resolve_shadow(hipStream_t s, const TraceParams &p);
hipError_t launch_finalize(hipStream_t s, const TraceParams &p);
// guided training: the pass's samples in a fixed order (after finalize)
hipError_t launch_guide_collect(hipStream_t s, const TraceParams &p);
hipError_t launch_write_pixels(hipStream_t s, const TraceParams &p,
                               float *image);
// the pass's pixels in slot order: out[slot] (single-scene launches)
hipError_t launch_write_slots(hipStream_t s, const TraceParams &p, float *out);
hipError_t launch_clear_counters(hipStream_t s, uint32_t *counters, int n);
hipError_t launch_set_counter(hipStream_t s, uint32_t *counter, uint32_t v);
hipError_t launch_clear_floats(hipStream_t s, float *p, size_t n);
hipError_t launch_guide_infer(hipStream_t s, const TraceParams &p, int q);
hipError_t launch_accumulate(hipStream_t s, float4 *dst, const float4 *src,
                             uint32_t pixels, uint32_t before, uint32_t added);
hipError_t launch_tonemap(hipStream_t s, const float4 *src, uint32_t pixels,
                          float exposure, bool dbl, uchar4 *out);
hipError_t launch_irradiance(hipStream_t s, const float4 *src, uint32_t ew,
                             uint32_t eh, uint32_t w, uint32_t h, const float *phis,
                             uint32_t nPhi, const float *thetas, uint32_t nTheta,
                             float clampValue, bool dbl, float4 *out);
// guide network (spg_guide.hip)
hipError_t launch_guide_eval(hipStream_t s, const _Float16 *w16, const float *bias,
                             const float *in9, uint32_t n, float *out48,
                             float *lobes40);
hipError_t launch_guide_convert(hipStream_t s, const GuideTrainArgs &a);
hipError_t launch_guide_gradient(hipStream_t s, const GuideTrainArgs &a,
                                 uint32_t cus);
hipError_t launch_guide_adam(hipStream_t s, const GuideTrainArgs &a,
                             const GuideAdam &h);
// device Adam step count: += 1 before an update
hipError_t launch_guide_step(hipStream_t s, uint32_t *step);
}  // namespace spg
