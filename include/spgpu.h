/*
 * spgpu -- MI355X (gfx950) wavefront path tracer behind a C ABI.
 *
 * A drop-in for the reference's tile-based CPU tracer (deadVertex/
 * vk_cinematic, `sp_` prefix).  Every entry point below names the reference
 * interface it stands in for.  Conventions:
 *
 *   - Every call returns SPG_OK (0) or a negative SPG_E_* code; nothing
 *     aborts.  spg_last_error() returns the calling thread's last message.
 *   - Plain pointers and sizes only.  Host pointers unless the name ends in
 *     _device; device pointers belong to the device passed in and are used on
 *     that device's stream.
 *   - Layouts match the reference: VertexPNT is 8 floats (position, normal,
 *     uv; src/mesh.h:10-15), quaternions are xyzw (src/math_lib.h:99),
 *     pixels are RGBA32F row-major x + y*width (src/simd_path_tracer.h:3-8).
 *   - Data passed in is copied to HBM (the reference borrows the caller's
 *     arrays, src/sp_scene.cpp:8-19).
 *   - Handles may be shared between host threads.  Device work of one
 *     device is serialized by the library, except that concurrent
 *     spg_path_trace_tile / spg_path_trace_tiles calls with the same scene,
 *     camera, options and host image (the reference's worker pool,
 *     src/main.cpp:731-759) are traced together as one wavefront.
 */
#ifndef SPGPU_H
#define SPGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPG_OK 0
#define SPG_E_INVALID -1    /* bad argument */
#define SPG_E_HIP -2        /* a HIP runtime call failed */
#define SPG_E_CAPACITY -3   /* a fixed limit was exceeded */
#define SPG_E_STATE -4      /* call order (e.g. trace before build) */
#define SPG_E_NOT_FOUND -5  /* unknown id */

#define SPG_NO_TEXTURE 0xFFFFFFFFu
#define SPG_MISS 0xFFFFFFFFu

/* ---- enums ---- */
enum spg_RngMode {
    SPG_RNG_XORSHIFT_PIXEL = 1, /* xorshift32 per pixel, seeded by
                                   spg_pixel_seed(seed, y*width+x); samples of
                                   a pixel draw from one stream in order */
    SPG_RNG_PCG = 2             /* PCG32, stream = pixel index, sample s starts
                                   2^20*s draws into the stream */
};
enum spg_Estimator {
    SPG_EST_PARITY = 0,   /* the reference estimator (ComputeRadianceForPath,
                             src/simd_path_tracer.cpp:107-175) */
    SPG_EST_UNBIASED = 1, /* same materials, cosine sampling with 1/pdf */
    SPG_EST_GUIDED = 2,   /* unbiased + learned guiding, one-sample MIS */
    SPG_EST_NEE = 3,      /* unbiased + next-event estimation toward the
                             background's cube-map emission (importance
                             sampled by texel) and toward emissive triangles
                             (by emission x area), one light sample and one
                             shadow ray per vertex, combined with BSDF
                             sampling by the power heuristic; no reference
                             counterpart */
    SPG_EST_GUIDED_NEE = 4 /* the guided estimator plus the same
                             next-event estimation: each light sample is
                             weighted against the one-sample MIS density of
                             BSDF and guide, and emission reached by a
                             scattered ray against the light sampler (three
                             techniques, power heuristic); training passes
                             also learn the sky/emitter split of the light
                             samples (spg_guide_light_split); needs a guide */
};
enum spg_Transcendentals {
    SPG_TX_DOUBLE = 1,    /* sin/cos/atan2/pow evaluated in double and
                             rounded once to float (bit-reproducible) */
    SPG_TX_FAST = 2       /* device float intrinsics (not bit-reproducible) */
};

/* sp_Metrics counters (src/sp_metrics.h:3-41).  Index 0 holds nanoseconds
 * instead of rdtsc cycles; the other per-stage cycle counters stay 0. */
enum {
    SPG_METRIC_ELAPSED_NS = 0,
    SPG_METRIC_PATHS_TRACED = 1,
    SPG_METRIC_RAYS_TRACED = 2,
    SPG_METRIC_RAY_HIT_COUNT = 3,
    SPG_METRIC_RAY_MISS_COUNT = 4,
    SPG_METRIC_MIDPHASE_AABB_TESTS = 10,
    SPG_METRIC_MESH_TESTS = 11,
    SPG_MAX_METRICS = 12
};

/* ---- POD types ---- */
typedef struct spg_Device spg_Device;
typedef struct spg_Scene spg_Scene;
typedef struct spg_Guide spg_Guide;

/* Tile, src/tile.h:3-9 */
typedef struct { uint32_t minX, minY, maxX, maxY; } spg_Tile;

/* sp_Metrics, src/sp_metrics.h:45-48 */
typedef struct { uint64_t values[SPG_MAX_METRICS]; } spg_Metrics;

/* sp_Material, src/sp_material_system.h:3-12 */
typedef struct {
    float albedo[3];
    uint32_t albedoTexture;
    float emission[3];
    uint32_t emissionTexture;
    float roughness;
} spg_Material;

/* sp_Camera, src/simd_path_tracer.h:17-32 (the image plane is implied by
 * width/height; pixels are passed to the trace calls). */
typedef struct {
    float right[3], up[3], forward[3];
    float position[3], filmCenter[3];
    uint32_t width, height;
    float halfPixelWidth, halfPixelHeight;
    float halfFilmWidth, halfFilmHeight;
} spg_Camera;

/* Per-call tracing options (the reference's compile-time switches,
 * src/config.h:25-37, and its hard-coded bounce count,
 * src/simd_path_tracer.cpp:195). */
typedef struct {
    uint32_t samplesPerPixel;  /* SAMPLES_PER_PIXEL */
    uint32_t bounces;          /* rays per path, 3 in the reference */
    float radianceClamp;       /* RADIANCE_CLAMP, 0 = off */
    uint32_t estimator;        /* spg_Estimator */
    uint32_t rngMode;          /* spg_RngMode */
    uint32_t seed;             /* base seed of the per-pixel streams */
    uint32_t transcendentals;  /* spg_Transcendentals */
    uint32_t sampleOffset;     /* first sample index (progressive passes) */
    uint32_t collectMetrics;   /* count AABB/triangle/mesh tests */
    uint32_t guideTrain;       /* SPG_EST_GUIDED(_NEE): collect training
                                  samples and run the guide's training steps
                                  after each pass */
    spg_Guide *guide;          /* SPG_EST_GUIDED(_NEE): the guide
                                  (spg_guide_create) */
} spg_TraceConfig;

/* sp_Context, src/simd_path_tracer.h:34-46 */
typedef struct {
    const spg_Camera *camera;
    spg_Scene *scene;
    spg_TraceConfig config;
} spg_Context;

/* Scene build options (replaces sp_BuildMeshMidphase's greedy 4-ary tree,
 * src/bvh.cpp:51-200, with a binned SAH BVH2). */
typedef struct {
    uint32_t maxLeafSize;   /* triangles per leaf, 1..8 (default 4) */
    uint32_t sahBins;       /* SAH bins per axis (default 16) */
    uint32_t threads;       /* host build threads, 0 = all cores */
} spg_BuildOpts;

/* Counters of one closest-hit launch (no reference counterpart beyond
 * bvh_IntersectRayResult.aabbTestCount, src/bvh.h:20-25). */
typedef struct {
    uint64_t rays, boxTests, triTests, nodeFetches, hits;
} spg_TraversalStats;

/* ---- errors ---- */
const char *spg_last_error(void);

/* ---- lifecycle ---- */
int spg_device_create(int ordinal, spg_Device **out);
int spg_device_destroy(spg_Device *dev);
int spg_device_synchronize(spg_Device *dev);
/* Device memory helpers for callers without their own allocator. */
int spg_device_alloc(spg_Device *dev, size_t bytes, void **out);
int spg_device_free(spg_Device *dev, void *ptr);
int spg_copy_to_device(spg_Device *dev, void *dst, const void *src, size_t bytes);
int spg_copy_to_host(spg_Device *dev, void *dst, const void *src, size_t bytes);

/* ---- scene: sp_InitializeScene / sp_CreateMesh / sp_BuildMeshMidphase /
 *      sp_AddObjectToScene / sp_BuildSceneBroadphase (src/sp_scene.cpp:1-125) */
int spg_scene_create(spg_Device *dev, spg_Scene **out);
int spg_scene_destroy(spg_Scene *scene);
/* sp_CreateMesh(VertexPNT*, u32, u32*, u32, b32) + sp_BuildMeshMidphase:
 * verts8 is nv*8 floats, idx is ni u32 (ni % 3 == 0). */
int spg_mesh_create(spg_Scene *scene, const float *verts8, uint32_t nv,
                    const uint32_t *idx, uint32_t ni, int smoothShading,
                    uint32_t *meshId);
/* sp_AddObjectToScene(scene, mesh, material, position, orientation, scale) */
int spg_scene_add_object(spg_Scene *scene, uint32_t meshId,
                         uint32_t materialId, const float pos[3],
                         const float quat[4], const float scale[3],
                         uint32_t *objectId);
/* sp_BuildSceneBroadphase + the per-mesh midphase builds; uploads to HBM. */
int spg_scene_build(spg_Scene *scene, const spg_BuildOpts *opts);
/* Learned light split of SPG_EST_NEE (no reference counterpart): with
 * enable != 0 each NEE call on this scene adds the luminance its reached
 * light samples delivered, sky and emitters apart, and later calls draw
 * the sky with the learned share (clamped to [0.05, 0.95]) instead of 0.5
 * when the scene has both; enable = 0 stops and forgets it (the default).
 * spg_scene_light_split reads it (-1 before any).  Batches draw with 0.5. */
int spg_scene_learn_light_split(spg_Scene *scene, uint32_t enable);
int spg_scene_light_split(spg_Scene *scene, float *skyFraction);
/* Host-only SAH build of a triangle list (tris9 = a,b,c per triangle), no
 * device involved: nodes16 receives the 64-byte nodes (layout as
 * spg_mesh_bvh_export), order the leaf primitive order. */
int spg_bvh_build_host(const float *tris9, uint32_t n, const spg_BuildOpts *opts,
                       float *nodes16, uint32_t nodeCap, uint32_t *order,
                       uint32_t *nodeCount, uint32_t *maxDepth);
/* The same build collapsed to the traversal's four-wide nodes (64 bytes:
 * origin, per-axis scale exponents, four child references, 8-bit child
 * boxes field-major; layout as spg_mesh_wide_export), leaf ranges into
 * `order`.  maxPending: the walk's stack bound for this tree. */
int spg_bvh_build_wide_host(const float *tris9, uint32_t n,
                            const spg_BuildOpts *opts, uint8_t *nodes64,
                            uint32_t nodeCap, uint32_t *order,
                            uint32_t *nodeCount, uint32_t *maxPending);
/* Node/triangle counts and depth of a built mesh BVH. */
int spg_mesh_bvh_info(spg_Scene *scene, uint32_t meshId, uint32_t *nodes,
                      uint32_t *leaves, uint32_t *maxDepth, double *sahCost);
/* Host copy of a built mesh BVH (for checkers): nodes as 16 floats each
 * (child0 lo.xyz, child0 index bits, child0 hi.xyz, child1 index bits,
 *  child1 lo.xyz, pad, child1 hi.xyz, pad), primitive order, root. */
int spg_mesh_bvh_export(spg_Scene *scene, uint32_t meshId, float *nodes16,
                        uint32_t nodeCap, uint32_t *primOrder,
                        uint32_t primCap, int32_t *root);
/* Host copy of the four-wide hierarchy the traversal walks (for
 * checkers): 64-byte nodes -- origin p (3 floats), per-axis scale
 * exponents e (s = 2^(e-127), 3 bytes, 1 pad), four child references
 * (int32, 0x7FFFFFFF absent, leaves as in spg_mesh_bvh_export), then the
 * child boxes as bytes q[f][c] (field f = lo.x lo.y lo.z hi.x hi.y hi.z of
 * child c) decoding to p + q * s, 8 pad bytes -- and the primitive order.
 * With nodes64 and primOrder NULL only the counts are returned;
 * maxPending is the walk's stack bound for this tree. */
int spg_mesh_wide_export(spg_Scene *scene, uint32_t meshId, uint8_t *nodes64,
                         uint32_t nodeCap, uint32_t *primOrder,
                         uint32_t primCap, uint32_t *nodeCount,
                         uint32_t *primCount, uint32_t *maxPending);

/* ---- materials: sp_RegisterMaterial / sp_RegisterTexture
 *      (src/sp_material_system.cpp:1-57) and backgroundMaterialId
 *      (src/sp_material_system.h:45) ---- */
int spg_material_register(spg_Scene *scene, const spg_Material *material,
                          uint32_t id);
/* 2D RGBA32F texture (albedo by uv; as emission it is an equirect map). */
int spg_texture_register(spg_Scene *scene, const float *rgba, uint32_t width,
                         uint32_t height, uint32_t id);
/* Cube-map RGBA32F texture, 6 faces of width*height in the face order and
 * basis of src/cubemap.cpp:55-103; as emission it is looked up by direction. */
int spg_cubemap_register(spg_Scene *scene, const float *rgba6, uint32_t width,
                         uint32_t height, uint32_t id);
int spg_set_background_material(spg_Scene *scene, uint32_t id);

/* ---- camera: sp_ConfigureCamera (src/simd_path_tracer.cpp:1-36) ---- */
int spg_camera_configure(spg_Camera *camera, uint32_t width, uint32_t height,
                         const float pos[3], const float quat[4],
                         float filmDistance);

/* ---- tiles: ComputeTiles (src/tile.h:11-42) ---- */
int spg_compute_tiles(uint32_t width, uint32_t height, uint32_t tileWidth,
                      uint32_t tileHeight, spg_Tile *tiles, uint32_t maxTiles,
                      uint32_t *count);

/* ---- tracing ---- */
/* sp_PathTraceTile(ctx, tile, rng, metrics), src/simd_path_tracer.cpp:178:
 * overwrites the tile's pixels of the host image (width*height RGBA32F).
 * Safe to call from many threads at once on one context (the reference's
 * WorkerThread, src/main.cpp:744): calls that arrive while a trace runs are
 * traced together by the next one, with the same pixels as one
 * spg_path_trace_tiles call over their tiles (per-pixel random streams).
 * metrics: paths exactly; rays, hits, misses and test counts of a combined
 * trace are shared out in proportion to each call's paths (the shares sum
 * to the combined counts); elapsed time from the call's own start.  A
 * guided call with training on is never combined. */
int spg_path_trace_tile(const spg_Context *ctx, spg_Tile tile,
                        spg_Metrics *metrics, float *hostPixels);
/* The work queue of src/main.cpp:819-844 in one call: all tiles are traced
 * in one wavefront batch. */
int spg_path_trace_tiles(const spg_Context *ctx, const spg_Tile *tiles,
                         uint32_t count, spg_Metrics *metrics,
                         float *hostPixels);
/* Same, image in HBM (width*height*4 floats), no host copies. */
int spg_path_trace_tiles_device(const spg_Context *ctx, const spg_Tile *tiles,
                                uint32_t count, spg_Metrics *metrics,
                                float *devicePixels);

/* Closest hit in mesh space for a batch of rays: sp_RayIntersectMesh
 * (src/sp_scene.cpp:127-227) over a SAH BVH.  Among triangles with a
 * Moller-Trumbore hit at t > 0 (src/ray_intersection.cpp:156-190) the
 * smallest t wins, equal t goes to the lower triangle index.  primId is
 * SPG_MISS and t is -1 on a miss.  u, v may be NULL. */
int spg_intersect_closest(spg_Scene *scene, uint32_t meshId,
                          const float *ox, const float *oy, const float *oz,
                          const float *dx, const float *dy, const float *dz,
                          uint32_t n, uint32_t *primId, float *t, float *u,
                          float *v);
int spg_intersect_closest_device(spg_Scene *scene, uint32_t meshId,
                                 const float *ox, const float *oy,
                                 const float *oz, const float *dx,
                                 const float *dy, const float *dz, uint32_t n,
                                 uint32_t *primId, float *t, float *u,
                                 float *v, spg_TraversalStats *stats);
/* spg_intersect_closest_device without any allocation.  No counterpart in
 * the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).  The
 * persistent launch's work counter and traversal-stack spill live in a
 * static device array of the library (SPG_NOALLOC_SCRATCH_BYTES on every
 * GPU it runs on, held until the process ends); no hipMalloc, hipFree or
 * allocating library call is made.  Limits: a mesh whose walk needs more
 * than SPG_NOALLOC_STACK stack entries is rejected with SPG_E_CAPACITY; n
 * is unbounded (the launch is persistent).  All calls on one GPU share the
 * scratch: calls that overlap on one GPU (from different spg_Device handles
 * of that GPU) race.  No traversal statistics. */
#define SPG_NOALLOC_STACK 56
#define SPG_NOALLOC_SCRATCH_BYTES (256u + 4096u * 64u * 24u * 4u)
int spg_intersect_closest_device_noalloc(spg_Scene *scene, uint32_t meshId,
                                         const float *ox, const float *oy,
                                         const float *oz, const float *dx,
                                         const float *dy, const float *dz,
                                         uint32_t n, uint32_t *primId, float *t,
                                         float *u, float *v);
/* Scene-level closest hit in world space: sp_RayIntersectScene
 * (src/sp_scene.cpp:229-339).  out7 per ray: t, materialId bits,
 * normal.xyz, uv (7 floats); object/prim may be NULL. */
int spg_intersect_scene(spg_Scene *scene, const float *rays6, uint32_t n,
                        float *out7, uint32_t *object, uint32_t *prim);

/* Times `iterations` launches of the closest-hit kernel on device-resident
 * rays with HIP events on the launch stream; returns the mean kernel ms. */
int spg_time_intersect_closest(spg_Scene *scene, uint32_t meshId,
                               const float *ox, const float *oy,
                               const float *oz, const float *dx,
                               const float *dy, const float *dz, uint32_t n,
                               uint32_t *primId, float *t, uint32_t iterations,
                               double *meanKernelMs);
/* Kernel timings of the last trace call on this device, in ms, summed over
 * its launches: [0] raygen [1] extend [2] shade [3] finalize [4] guide
 * [5] total wall; launches[i] counts launches. */
int spg_last_trace_timings(spg_Device *dev, double *ms6, uint32_t *launches6);
/* Raw counters of the last trace call on this device: [0] rays [1] hits
 * [2] misses [3] mesh-hierarchy slab tests [4] mesh (instance) entries
 * [5] triangle tests [6] instance-hierarchy slab tests [7] guide network
 * evaluations (SPG_EST_GUIDED: queued hits that will scatter).
 * The test counts are collected only when the call's config set
 * collectMetrics. */
int spg_last_trace_counters(spg_Device *dev, uint64_t *counters8);

/* ---- batched tracing (no reference counterpart: the reference renders one
 *      scene at a time, 64x64 tiles on 16 CPU threads, src/main.cpp:731-844)
 * A batch renders many scenes -- each with its own camera and image size --
 * as one wavefront: every pixel of every item is a slot, paths of all
 * items share the ray queues, and each closest-hit launch is one persistent
 * launch over rays of all scenes.  spg_batch_create builds the per-item
 * tables, picks the launch shape by timing candidates on the batch itself
 * (autotune != 0) and captures the whole render in a HIP graph;
 * spg_batch_run replays it.  Each item's image equals what
 * spg_path_trace_tiles_device renders for that camera and scene alone
 * (same per-pixel random streams).  The scenes, cameras' images and the
 * device must outlive the batch; scenes must not be rebuilt meanwhile. */
typedef struct spg_Batch spg_Batch;

typedef struct {
    const spg_Camera *camera;  /* copied at creation */
    spg_Scene *scene;          /* built, on the batch's device */
    float *deviceImage;        /* RGBA32F width x height, device memory */
} spg_BatchItem;

typedef struct {
    uint32_t pathsInFlight;      /* wavefront capacity (paths) */
    uint32_t extendBlocks;       /* resident one-wave traversal blocks
                                    (0 = as many as fit) */
    uint32_t candidates;         /* launch shapes timed by the autotuner
                                    (0 when the shape came from its cache) */
    float tunedMsPerSample;      /* the chosen shape's time for 1 spp */
    uint64_t pixels;             /* slots over all items */
    uint32_t graphNodes;         /* kernels in the captured graph */
    uint32_t refill;             /* idle lanes a traversal wave waits for
                                    before taking new rays (0 = default) */
    uint32_t cached;             /* 1: the shape tuned for an earlier batch
                                    of the same configuration was reused */
} spg_BatchTuning;

/* cfg: samplesPerPixel, bounces, radianceClamp, estimator, rngMode, seed,
 * transcendentals, sampleOffset, collectMetrics; guide and guideTrain for
 * SPG_EST_GUIDED.  With the guided estimator every item must trace the same
 * scene (one guide learns one scene); guide inference at every bounce and,
 * with guideTrain, the pass's sample collection and training steps are
 * part of the captured graph, so each replay trains the guide further.
 * The guide is pinned while the batch lives: spg_guide_destroy fails and
 * its training buffers cannot grow (spg_guide_train with more samples than
 * the batch's passes collect fails).  The autotuner's timing runs do not
 * train. */
int spg_batch_create(spg_Device *dev, const spg_BatchItem *items, uint32_t n,
                     const spg_TraceConfig *cfg, int autotune, spg_Batch **out);
/* Replays the render; metrics (may be NULL) as sp_PathTraceTile's. */
int spg_batch_run(spg_Batch *b, spg_Metrics *metrics);
int spg_batch_tuning(const spg_Batch *b, spg_BatchTuning *out);
/* GPU time of the last spg_batch_run in ms (HIP events around the graph). */
int spg_batch_last_ms(const spg_Batch *b, double *ms);
/* The same render run eagerly (not from the graph) with every launch
 * timed: ms6/launches6 per kernel class as spg_last_trace_timings ([5] =
 * host wall time of the call), counters8 as spg_last_trace_counters
 * (test counts when collect != 0).  Any pointer may be NULL. */
int spg_batch_profile(spg_Batch *b, int collect, double *ms6, uint32_t *launches6,
                      uint64_t *counters8);
int spg_batch_destroy(spg_Batch *b);

/* ---- learned path guiding (no reference counterpart: the reference
 *      samples directions uniformly, src/math_lib.h:932-946) ----
 * A guide is an MLP (64 -> 4 x 64 ReLU -> 48, FP16 weights, FP32
 * accumulation on MFMA) mapping a hit (position normalized to the scene
 * bounds and frequency-encoded, shading normal, view direction) to a
 * mixture of 8 von Mises-Fisher lobes.  SPG_EST_GUIDED samples each bounce
 * by one-sample MIS of cosine BSDF sampling (probability bsdfFraction) and
 * the mixture, so the estimator stays unbiased for any network.  Training
 * (online, inside the trace call when guideTrain is set) minimizes the
 * weighted negative log-likelihood of the sampled directions with weights
 * f cos Li / pdf, normalized to mean 1 per batch, with Adam. */
#define SPG_GUIDE_PARAMS 19760  /* floats: W0 b0 ... W4 b4, W[out][in] */
#define SPG_GUIDE_OUTPUTS 48
#define SPG_GUIDE_LOBE_FLOATS 40 /* pi[8] kappa[8] mu[8][3] */
#define SPG_GUIDE_SAMPLE_FLOATS 13 /* pn[3] n[3] v[3] wi[3] weight */

typedef struct {
    uint32_t seed;          /* parameter initialization */
    float learningRate;     /* Adam step size (default 3e-3) */
    uint32_t batchSize;     /* samples per training step, multiple of 64
                               (default 16384) */
    uint32_t stepsPerPass;  /* training steps after each traced pass
                               (default 4) */
    float maxWeight;        /* clamp of batch-normalized sample weights,
                               0 = off (default 50) */
    float bsdfFraction;     /* MIS probability of BSDF sampling (0.5) */
} spg_GuideConfig;

/* Fills the defaults above. */
void spg_guide_default_config(spg_GuideConfig *cfg);
/* A guide on dev, initialized from cfg->seed (NULL cfg = defaults). */
int spg_guide_create(spg_Device *dev, const spg_GuideConfig *cfg,
                     spg_Guide **out);
int spg_guide_destroy(spg_Guide *g);
/* FP32 master parameters (SPG_GUIDE_PARAMS floats); set also resets the
 * Adam state. */
int spg_guide_get_params(spg_Guide *g, float *params, uint32_t count);
int spg_guide_set_params(spg_Guide *g, const float *params, uint32_t count);
/* The network on n hits given as rows of 9 floats (pn[3] in [0,1], n[3],
 * v[3]): raw outputs (n x 48) and the mixtures (n x 40), either may be
 * NULL.  Host arrays. */
int spg_guide_evaluate(spg_Guide *g, const float *inputs9, uint32_t n,
                       float *outputs48, float *lobes40);
/* Gradient of L = (1/n) sum_j w_j (-log q(wi_j | x_j)) over n samples (rows
 * of SPG_GUIDE_SAMPLE_FLOATS) computed by the training kernels, without a
 * parameter update.  normalize != 0 scales the weights to mean 1 and
 * maxWeight > 0 clamps them, as training does.  grad: SPG_GUIDE_PARAMS
 * floats; loss may be NULL.  n <= 65536. */
int spg_guide_gradient(spg_Guide *g, const float *samples13, uint32_t n,
                       int normalize, float maxWeight, float *grad,
                       float *loss);
/* `steps` Adam steps on host samples, consecutive batches of batchSize
 * (wrapping); mean loss of the last step in *loss (may be NULL). */
int spg_guide_train(spg_Guide *g, const float *samples13, uint32_t n,
                    uint32_t steps, float *loss);
/* Training state: samples collected in the last traced pass, Adam steps
 * taken, mean loss of the last step. */
int spg_guide_stats(spg_Guide *g, uint32_t *collected, uint32_t *steps,
                    float *loss);
/* The training samples collected by the last traced pass, in the order the
 * training steps read them (rows of SPG_GUIDE_SAMPLE_FLOATS: pn, n, v, wi,
 * weight); *count = samples available (at most the pass's sample
 * capacity), rows copied = min(*count, cap). */
int spg_guide_samples(spg_Guide *g, float *rows13, uint32_t cap, uint32_t *count);
/* Verification hooks of the guided estimator (no reference counterpart; the
 * tests use them to check the GPU against the oracle).  flags
 * SPG_GUIDE_ANALYTIC_MIXTURE: every vertex samples a fixed analytic mixture
 * instead of the network's (lobe 0 toward the point (0.3, 2.5, -0.4),
 * kappa 40, weight 0.5; lobe 1 about the normal, kappa 3, weight 0.3;
 * lobe 2 about the mirror direction 2 (n.v) n - v, kappa 120, weight 0.2).
 * recordCap > 0: from now on the guided estimator records its first
 * recordCap scatter events (SPG_GUIDE_RECORD_FLOATS floats each: pn[3]
 * n[3] v[3] roughness u0 u1 u2 lobes[40] wi[3] pdf hit[3] pad[4]); 0
 * stops recording.  Each call restarts the record count. */
#define SPG_GUIDE_ANALYTIC_MIXTURE 1u
#define SPG_GUIDE_RECORD_FLOATS 64
int spg_guide_debug(spg_Guide *g, uint32_t flags, uint32_t recordCap);
/* The recorded scatter events: *count = recorded so far, rows copied =
 * min(*count, cap). */
int spg_guide_debug_records(spg_Guide *g, float *records, uint32_t cap,
                            uint32_t *count);
/* Guide checkpoint (no reference counterpart; the optimizer state that
 * must be serializable): the configuration, the FP32 master parameters,
 * both Adam moments, the step count and the learned light split's sums, in
 * a little-endian file ("SPGGUIDE", version 2).  Training steps are
 * deterministic (fixed-order reductions, integer light sums), so a guide
 * loaded from a checkpoint continues training bit-identically to the guide
 * that wrote it. */
int spg_guide_save(spg_Guide *g, const char *path);
/* Restores a checkpoint into g; its configuration replaces g's.  Version-1
 * files (no light sums) load with the light split unlearned.
 * SPG_E_INVALID for a file that is not a checkpoint of this network. */
int spg_guide_load(spg_Guide *g, const char *path);
/* Learned light split (no reference counterpart): SPG_EST_GUIDED_NEE
 * draws a light sample from the sky with this probability (else from the
 * emitters) when the scene has both.  Training passes learn it as the
 * sky's share of the luminance their reached light samples delivered
 * (clamped to [0.05, 0.95]); before any, the scene's default split (0.5)
 * is used and *skyFraction reads -1.  spg_guide_set_light_split fixes it
 * (0.05..0.95) or, with -1, forgets it. */
int spg_guide_light_split(spg_Guide *g, float *skyFraction);
int spg_guide_set_light_split(spg_Guide *g, float skyFraction);

/* ---- after the path ---- */
/* Progressive accumulation across trace calls (the reference's open TODO,
 * src/main.cpp:75): dst = (dst * before + src * added) / (before + added)
 * per RGB channel, alpha 1, on device RGBA32F images of `pixels` texels
 * (e.g. before = samples already in dst, added = the new call's spp). */
int spg_accumulate_device(spg_Device *dev, float *dst, const float *src,
                          uint32_t pixels, uint32_t before, uint32_t added);
/* PerformToneMapping (src/shaders/post_processing.frag.glsl:19-26), the
 * viewer's post step for the ray-traced image: c *= exposure,
 * c = c / (1 + c), c = pow(c, 1/2.2) per RGB channel, then 8 bits as
 * round-to-nearest(255 c), alpha 255.  rgba: device RGBA32F; out: device
 * RGBA8.  transcendentals: SPG_TX_DOUBLE (pow in double, rounded once) or
 * SPG_TX_FAST. */
int spg_tonemap_device(spg_Device *dev, const float *rgba, uint32_t pixels,
                       float exposure, int transcendentals, uint8_t *out);
/* Host-only: sin and cos of each x as SPG_TX_DOUBLE evaluates them on the
 * device (the same code: a double-precision evaluation rounded once to
 * float), for checking them against a correctly rounded library. */
int spg_sincos_host(const float *x, uint32_t n, float *s, float *c);
/* Host-only: the importance-sampling table SPG_EST_NEE builds for a
 * background cube map (rgba: 6 faces of faceWidth x faceHeight RGBA32F in
 * the layout spg_cubemap_register takes).  Texel i (face-major, then rows,
 * then columns) is chosen with probability P_i proportional to its mean
 * channel value times its face-plane rectangle's area over
 * (1 + fx^2 + fy^2)^(3/2) at the rectangle's centre; pdf[i] = P_i / area_i
 * (times (1 + fx^2 + fy^2)^(3/2) this is the density per steradian of a
 * direction drawn uniformly in the rectangle); prob/alias: Vose's alias
 * table.  Arrays of 6 * faceWidth * faceHeight entries.  SPG_E_INVALID if
 * no texel has a positive finite value. */
int spg_env_table_build(const float *rgba, uint32_t faceWidth,
                        uint32_t faceHeight, float *pdf, float *prob,
                        uint32_t *alias);
/* Portable float map (PFM, "PF" colour, little-endian, rows bottom to top)
 * of a host RGBA32F image in the reference's layout (x + y * width, row 0
 * at the top); alpha is dropped. */
int spg_write_pfm(const char *path, uint32_t width, uint32_t height,
                  const float *rgba);
/* Times `iterations` accumulate launches, then `iterations` tone-map
 * launches, with HIP events on the launch stream (arguments as in the two
 * calls above); returns each kernel's mean ms.  dst keeps being averaged. */
int spg_time_post(spg_Device *dev, float *dst, const float *src,
                  uint32_t pixels, float exposure, int transcendentals,
                  uint8_t *out, uint32_t iterations, double *accumulateMs,
                  double *tonemapMs);

/* ---- the parked Vulkan compute tracer (src/shaders/test_compute.comp.glsl),
 *      a brute-force parity harness: no hierarchy, the shader's hard-coded
 *      materials ----
 * Traces the listed tiles (index t: the 16x16 tile (t % 16, t / 16) of the
 * shader's fixed 256x192 image) as the shader's main/PathTraceTile do: one
 * xorshift32 stream per tile seeded (t + 1) * 0xF51C0E49, one sample per
 * pixel, up to 3 rays per path, every object of the scene an entity tested
 * triangle by triangle (no back-face culling, flat normals), bounce
 * direction normalize(n + (B, B, B)) into n's hemisphere, materials by the
 * object's material id: 0 red albedo (0.18, 0.1, 0.1), 2 the 256x256
 * checkerboard (src/image.cpp:134-160) sampled bilinearly with repeat,
 * 5 white light (emission 1), others black.  cameraTransform: the
 * uniform's column-major mat4 (columns right, up, forward, position).
 * image: host RGBA32F 256x192; only the tiles' pixels are written. */
int spg_compute_tracer_tiles(spg_Scene *scene, const float cameraTransform[16],
                             const uint32_t *tileIndices, uint32_t count,
                             float *image);

/* ---- mesh files ---- */
/* LoadMesh (src/mesh.cpp:5-62) for Wavefront OBJ text: the triangles
 * assimp's import yields there (aiProcess_Triangulate: polygons split into
 * fans; aiProcess_JoinIdenticalVertices: one vertex per distinct position +
 * normal pair, in order of first use), positions and normals of the first
 * object ('o'), texture coordinates left 0 as LoadMesh leaves them.
 * Corners without a normal are rejected (LoadMesh asserts normals).  Call
 * with verts8 = idx = NULL for nv/ni, then with arrays of that size; the
 * result feeds spg_mesh_create. */
int spg_obj_parse(const char *text, size_t len, float *verts8, uint32_t vertCap,
                  uint32_t *idx, uint32_t idxCap, uint32_t *nv, uint32_t *ni);

/* ---- environment maps ---- */
/* CreateIrradianceCubeMap(hdri, arena, faceWidth, faceHeight)
 * (src/cubemap.cpp:108-233, called at src/main.cpp:1313 with 32x32) with
 * the reference's production switches: uniform sampling of the hemisphere
 * about each texel's direction (phi, theta stepped by 0.1) and radiance
 * clamped to radianceClamp (RADIANCE_CLAMP, src/config.h:37; <= 0: no
 * clamp).  equirect: RGBA32F equirectangular image (HdrImage pixels,
 * ew*eh*4 floats, host memory); out: 6 faces of faceWidth*faceHeight
 * RGBA32F texels (host memory, face-major like HdrCubeMap.images).
 * transcendentals: SPG_TX_DOUBLE (sums bit-identical to the oracle's
 * double-evaluated restatement) or SPG_TX_FAST.  kernelMs (may be NULL)
 * receives the kernel's time. */
int spg_create_irradiance_cubemap(spg_Device *dev, const float *equirect,
                                  uint32_t ew, uint32_t eh, uint32_t faceWidth,
                                  uint32_t faceHeight, float radianceClamp,
                                  int transcendentals, float *out,
                                  double *kernelMs);

/* ---- generators used by the benchmark scenes (no reference counterpart
 *      except the icosphere, CreateIcosahedronMesh,
 *      src/mesh_generation.cpp:220-337, and the plane, :1-21) ---- */
int spg_gen_icosphere(uint32_t level, float *verts8, uint32_t vertCap,
                      uint32_t *idx, uint32_t idxCap, uint32_t *nv,
                      uint32_t *ni);
int spg_gen_plane(float *verts8, uint32_t *idx);
/* Triangle soup: centroids uniform in [0,1]^3, equilateral triangles of
 * edge exp(N(mu, sigma)) in a random orientation; tris9 = a,b,c. */
int spg_gen_triangle_soup(uint64_t seed, uint32_t n, float mu, float sigma,
                          float *tris9);
/* Sky gradient + sun disk environment baked into a cube map
 * (6*size*size RGBA32F); sunDir need not be normalized. */
int spg_gen_sky_cubemap(uint32_t size, const float sunDir[3],
                        float sunAngularRadius, float sunRadiance,
                        float *rgba6);
/* The same environment as an equirect image for the reference's lookup. */
int spg_gen_sky_equirect(uint32_t width, uint32_t height,
                         const float sunDir[3], float sunAngularRadius,
                         float sunRadiance, float *rgba);

/* Self-check of the device arithmetic that bit-exact results rely on (no
 * reference counterpart): the tracer's fast reciprocal (hardware reciprocal
 * plus one FMA Newton step, used where x and 1/x are normal) against the
 * IEEE division for all 2^32 float bit patterns; counts2[0] = patterns in
 * the fast range tested, counts2[1] = results that differ (0 expected). */
int spg_check_arithmetic(spg_Device *dev, uint64_t *counts2);

/* xorshift32 seed of pixel `index` in SPG_RNG_XORSHIFT_PIXEL mode. */
uint32_t spg_pixel_seed(uint32_t base, uint32_t index);

#ifdef __cplusplus
}
#endif
#endif /* SPGPU_H */
