/* ORACLE -- test infrastructure only.
 *
 * A plain-C restatement of the reference CPU path tracer (deadVertex/
 * vk_cinematic, the `sp_` tracer) used by tests/, by bench.py's cpu_baseline
 * leg and by __graft_entry__.smoke() as the checker of the HIP product.  The
 * product never includes, links or calls anything in this directory.
 *
 * Every function names the reference file:line it restates.  Arithmetic is
 * written operation-for-operation in the reference's association order and
 * compiled with -ffp-contract=off, so that with SPO_TX_LIBM it reproduces the
 * reference's FP32 results bit for bit (checked against the reference's own
 * sources compiled into oracle/_ref and against golden vectors).
 *
 * Modes (spo_Config):
 *   rng   SPO_RNG_XORSHIFT_TILE  one xorshift32 stream through the tile in
 *                                pixel order (src/simd_path_tracer.cpp:178-345)
 *         SPO_RNG_XORSHIFT_PIXEL one xorshift32 stream per pixel, seeded by
 *                                spo_pixel_seed()
 *         SPO_RNG_PCG            one PCG32 stream per pixel, sample s starting
 *                                2^20*s draws into it
 *   tx    SPO_TX_LIBM  sinf/cosf/atan2f/powf from libm (the reference's)
 *         SPO_TX_CR    the same functions evaluated in double and rounded
 *                      once to float (what the GPU path computes; x^5 by
 *                      double multiplications)
 *   hit   SPO_HIT_REFERENCE  candidates from the reference's greedy 4-ary BVH,
 *                            first-visited wins ties
 *         SPO_HIT_BRUTE      every triangle is a candidate, ties go to the
 *                            lower triangle index / lower object index
 *
 * Parity status: every function is pinned by a test -- the reference's
 * golden vectors or its own build in oracle/_ref where it has a reference
 * counterpart; closed forms, brute force, quadrature, finite differences or
 * distribution checks where it has none (tests/test_oracle_*.py,
 * tests/test_wide_bvh.py, tests/test_nee.py, tests/test_post.py).  No
 * function is left parity unpinned.
 */
#ifndef SPO_H
#define SPO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { SPO_RNG_XORSHIFT_TILE = 0, SPO_RNG_XORSHIFT_PIXEL = 1, SPO_RNG_PCG = 2 };
enum { SPO_TX_LIBM = 0, SPO_TX_CR = 1 };
enum { SPO_HIT_REFERENCE = 0, SPO_HIT_BRUTE = 1 };
enum { SPO_EST_PARITY = 0, SPO_EST_UNBIASED = 1, SPO_EST_GUIDED = 2, SPO_EST_NEE = 3,
       SPO_EST_GUIDED_NEE = 4 };

#define SPO_MAX_METRICS 12
#define SPO_NO_TEXTURE 0xFFFFFFFFu

typedef struct { float x, y, z; } spo_v3;
typedef struct { float x, y; } spo_v2;

/* VertexPNT, src/mesh.h:10-15 (32 bytes) */
typedef struct { spo_v3 p, n; spo_v2 uv; } spo_Vertex;

/* bvh_Node, src/bvh.h:5-12, with child indices instead of pointers */
typedef struct {
    spo_v3 min, max;
    int32_t children[4];
    uint32_t leafIndex;
} spo_BvhNode;

typedef struct {
    spo_BvhNode *nodes;
    uint32_t count, capacity;
    int32_t root; /* -1 = empty tree */
} spo_Bvh;

typedef struct {
    uint32_t errorOccurred, count, aabbTestCount;
} spo_BvhResult;

/* sp_Mesh, src/sp_scene.h:3-12 (owns copies of its arrays).  A mesh may
 * instead carry an externally built BVH2 (the product's SAH hierarchy, 16
 * floats per node, see spo_bvh2.c); its query is then the closest-hit walk
 * of that hierarchy with equal t going to the lower triangle index. */
typedef struct {
    spo_Vertex *vertices;
    uint32_t *indices;
    uint32_t vertexCount, indexCount;
    int useSmoothShading;
    spo_Bvh midphase;
    float *bvh2Nodes;
    uint32_t *bvh2Order;
    uint32_t bvh2Count;
} spo_Mesh;

typedef struct { float m[16]; } spo_m4; /* column-major, src/math_lib.h:93 */

typedef struct {
    uint32_t mesh, material;
    spo_m4 model, invModel;
    spo_v3 aabbMin, aabbMax;
} spo_Object;

/* sp_Material, src/sp_material_system.h:3-12 */
typedef struct {
    spo_v3 albedo;
    uint32_t albedoTexture;
    spo_v3 emission;
    uint32_t emissionTexture;
    float roughness;
} spo_Material;

typedef struct {
    const float *pixels; /* RGBA32F; cube maps hold 6 faces in order */
    uint32_t width, height;
    int isCube;
} spo_Image;

/* storage; the registry accepts materialCapacity entries, 32 by default
 * (SP_MAX_MATERIALS, src/sp_material_system.h:32) */
#define SPO_MAX_MATERIALS 4096
#define SPO_MAX_IMAGES 16

typedef struct spo_Scene {
    spo_Mesh *meshes;
    uint32_t meshCount, meshCapacity;
    spo_Object *objects;
    uint32_t objectCount, objectCapacity;
    spo_Bvh broadphase;
    /* sp_MaterialSystem, src/sp_material_system.h:35-46 */
    uint32_t materialKeys[SPO_MAX_MATERIALS];
    spo_Material materials[SPO_MAX_MATERIALS];
    uint32_t materialCount, materialCapacity;
    uint32_t imageKeys[SPO_MAX_IMAGES];
    spo_Image images[SPO_MAX_IMAGES];
    uint32_t imageCount;
    uint32_t backgroundMaterialId;
    /* limits that the reference hard-codes (src/sp_scene.cpp:139,:246) */
    uint32_t maxBroadphaseHits, maxMidphaseHits;
    /* next-event estimation table of the background's cube map
     * (spo_scene_build_env); envPdf NULL: the environment is not sampled */
    float *envPdf, *envProb;
    uint32_t *envAlias;
    uint32_t envW, envH;
    /* emissive triangles (spo_scene_build_env); emitCount 0: none */
    float *emitPdfA, *emitProb;
    uint32_t *emitAlias, *emitPair;
    int32_t *emitBase;
    uint32_t emitCount;
    /* the spatial emitter choice (emitInstCount > 0; spo_scene_emitter_
     * instances): 8 floats per emitter instance, emitter index per object,
     * per-entry alias tables over each instance's triangles by area */
    float *emitInst, *emitProbL;
    int32_t *emitSlot;
    uint32_t *emitAliasL;
    uint32_t emitInstCount;
    float envSelect;     /* probability of sampling the environment */
} spo_Scene;

/* sp_Camera, src/simd_path_tracer.h:17-32 */
typedef struct {
    spo_v3 right, up, forward;
    spo_v3 position, filmCenter;
    uint32_t width, height;
    float halfPixelWidth, halfPixelHeight;
    float halfFilmWidth, halfFilmHeight;
} spo_Camera;

typedef struct {
    int rng, tx, hit, estimator;
    uint32_t samplesPerPixel, bounces;
    float radianceClamp;      /* 0 = no clamp */
    int jitterYFirst;         /* argument evaluation order of the jitter draws */
    uint32_t seed;            /* xorshift tile state, or base seed */
    float bsdfFraction;       /* SPO_EST_GUIDED: MIS probability of the BSDF */
} spo_Config;

/* closest-hit record, sp_RayIntersectSceneResult (src/sp_scene.h:39-53)
 * plus the object and triangle that produced it */
typedef struct {
    float t;
    uint32_t materialId;
    spo_v3 normal;
    spo_v2 uv;
    uint32_t object, prim;
} spo_SceneHit;

/* ---- primitives ---- */
float spo_rt_mt(const float *o, const float *d, const float *a, const float *b,
                const float *c, float *out6);
float spo_rt_triangle_slow(const float *o, const float *d, const float *a,
                           const float *b, const float *c, float *out6);
float spo_rt_aabb_scalar(const float *mn, const float *mx, const float *o,
                         const float *d);
float spo_rt_aabb_simd1(const float *mn, const float *mx, const float *o,
                        const float *d);
uint32_t spo_rt_aabb4(const float *mins12, const float *maxs12, const float *o,
                      const float *invD);
double spo_rt_mt_f64(const double *o, const double *d, const double *a,
                     const double *b, const double *c);

uint32_t spo_xorshift32(uint32_t *state);
float spo_random_unilateral(uint32_t *state);
float spo_random_bilateral(uint32_t *state);
uint32_t spo_pixel_seed(uint32_t base, uint32_t pixelIndex);
void spo_pcg32_seed(uint64_t *state, uint64_t *inc, uint64_t initstate,
                    uint64_t initseq);
uint32_t spo_pcg32_next(uint64_t *state, uint64_t inc);
void spo_pcg32_advance(uint64_t *state, uint64_t inc, uint64_t delta);

void spo_hemisphere(const float *normal, uint32_t *state, int tx, float *out);
void spo_to_spherical(const float *v, int tx, float *out2);
void spo_map_to_equirect(const float *sc, int tx, float *out2);
void spo_spherical_to_cartesian(const float *sc, int tx, float *out3);
void spo_sample_nearest(const float *pixels, uint32_t w, uint32_t h,
                        const float *uv, float *out4);
void spo_sample_bilinear(const float *pixels, uint32_t w, uint32_t h,
                         const float *uv, float *out4);
void spo_cubemap_basis(uint32_t face, float *out9);
void spo_cubemap_lookup(const float *faces, uint32_t w, uint32_t h,
                        const float *dir, float *out4);
/* sin and cos in double, rounded once to float, of an array (SPO_TX_CR) */
void spo_sincos_cr_array(const float *x, uint32_t n, float *s, float *c);
/* ---- environment importance sampling (next-event estimation) ---- */
int spo_env_table_build(const float *rgba, uint32_t w, uint32_t h, float *pdf,
                        float *prob, uint32_t *alias);
float spo_env_pdf(const float *pdfTable, uint32_t w, uint32_t h, const float *dir);
void spo_env_direction(const float *prob, const uint32_t *alias, uint32_t w,
                       uint32_t h, uint32_t r, float e1, float e2, float e3,
                       float *out3);
/* Builds the scene's light tables: the background material's cube-map
 * emission (1 if built, else the sky is not sampled) and the emissive
 * triangles of instances whose material has a constant emission; sets the
 * environment's selection probability (0.5 with both, else 1 or 0). */
int spo_scene_build_env(spo_Scene *s);
/* Overrides the sky share of light samples (scenes with sky and emitters). */
void spo_scene_set_light_split(spo_Scene *s, float skyFraction);
/* The emitter table alone (arrays of the scene's emissive triangle count,
 * base: one entry per object); returns the entry count. */
uint32_t spo_scene_emitter_table(const spo_Scene *s, float *pdfA, float *prob,
                                 uint32_t *alias, uint32_t *pair, int32_t *base);
/* The spatial emitter choice's tables (next-event estimation; no reference
 * counterpart), given spo_scene_emitter_table's base: per emitter instance
 * 8 floats in rec (bounding-sphere centre xyz and radius of its world
 * vertices, power = mean emission x world area, 1 / area, first table
 * entry and triangle count as u32 bits); slot: per object its emitter index
 * or -1; probL / aliasL (arrays of the table's entry count): per entry an
 * alias table over its object's triangles by area, local indices.  Objects
 * with zero total area get no record; more than 256 records give none.
 * Returns the record count. */
#define SPO_MAX_SPATIAL_EMITTERS 256
uint32_t spo_scene_emitter_instances(const spo_Scene *s, const int32_t *base,
                                     uint32_t entries, float *rec, int32_t *slot,
                                     float *probL, uint32_t *aliasL);
void spo_create_cubemap(const float *equirect, uint32_t ew, uint32_t eh,
                        uint32_t w, uint32_t h, float *out);
void spo_create_irradiance_cubemap(const float *equirect, uint32_t ew, uint32_t eh,
                                   uint32_t w, uint32_t h, float clampValue, int tx,
                                   float *out);
void spo_tonemap(const float *rgba, uint32_t pixels, float exposure, int tx,
                 uint8_t *out);
void spo_accumulate(float *dst, const float *src, uint32_t pixels, uint32_t before,
                    uint32_t added);
uint32_t spo_compute_tiles(uint32_t W, uint32_t H, uint32_t tw, uint32_t th,
                           uint32_t maxTiles, uint32_t *out4);
void spo_transform_aabb(const float *mn, const float *mx, const float *pos,
                        const float *q, const float *scale, float *out6);
void spo_object_matrices(const float *pos, const float *q, const float *scale,
                         float *model16, float *inv16);
void spo_rotate_vector(const float *v, const float *q, float *out3);
void spo_quat_axis_angle(const float *axis, float angle, int tx, float *out4);
void spo_quat_mul(const float *p, const float *q, float *out4);

/* ---- meshes ---- */
void spo_icosphere(uint32_t level, float *verts8, uint32_t *idx, uint32_t *nv,
                   uint32_t *ni);
void spo_plane(float *verts8, uint32_t *idx, uint32_t *nv, uint32_t *ni);

/* ---- reference BVH (src/bvh.cpp) ---- */
spo_Bvh *spo_bvh_create(const float *mins, const float *maxs, uint32_t n);
void spo_bvh_destroy(spo_Bvh *t);
uint32_t spo_bvh_intersect(const spo_Bvh *t, const float *o, const float *d,
                           uint32_t maxHits, uint32_t *leaves,
                           uint32_t *aabbTests, uint32_t *errorOccurred);
uint32_t spo_bvh_dump(const spo_Bvh *t, float *boxes, int32_t *children,
                      uint32_t *leafIndex, uint32_t cap);

/* ---- scene / materials / camera ---- */
spo_Scene *spo_scene_create(void);
void spo_scene_destroy(spo_Scene *s);
int spo_mesh_create(spo_Scene *s, const float *verts8, uint32_t nv,
                    const uint32_t *idx, uint32_t ni, int smooth,
                    int buildMidphase);
int spo_scene_add_object(spo_Scene *s, uint32_t mesh, uint32_t material,
                         const float *pos, const float *q, const float *scale);
/* Attach an external BVH2 to a mesh (copied). */
int spo_mesh_set_bvh2(spo_Scene *s, uint32_t mesh, const float *nodes16,
                      uint32_t nNodes, const uint32_t *order);
/* The reference's fixed limits (32 objects and 128 leaves per query, 32
 * materials) can be raised for scenes beyond the reference's scale
 * (materials up to SPO_MAX_MATERIALS). */
void spo_scene_set_limits(spo_Scene *s, uint32_t maxBroadphase,
                          uint32_t maxMidphase, uint32_t maxMaterials);
void spo_scene_build(spo_Scene *s);
int spo_material_register(spo_Scene *s, const float *albedo,
                          uint32_t albedoTexture, const float *emission,
                          uint32_t emissionTexture, float roughness,
                          uint32_t id);
int spo_texture_register(spo_Scene *s, const float *rgba, uint32_t w,
                         uint32_t h, uint32_t id, int isCube);
void spo_set_background(spo_Scene *s, uint32_t id);
/* GGX half-vector sampling and its density (unbiased estimator). */
void spo_ggx_direction(const float *n, const float *V, float roughness, float u1,
                       float u2, int tx, float *out3);
float spo_ggx_pdf(const float *n, const float *V, const float *L, float roughness);
void spo_evaluate_material(const spo_Scene *s, uint32_t id, const float *uv,
                           const float *outgoingDir, int tx, float *out7);
void spo_compute_radiance_for_path(const spo_Scene *s, const float *path15,
                                   uint32_t len, int tx, float clampValue,
                                   float *out3);

void spo_camera_configure(spo_Camera *c, uint32_t w, uint32_t h,
                          const float *pos, const float *q,
                          float filmDistance);
void spo_film_position(const spo_Camera *c, float px, float py, float *out3);

int spo_scene_intersect(const spo_Scene *s, const float *o, const float *d,
                        int hitMode, int tx, spo_SceneHit *out,
                        uint64_t *metrics12);
/* The guided estimator's scatter (FP32, the product's operation order):
 * with probability fb the BSDF lobe (cosine, or GGX when u0 < fb * ps on
 * rough materials), otherwise the vMF mixture lobes40 (pi[8] kappa[8]
 * mu[8][3]) with u0' = (u0 - fb) / (1 - fb); pdf = fb * p_bsdf + (1 - fb) *
 * q(wi) where cos(n, wi) > 0, else 0. */
void spo_guided_scatter(const float *lobes40, const float *n, const float *v,
                        float roughness, float fb, float u0, float u1, float u2,
                        int tx, float *wi3, float *pdf);
/* The verification mixture of the product's SPG_GUIDE_ANALYTIC_MIXTURE at
 * hit point P with normal n and view direction v. */
void spo_guide_analytic_lobes(const float *P, const float *n, const float *v,
                              float *lobes40);
/* SPO_EST_GUIDED renders with the verification mixture; candidate training
 * samples (every scattering vertex with a positive weight, in the order the
 * paths are traced) are appended to this sink, 16 floats each: hit point
 * P[3], n[3], v[3], wi[3], weight, pixel index, sample index, vertex index
 * (the last three as float).  cap 0 turns the sink off. */
void spo_guided_sink(float *rows16, uint32_t cap);
uint32_t spo_guided_sink_count(void);
/* 1 if some object's closest mesh hit lies at a positive world t (the
 * shadow test toward the sky), 0 if none, -1 on a broadphase overflow */
int spo_scene_occluded(const spo_Scene *s, const float *o, const float *d, int hitMode,
                       uint64_t *metrics12);
float spo_mesh_intersect(const spo_Scene *s, uint32_t mesh, const float *o,
                         const float *d, int hitMode, float *out6,
                         uint32_t *prim, uint64_t *metrics12);

/* batch closest-hit over one mesh in mesh space (brute force, threaded) */
void spo_closest_hit_brute(const float *tris9, uint32_t ntris,
                           const float *rays6, uint32_t nrays,
                           uint32_t threads, uint32_t *primOut, float *tOut);
/* same rays in FP64 (the t-error oracle) */
void spo_closest_hit_brute_f64(const float *tris9, uint32_t ntris,
                               const float *rays6, uint32_t nrays,
                               uint32_t threads, uint32_t *primOut,
                               double *tOut);

/* FP64 Moller-Trumbore t for given (ray, triangle) pairs (-1 = no hit) */
void spo_mt_f64_pairs(const float *tris9, const float *rays6,
                      const uint32_t *prim, uint32_t n, double *tOut);

/* ---- the GPU's BVH2 traversal restated (spo_bvh2.c) ---- */
void spo_bvh2_closest(const float *nodes16, const uint32_t *order,
                      const float *tris9, const float *rays6, uint32_t n,
                      uint32_t threads, uint32_t *primOut, float *tOut,
                      uint64_t *boxTests, uint64_t *triTests);
int spo_bvh2_check(const float *nodes16, uint32_t nNodes, const uint32_t *order,
                   const float *tris9, uint32_t nTris);
/* ---- the GPU's four-wide quantized hierarchy walk (spo_bvh2.c) ---- */
void spo_wide_closest(const uint8_t *nodes64, const uint32_t *order,
                      const float *tris9, const float *rays6, uint32_t n,
                      uint32_t threads, uint32_t *primOut, float *tOut,
                      uint64_t *boxTests, uint64_t *triTests);
int spo_wide_check(const uint8_t *nodes64, uint32_t nNodes, const uint32_t *order,
                   const float *tris9, uint32_t nTris, uint32_t *maxPending);

/* ---- path tracing ---- */
int spo_path_trace_tile(const spo_Scene *s, const spo_Camera *c,
                        const spo_Config *cfg, uint32_t minX, uint32_t minY,
                        uint32_t maxX, uint32_t maxY, uint32_t *rngState,
                        float *pixels, uint64_t *metrics12);
/* Whole image with a pool of worker threads pulling tiles from an atomic
 * counter, as src/main.cpp:731-844 does; returns wall seconds. */
double spo_render_threaded(const spo_Scene *s, const spo_Camera *c,
                           const spo_Config *cfg, uint32_t tileW,
                           uint32_t tileH, uint32_t threads, float *pixels,
                           uint64_t *metrics12);

#ifdef __cplusplus
}
#endif
#endif
