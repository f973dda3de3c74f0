// Test infrastructure only: a harness that compiles the reference tracer's own
// hot-path sources (as its unit tests do: src/*.cpp included into one
// translation unit, see unit_tests/test_simd_path_tracer.cpp:1-31) and exports
// a C ABI so that tests/golden/make_golden.py can record golden vectors and
// tests can cross-check the oracle restatement.  Nothing here ships in the
// product, and no reference source is copied: the sources are read from the
// reference checkout through the include path given by oracle/Makefile.
//
// Two variants are built from this file:
//   libref.so      production switches (src/config.h: Moller-Trumbore
//                  triangles, 4-wide AABB test, RADIANCE_CLAMP 10)
//   libref_test.so the switches the reference's unit tests compile with
//                  (no config.h: slow triangle test, scalar AABB, no clamp)

#include <cstdarg>
#include <cstring>

#ifndef REF_TEST_CONFIG
#include "config.h"
#undef SAMPLES_PER_PIXEL
#endif
static unsigned g_refSamplesPerPixel = 1;
#define SAMPLES_PER_PIXEL g_refSamplesPerPixel

#include "platform.h"
#include "math_lib.h"
#include "tile.h"
#include "memory_pool.h"
#include "bvh.h"
#include "ray_intersection.h"
#include "asset_loader/asset_loader.h"
#include "image.h"
#include "mesh.h"
#include "sp_scene.h"
#include "sp_material_system.h"
#include "simd_path_tracer.h"
#include "sp_metrics.h"
#include "simd.h"
#include "aabb.h"

#include "memory_pool.cpp"
#include "ray_intersection.cpp"
#include "bvh.cpp"
#include "sp_scene.cpp"
#include "sp_material_system.cpp"
#include "simd_path_tracer.cpp"
#include "mesh_generation.cpp"
#ifndef REF_TEST_CONFIG
#include "cubemap.cpp" // needs RADIANCE_CLAMP from config.h
#endif

static void RefLog(const char *fmt, ...)
{
    va_list args;
    va_start(args, fmt);
    vfprintf(stderr, fmt, args);
    va_end(args);
    fputc('\n', stderr);
}

struct RefState
{
    MemoryArena arena;      // meshes, BVH nodes, scene
    MemoryArena tempArena;
    sp_Scene scene;
    sp_MaterialSystem materials;
    sp_Camera camera;
    ImagePlane imagePlane;
    sp_Mesh meshes[64];
    u32 meshCount;
    // owned copies of mesh data (the reference borrows pointers)
    VertexPNT *meshVertices[64];
    u32 *meshIndices[64];
    float *texturePixels[SP_MAX_IMAGES];
    u32 textureCount;
};

static RefState *g_ref = NULL;

static vec3 V3(const float *p) { return Vec3(p[0], p[1], p[2]); }
static void Put3(float *dst, vec3 v) { dst[0] = v.x; dst[1] = v.y; dst[2] = v.z; }

extern "C" {

int ref_is_test_config(void)
{
#ifdef REF_TEST_CONFIG
    return 1;
#else
    return 0;
#endif
}

void ref_reset(unsigned long long arenaBytes)
{
    LogMessage = RefLog;
    if (g_ref)
    {
        free(g_ref->arena.base);
        free(g_ref->tempArena.base);
        for (u32 i = 0; i < g_ref->meshCount; ++i)
        {
            free(g_ref->meshVertices[i]);
            free(g_ref->meshIndices[i]);
        }
        for (u32 i = 0; i < g_ref->textureCount; ++i)
            free(g_ref->texturePixels[i]);
        free(g_ref->imagePlane.pixels);
        free(g_ref);
    }
    g_ref = (RefState *)calloc(1, sizeof(RefState));
    InitializeMemoryArena(&g_ref->arena, calloc(1, arenaBytes), arenaBytes);
    InitializeMemoryArena(
        &g_ref->tempArena, calloc(1, arenaBytes / 4), arenaBytes / 4);
    sp_InitializeScene(&g_ref->scene, &g_ref->arena);
}

// ---------------------------------------------------------------- primitives
void ref_mt(const float *o, const float *d, const float *a, const float *b,
    const float *c, float *out /* t,u,v,nx,ny,nz */)
{
    RayIntersectTriangleResult r =
        RayIntersectTriangleMT(V3(o), V3(d), V3(a), V3(b), V3(c));
    out[0] = r.t; out[1] = r.uv.x; out[2] = r.uv.y;
    Put3(out + 3, r.normal);
}

void ref_triangle(const float *o, const float *d, const float *a,
    const float *b, const float *c, float *out)
{
    RayIntersectTriangleResult r =
        RayIntersectTriangle(V3(o), V3(d), V3(a), V3(b), V3(c));
    out[0] = r.t; out[1] = r.uv.x; out[2] = r.uv.y;
    Put3(out + 3, r.normal);
}

unsigned ref_aabb4(const float *mins12, const float *maxs12, const float *o,
    const float *invD)
{
    vec3 mn[4], mx[4];
    for (int i = 0; i < 4; ++i)
    {
        mn[i] = V3(mins12 + 3 * i);
        mx[i] = V3(maxs12 + 3 * i);
    }
    return simd_RayIntersectAabb4(mn, mx, V3(o), V3(invD));
}

float ref_aabb_scalar(const float *mn, const float *mx, const float *o,
    const float *d)
{
    return RayIntersectAabb(V3(mn), V3(mx), V3(o), V3(d));
}

float ref_aabb_simd1(const float *mn, const float *mx, const float *o,
    const float *d)
{
    return simd_RayIntersectAabb(V3(mn), V3(mx), V3(o), V3(d));
}

void ref_inverse(const float *d, float *out) { Put3(out, Inverse(V3(d))); }

unsigned ref_xorshift(unsigned *state)
{
    RandomNumberGenerator rng = {*state};
    u32 r = XorShift32(&rng);
    *state = rng.state;
    return r;
}

float ref_random_unilateral(unsigned *state)
{
    RandomNumberGenerator rng = {*state};
    f32 r = RandomUnilateral(&rng);
    *state = rng.state;
    return r;
}

float ref_random_bilateral(unsigned *state)
{
    RandomNumberGenerator rng = {*state};
    f32 r = RandomBilateral(&rng);
    *state = rng.state;
    return r;
}

void ref_random_direction_on_hemisphere(const float *n, unsigned *state,
    float *out)
{
    RandomNumberGenerator rng = {*state};
    Put3(out, RandomDirectionOnHemisphere(V3(n), &rng));
    *state = rng.state;
}

void ref_to_spherical(const float *v, float *out)
{
    vec2 r = ToSphericalCoordinates(V3(v));
    out[0] = r.x; out[1] = r.y;
}

void ref_map_to_equirect(const float *sc, float *out)
{
    vec2 r = MapToEquirectangular(Vec2(sc[0], sc[1]));
    out[0] = r.x; out[1] = r.y;
}

void ref_spherical_to_cartesian(const float *sc, float *out)
{
    Put3(out, MapSphericalToCartesianCoordinates(Vec2(sc[0], sc[1])));
}

void ref_normalize(const float *v, float *out) { Put3(out, Normalize(V3(v))); }

void ref_sample_nearest(const float *pixels, unsigned w, unsigned h,
    const float *uv, float *out)
{
    HdrImage img = {(float *)pixels, w, h};
    vec4 s = SampleImageNearest(img, Vec2(uv[0], uv[1]));
    memcpy(out, s.data, 16);
}

void ref_sample_bilinear(const float *pixels, unsigned w, unsigned h,
    const float *uv, float *out)
{
    HdrImage img = {(float *)pixels, w, h};
    vec4 s = SampleImageBilinear(img, Vec2(uv[0], uv[1]));
    memcpy(out, s.data, 16);
}

#ifndef REF_TEST_CONFIG
void ref_cubemap_basis(unsigned face, float *out /* fwd,up,right */)
{
    BasisVectors b = MapCubeMapLayerIndexToBasisVectors(face);
    Put3(out, b.forward); Put3(out + 3, b.up); Put3(out + 6, b.right);
}

// Resamples an equirectangular RGBA32F image into 6 faces of w*h texels
// (src/cubemap.cpp CreateCubeMap); out holds 6*w*h*4 floats.
void ref_create_cubemap(const float *equirect, unsigned ew, unsigned eh,
    unsigned w, unsigned h, float *out)
{
    MemoryArena arena;
    u64 bytes = (u64)6 * w * h * 16 + 64;
    InitializeMemoryArena(&arena, calloc(1, bytes), bytes);
    HdrImage src = {(float *)equirect, ew, eh};
    HdrCubeMap cm = CreateCubeMap(src, &arena, w, h);
    for (u32 f = 0; f < 6; ++f)
        memcpy(out + (u64)f * w * h * 4, cm.images[f].pixels, (u64)w * h * 16);
    free(arena.base);
}

// Irradiance cube map of an equirectangular RGBA32F image (src/cubemap.cpp
// CreateIrradianceCubeMap, production switches); out holds 6*w*h*4 floats.
void ref_create_irradiance_cubemap(const float *equirect, unsigned ew, unsigned eh,
    unsigned w, unsigned h, float *out)
{
    MemoryArena arena;
    u64 bytes = (u64)6 * w * h * 16 + 64;
    InitializeMemoryArena(&arena, calloc(1, bytes), bytes);
    HdrImage src = {(float *)equirect, ew, eh};
    HdrCubeMap cm = CreateIrradianceCubeMap(src, &arena, w, h);
    for (u32 f = 0; f < 6; ++f)
        memcpy(out + (u64)f * w * h * 4, cm.images[f].pixels, (u64)w * h * 16);
    free(arena.base);
}
#endif

unsigned ref_compute_tiles(unsigned W, unsigned H, unsigned tw, unsigned th,
    unsigned maxTiles, unsigned *out /* 4 per tile */)
{
    Tile *tiles = (Tile *)calloc(maxTiles ? maxTiles : 1, sizeof(Tile));
    u32 n = ComputeTiles(W, H, tw, th, tiles, maxTiles);
    memcpy(out, tiles, sizeof(Tile) * n);
    free(tiles);
    return n;
}

void ref_transform_aabb(const float *mn, const float *mx, const float *pos,
    const float *q, const float *scale, float *out)
{
    quat qq = {q[0], q[1], q[2], q[3]};
    Aabb r = TransformAabb(V3(mn), V3(mx), V3(pos), qq, V3(scale));
    Put3(out, r.min); Put3(out + 3, r.max);
}

void ref_rotate_vector(const float *v, const float *q, float *out)
{
    quat qq = {q[0], q[1], q[2], q[3]};
    Put3(out, RotateVector(V3(v), qq));
}

void ref_quat_axis_angle(const float *axis, float angle, float *out)
{
    quat r = Quat(V3(axis), angle);
    memcpy(out, r.data, 16);
}

void ref_quat_mul(const float *p, const float *q, float *out)
{
    quat a = {p[0], p[1], p[2], p[3]};
    quat b = {q[0], q[1], q[2], q[3]};
    quat r = a * b;
    memcpy(out, r.data, 16);
}

// Model and inverse model matrices exactly as sp_AddObjectToScene forms them.
void ref_object_matrices(const float *pos, const float *q, const float *scale,
    float *model16, float *inv16)
{
    quat qq = {q[0], q[1], q[2], q[3]};
    mat4 m = Translate(V3(pos)) * Rotate(qq) * Scale(V3(scale));
    mat4 mi = Scale(Inverse(V3(scale))) * Rotate(Conjugate(qq)) *
              Translate(-V3(pos));
    memcpy(model16, m.data, 64);
    memcpy(inv16, mi.data, 64);
}

// ---------------------------------------------------------------- meshes
// Writes vertices as 8 floats (position, normal, uv) and indices.
void ref_icosphere(unsigned level, float *verts, unsigned *idx,
    unsigned *nv, unsigned *ni)
{
    MemoryArena arena;
    u64 bytes = Megabytes(8);
    InitializeMemoryArena(&arena, calloc(1, bytes), bytes);
    MeshData md = CreateIcosahedronMesh(level, &arena);
    memcpy(verts, md.vertices, sizeof(VertexPNT) * md.vertexCount);
    memcpy(idx, md.indices, 4 * md.indexCount);
    *nv = md.vertexCount;
    *ni = md.indexCount;
    free(arena.base);
}

void ref_plane(float *verts, unsigned *idx, unsigned *nv, unsigned *ni)
{
    MemoryArena arena;
    u64 bytes = 4096;
    InitializeMemoryArena(&arena, calloc(1, bytes), bytes);
    MeshData md = CreatePlaneMesh(&arena);
    memcpy(verts, md.vertices, sizeof(VertexPNT) * md.vertexCount);
    memcpy(idx, md.indices, 4 * md.indexCount);
    *nv = md.vertexCount;
    *ni = md.indexCount;
    free(arena.base);
}

// ---------------------------------------------------------------- bvh
// Builds a tree over n boxes in the state arena and returns an opaque handle.
void *ref_bvh_create(const float *mins, const float *maxs, unsigned n)
{
    vec3 *mn = AllocateArray(&g_ref->tempArena, vec3, n);
    vec3 *mx = AllocateArray(&g_ref->tempArena, vec3, n);
    for (u32 i = 0; i < n; ++i)
    {
        mn[i] = V3(mins + 3 * i);
        mx[i] = V3(maxs + 3 * i);
    }
    bvh_Tree *tree = AllocateStruct(&g_ref->arena, bvh_Tree);
    *tree = bvh_CreateTree(&g_ref->arena, mn, mx, n);
    return tree;
}

// Returns the visit-ordered leaf indices of a query; *errorOccurred as the
// reference reports it.
unsigned ref_bvh_intersect(void *handle, const float *o, const float *d,
    unsigned maxHits, unsigned *leaves, unsigned *aabbTests,
    unsigned *errorOccurred)
{
    bvh_Tree *tree = (bvh_Tree *)handle;
    bvh_Node **nodes = (bvh_Node **)calloc(maxHits ? maxHits : 1, sizeof(void *));
    bvh_IntersectRayResult r =
        bvh_IntersectRay(tree, V3(o), V3(d), nodes, maxHits);
    for (u32 i = 0; i < r.count; ++i)
        leaves[i] = nodes[i]->leafIndex;
    *aabbTests = r.aabbTestCount;
    *errorOccurred = r.errorOccurred;
    free(nodes);
    return r.count;
}

static u32 RefDumpNode(bvh_Node *node, float *boxes, int *children,
    unsigned *leafIndex, u32 *count, u32 cap)
{
    u32 me = (*count)++;
    if (me >= cap) return me;
    Put3(boxes + 6 * me, node->min);
    Put3(boxes + 6 * me + 3, node->max);
    leafIndex[me] = node->leafIndex;
    for (int i = 0; i < 4; ++i)
    {
        children[4 * me + i] = -1;
    }
    for (int i = 0; i < 4; ++i)
    {
        if (node->children[i])
            children[4 * me + i] = (int)RefDumpNode(
                node->children[i], boxes, children, leafIndex, count, cap);
    }
    return me;
}

// Pre-order dump of the tree: per node 6 floats (min,max), 4 child slots
// (-1 = empty) and the leaf index (U32_MAX for inner nodes).
unsigned ref_bvh_dump(void *handle, float *boxes, int *children,
    unsigned *leafIndex, unsigned cap)
{
    bvh_Tree *tree = (bvh_Tree *)handle;
    u32 count = 0;
    if (tree->root)
        RefDumpNode(tree->root, boxes, children, leafIndex, &count, cap);
    return count;
}

// ---------------------------------------------------------------- scene
unsigned ref_mesh_create(const float *verts8, unsigned nv, const unsigned *idx,
    unsigned ni, int smooth)
{
    u32 m = g_ref->meshCount++;
    g_ref->meshVertices[m] = (VertexPNT *)malloc(sizeof(VertexPNT) * nv);
    memcpy(g_ref->meshVertices[m], verts8, sizeof(VertexPNT) * nv);
    g_ref->meshIndices[m] = (u32 *)malloc(4 * ni);
    memcpy(g_ref->meshIndices[m], idx, 4 * ni);
    g_ref->meshes[m] = sp_CreateMesh(
        g_ref->meshVertices[m], nv, g_ref->meshIndices[m], ni, smooth);
    ResetMemoryArena(&g_ref->tempArena);
    sp_BuildMeshMidphase(&g_ref->meshes[m], &g_ref->arena, &g_ref->tempArena);
    return m;
}

void ref_scene_add_object(unsigned mesh, unsigned material, const float *pos,
    const float *q, const float *scale)
{
    quat qq = {q[0], q[1], q[2], q[3]};
    sp_AddObjectToScene(&g_ref->scene, g_ref->meshes[mesh], material, V3(pos),
        qq, V3(scale));
}

void ref_scene_build(void) { sp_BuildSceneBroadphase(&g_ref->scene); }

// out: t, materialId(bits), normal xyz, uv; metrics12 accumulates.
void ref_scene_intersect(const float *o, const float *d, float *out,
    unsigned long long *metrics12)
{
    sp_Metrics metrics = {};
    sp_RayIntersectSceneResult r =
        sp_RayIntersectScene(&g_ref->scene, V3(o), V3(d), &metrics);
    out[0] = r.t;
    memcpy(out + 1, &r.materialId, 4);
    Put3(out + 2, r.normal);
    out[5] = r.uv.x; out[6] = r.uv.y;
    for (int i = 0; i < SP_MAX_METRICS; ++i) metrics12[i] += metrics.values[i];
}

// Mesh-level query in mesh space (sp_RayIntersectMesh).  out: t,u,v,n.xyz
void ref_mesh_intersect(unsigned mesh, const float *o, const float *d,
    float *out, unsigned long long *metrics12)
{
    sp_Metrics metrics = {};
    sp_RayIntersectMeshResult r =
        sp_RayIntersectMesh(g_ref->meshes[mesh], V3(o), V3(d), &metrics);
    out[0] = r.triangleIntersection.t;
    out[1] = r.triangleIntersection.uv.x;
    out[2] = r.triangleIntersection.uv.y;
    Put3(out + 3, r.triangleIntersection.normal);
    for (int i = 0; i < SP_MAX_METRICS; ++i) metrics12[i] += metrics.values[i];
}

// ---------------------------------------------------------------- materials
int ref_material_register(const float *albedo, unsigned albedoTexture,
    const float *emission, unsigned emissionTexture, float roughness,
    unsigned id)
{
    sp_Material m = {};
    m.albedo = V3(albedo);
    m.albedoTexture = albedoTexture;
    m.emission = V3(emission);
    m.emissionTexture = emissionTexture;
    m.roughness = roughness;
    return sp_RegisterMaterial(&g_ref->materials, m, id);
}

int ref_texture_register(const float *rgba, unsigned w, unsigned h,
    unsigned id)
{
    u32 t = g_ref->textureCount++;
    g_ref->texturePixels[t] = (float *)malloc((u64)w * h * 16);
    memcpy(g_ref->texturePixels[t], rgba, (u64)w * h * 16);
    HdrImage img = {g_ref->texturePixels[t], w, h};
    return sp_RegisterTexture(&g_ref->materials, img, id);
}

void ref_set_background(unsigned id)
{
    g_ref->materials.backgroundMaterialId = id;
}

// path: per vertex 15 floats: materialId(bits), worldPosition, outgoingDir,
// incomingDir, normal, uv(2)  -> 1+3+3+3+3+2 = 15
void ref_compute_radiance_for_path(const float *path, unsigned len, float *out)
{
    sp_PathVertex v[8] = {};
    for (u32 i = 0; i < len && i < 8; ++i)
    {
        const float *p = path + 15 * i;
        memcpy(&v[i].materialId, p, 4);
        v[i].worldPosition = V3(p + 1);
        v[i].outgoingDir = V3(p + 4);
        v[i].incomingDir = V3(p + 7);
        v[i].normal = V3(p + 10);
        v[i].uv = Vec2(p[13], p[14]);
    }
    Put3(out, ComputeRadianceForPath(v, len, &g_ref->materials));
}

void ref_evaluate_material(unsigned id, const float *uv,
    const float *outgoingDir, float *out /* albedo, emission, roughness */)
{
    sp_Material *m = sp_FindMaterialById(&g_ref->materials, id);
    sp_PathVertex v = {};
    v.uv = Vec2(uv[0], uv[1]);
    v.outgoingDir = V3(outgoingDir);
    sp_MaterialOutput r = {};
    if (m) r = sp_EvaluateMaterial(&g_ref->materials, m, &v);
    else r.emission = Vec3(1, 0, 1);
    Put3(out, r.albedo); Put3(out + 3, r.emission); out[6] = r.roughness;
}

// ---------------------------------------------------------------- camera
void ref_camera_configure(unsigned w, unsigned h, const float *pos,
    const float *q, float filmDistance)
{
    free(g_ref->imagePlane.pixels);
    g_ref->imagePlane.pixels = (vec4 *)calloc((u64)w * h, sizeof(vec4));
    g_ref->imagePlane.width = w;
    g_ref->imagePlane.height = h;
    quat qq = {q[0], q[1], q[2], q[3]};
    sp_ConfigureCamera(
        &g_ref->camera, &g_ref->imagePlane, V3(pos), qq, filmDistance);
}

// Camera state: basis(9), position, filmCenter, halfPixelW/H, halfFilmW/H.
void ref_camera_state(float *out)
{
    sp_Camera *c = &g_ref->camera;
    Put3(out, c->basis.right); Put3(out + 3, c->basis.up);
    Put3(out + 6, c->basis.forward); Put3(out + 9, c->position);
    Put3(out + 12, c->filmCenter);
    out[15] = c->halfPixelWidth; out[16] = c->halfPixelHeight;
    out[17] = c->halfFilmWidth; out[18] = c->halfFilmHeight;
}

void ref_film_position(float px, float py, float *out)
{
    vec3 filmP = {};
    vec2 p = Vec2(px, py);
    sp_CalculateFilmPositions(&g_ref->camera, &filmP, &p, 1);
    Put3(out, filmP);
}

// Traces one tile with the given xorshift state (updated in place) and
// samples per pixel; the whole image plane (w*h*4 floats) is copied to
// pixelsOut so that writes outside the tile are visible.
void ref_path_trace_tile(unsigned minX, unsigned minY, unsigned maxX,
    unsigned maxY, unsigned *rngState, unsigned spp,
    unsigned long long *metrics12, float *pixelsOut)
{
    g_refSamplesPerPixel = spp;
    sp_Context ctx = {};
    ctx.camera = &g_ref->camera;
    ctx.scene = &g_ref->scene;
    ctx.materialSystem = &g_ref->materials;
    RandomNumberGenerator rng = {*rngState};
    sp_Metrics metrics = {};
    Tile tile = {minX, minY, maxX, maxY};
    sp_PathTraceTile(&ctx, tile, &rng, &metrics);
    *rngState = rng.state;
    for (int i = 0; i < SP_MAX_METRICS; ++i) metrics12[i] += metrics.values[i];
    if (pixelsOut)
        memcpy(pixelsOut, g_ref->imagePlane.pixels,
            (u64)g_ref->imagePlane.width * g_ref->imagePlane.height * 16);
}

// Per-pixel streams: every pixel of the rectangle is traced as its own 1x1
// tile with the xorshift state seeds[(y-minY)*(maxX-minX)+(x-minX)].
void ref_path_trace_pixels(unsigned minX, unsigned minY, unsigned maxX,
    unsigned maxY, const unsigned *seeds, unsigned spp,
    unsigned long long *metrics12, float *pixelsOut)
{
    g_refSamplesPerPixel = spp;
    sp_Context ctx = {};
    ctx.camera = &g_ref->camera;
    ctx.scene = &g_ref->scene;
    ctx.materialSystem = &g_ref->materials;
    u32 tw = maxX - minX;
    for (u32 y = minY; y < maxY; ++y)
        for (u32 x = minX; x < maxX; ++x)
        {
            RandomNumberGenerator rng = {seeds[(y - minY) * tw + (x - minX)]};
            sp_Metrics metrics = {};
            Tile tile = {x, y, x + 1, y + 1};
            sp_PathTraceTile(&ctx, tile, &rng, &metrics);
            for (int i = 0; i < SP_MAX_METRICS; ++i)
                metrics12[i] += metrics.values[i];
        }
    if (pixelsOut)
        memcpy(pixelsOut, g_ref->imagePlane.pixels,
            (u64)g_ref->imagePlane.width * g_ref->imagePlane.height * 16);
}

} // extern "C"
