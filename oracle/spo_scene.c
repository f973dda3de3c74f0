/* ORACLE -- test infrastructure only (see spo.h).
 * Scene assembly and the two-level closest-hit query (src/sp_scene.cpp),
 * the material system (src/sp_material_system.cpp) and the camera
 * (src/simd_path_tracer.cpp:1-63). */
#include <stdlib.h>
#include <string.h>
#include "spo_math.h"

spo_Scene *spo_scene_create(void)
{
    spo_Scene *s = (spo_Scene *)calloc(1, sizeof *s);
    s->broadphase.root = -1;
    s->maxBroadphaseHits = 32;  /* src/sp_scene.cpp:246 */
    s->maxMidphaseHits = 128;   /* src/sp_scene.cpp:139 */
    s->materialCapacity = 32;   /* src/sp_material_system.h:32 */
    return s;
}

void spo_scene_destroy(spo_Scene *s)
{
    if (!s) return;
    for (uint32_t i = 0; i < s->meshCount; ++i)
    {
        free(s->meshes[i].vertices);
        free(s->meshes[i].indices);
        free(s->meshes[i].midphase.nodes);
        free(s->meshes[i].bvh2Nodes);
        free(s->meshes[i].bvh2Order);
    }
    free(s->meshes);
    free(s->objects);
    free(s->broadphase.nodes);
    for (uint32_t i = 0; i < s->imageCount; ++i)
        free((void *)s->images[i].pixels);
    free(s->envPdf);
    free(s->envProb);
    free(s->envAlias);
    free(s->emitPdfA); free(s->emitProb); free(s->emitAlias);
    free(s->emitPair); free(s->emitBase);
    free(s->emitInst); free(s->emitProbL); free(s->emitSlot); free(s->emitAliasL);
    free(s);
}

/* sp_CreateMesh + sp_BuildMeshMidphase, src/sp_scene.cpp:8-75 (per-triangle
 * AABB = Min/Max of its three vertices).  The arrays are copied. */
int spo_mesh_create(spo_Scene *s, const float *verts8, uint32_t nv,
                    const uint32_t *idx, uint32_t ni, int smooth,
                    int buildMidphase)
{
    if (ni % 3 != 0) return -1;
    if (s->meshCount == s->meshCapacity)
    {
        s->meshCapacity = s->meshCapacity ? 2 * s->meshCapacity : 8;
        s->meshes = (spo_Mesh *)realloc(s->meshes,
                                        sizeof(spo_Mesh) * s->meshCapacity);
    }
    spo_Mesh *m = &s->meshes[s->meshCount];
    memset(m, 0, sizeof *m);
    m->vertices = (spo_Vertex *)malloc(sizeof(spo_Vertex) * (nv ? nv : 1));
    memcpy(m->vertices, verts8, sizeof(spo_Vertex) * nv);
    m->indices = (uint32_t *)malloc(4 * (ni ? ni : 1));
    memcpy(m->indices, idx, 4 * ni);
    m->vertexCount = nv;
    m->indexCount = ni;
    m->useSmoothShading = smooth;
    m->midphase.root = -1;
    if (buildMidphase && ni > 0)
    {
        uint32_t tc = ni / 3;
        float *mn = (float *)malloc(12 * tc), *mx = (float *)malloc(12 * tc);
        for (uint32_t i = 0; i < tc; ++i)
        {
            spo_v3 a = m->vertices[idx[3 * i]].p;
            spo_v3 b = m->vertices[idx[3 * i + 1]].p;
            spo_v3 c = m->vertices[idx[3 * i + 2]].p;
            v3put(mn + 3 * i, v3min(a, v3min(b, c)));
            v3put(mx + 3 * i, v3max(a, v3max(b, c)));
        }
        spo_Bvh *t = spo_bvh_create(mn, mx, tc);
        m->midphase = *t;
        free(t);
        free(mn);
        free(mx);
    }
    return (int)s->meshCount++;
}

int spo_mesh_set_bvh2(spo_Scene *s, uint32_t mesh, const float *nodes16,
                      uint32_t nNodes, const uint32_t *order)
{
    if (mesh >= s->meshCount || nNodes == 0) return -1;
    spo_Mesh *m = &s->meshes[mesh];
    free(m->bvh2Nodes);
    free(m->bvh2Order);
    m->bvh2Nodes = (float *)malloc(64 * (size_t)nNodes);
    memcpy(m->bvh2Nodes, nodes16, 64 * (size_t)nNodes);
    uint32_t nt = m->indexCount / 3;
    m->bvh2Order = (uint32_t *)malloc(4 * (size_t)(nt ? nt : 1));
    memcpy(m->bvh2Order, order, 4 * (size_t)nt);
    m->bvh2Count = nNodes;
    return 0;
}

void spo_scene_set_limits(spo_Scene *s, uint32_t maxBroadphase,
                          uint32_t maxMidphase, uint32_t maxMaterials)
{
    s->maxBroadphaseHits = maxBroadphase;
    s->maxMidphaseHits = maxMidphase;
    s->materialCapacity = maxMaterials < SPO_MAX_MATERIALS ? maxMaterials
                                                           : SPO_MAX_MATERIALS;
}

/* sp_AddObjectToScene, src/sp_scene.cpp:77-117 (ComputeAabb over all mesh
 * vertices, src/sp_scene.cpp:56-74, then TransformAabb). */
int spo_scene_add_object(spo_Scene *s, uint32_t mesh, uint32_t material,
                         const float *pos, const float *q, const float *scale)
{
    if (mesh >= s->meshCount) return -1;
    const spo_Mesh *m = &s->meshes[mesh];
    if (m->vertexCount == 0) return -1;
    if (s->objectCount == s->objectCapacity)
    {
        s->objectCapacity = s->objectCapacity ? 2 * s->objectCapacity : 32;
        s->objects = (spo_Object *)realloc(
            s->objects, sizeof(spo_Object) * s->objectCapacity);
    }
    spo_v3 lo = m->vertices[0].p, hi = lo;
    for (uint32_t i = 1; i < m->vertexCount; ++i)
    {
        lo = v3min(lo, m->vertices[i].p);
        hi = v3max(hi, m->vertices[i].p);
    }
    float bmin[3], bmax[3], box[6];
    v3put(bmin, lo);
    v3put(bmax, hi);
    spo_transform_aabb(bmin, bmax, pos, q, scale, box);
    spo_Object *o = &s->objects[s->objectCount];
    o->mesh = mesh;
    o->material = material;
    spo_object_matrices(pos, q, scale, o->model.m, o->invModel.m);
    o->aabbMin = v3p(box);
    o->aabbMax = v3p(box + 3);
    return (int)s->objectCount++;
}

/* sp_BuildSceneBroadphase, src/sp_scene.cpp:119-125 */
void spo_scene_build(spo_Scene *s)
{
    free(s->broadphase.nodes);
    memset(&s->broadphase, 0, sizeof s->broadphase);
    s->broadphase.root = -1;
    if (s->objectCount == 0) return;
    float *mn = (float *)malloc(12 * s->objectCount);
    float *mx = (float *)malloc(12 * s->objectCount);
    for (uint32_t i = 0; i < s->objectCount; ++i)
    {
        v3put(mn + 3 * i, s->objects[i].aabbMin);
        v3put(mx + 3 * i, s->objects[i].aabbMax);
    }
    spo_Bvh *t = spo_bvh_create(mn, mx, s->objectCount);
    s->broadphase = *t;
    free(t);
    free(mn);
    free(mx);
}

/* ------------------------------------------------------------ materials */
/* sp_RegisterMaterial, src/sp_material_system.cpp:1-13 */
int spo_material_register(spo_Scene *s, const float *albedo,
                          uint32_t albedoTexture, const float *emission,
                          uint32_t emissionTexture, float roughness,
                          uint32_t id)
{
    if (s->materialCount >= s->materialCapacity) return 0;
    uint32_t i = s->materialCount++;
    s->materialKeys[i] = id;
    s->materials[i].albedo = v3p(albedo);
    s->materials[i].albedoTexture = albedoTexture;
    s->materials[i].emission = v3p(emission);
    s->materials[i].emissionTexture = emissionTexture;
    s->materials[i].roughness = roughness;
    return 1;
}

/* sp_RegisterTexture, src/sp_material_system.cpp:43-57 (pixels copied) */
int spo_texture_register(spo_Scene *s, const float *rgba, uint32_t w,
                         uint32_t h, uint32_t id, int isCube)
{
    if (s->imageCount >= SPO_MAX_IMAGES) return 0;
    uint32_t i = s->imageCount++;
    uint64_t n = (uint64_t)w * h * 4 * (isCube ? 6 : 1);
    float *p = (float *)malloc(4 * n);
    memcpy(p, rgba, 4 * n);
    s->imageKeys[i] = id;
    s->images[i].pixels = p;
    s->images[i].width = w;
    s->images[i].height = h;
    s->images[i].isCube = isCube;
    return 1;
}

void spo_set_background(spo_Scene *s, uint32_t id)
{
    s->backgroundMaterialId = id;
}

/* sp_FindMaterialById, src/sp_material_system.cpp:15-28 */
static const spo_Material *find_material(const spo_Scene *s, uint32_t id)
{
    for (uint32_t i = 0; i < s->materialCount; ++i)
        if (s->materialKeys[i] == id) return &s->materials[i];
    return NULL;
}

/* sp_FindTexture, src/sp_material_system.cpp:30-41 */
static const spo_Image *find_texture(const spo_Scene *s, uint32_t id)
{
    for (uint32_t i = 0; i < s->imageCount; ++i)
        if (s->imageKeys[i] == id) return &s->images[i];
    return NULL;
}

/* Vose's alias table of P (n entries, sum 1): stacks filled in ascending
 * order, leftovers keep probability 1. */
static void alias_table(const double *P, size_t n, float *prob, uint32_t *alias)
{
    double *q = (double *)malloc((n ? n : 1) * sizeof(double));
    uint32_t *sm = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    uint32_t *lg = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    size_t ns = 0, nl = 0;
    for (size_t i = 0; i < n; ++i)
    {
        q[i] = P[i] * (double)n;
        prob[i] = 1.0f;
        alias[i] = (uint32_t)i;
        if (q[i] < 1.0) sm[ns++] = (uint32_t)i;
        else lg[nl++] = (uint32_t)i;
    }
    while (ns && nl)
    {
        uint32_t a = sm[--ns], b = lg[--nl];
        prob[a] = (float)q[a];
        alias[a] = b;
        q[b] = (q[b] + q[a]) - 1.0;
        if (q[b] < 1.0) sm[ns++] = b;
        else lg[nl++] = b;
    }
    free(q); free(sm); free(lg);
}

/* mean channel of an object's constant emission, 0 if none or textured */
static double object_emission(const spo_Scene *s, const spo_Object *o)
{
    const spo_Material *m = find_material(s, o->material);
    if (!m || find_texture(s, m->emissionTexture)) return 0.0;
    return ((double)m->emission.x + (double)m->emission.y + (double)m->emission.z) / 3.0;
}

/* Emissive triangles of the scene's objects (next-event estimation toward
 * emitting geometry; no reference counterpart): every triangle of an
 * object with a positive constant emission, in object then index order;
 * weight = mean emission x world-space area (vertices through the
 * column-major model matrix in double); P = weight / sum; pdfA = P / area
 * (0 for a degenerate triangle); alias table; pair = (object, triangle). */
uint32_t spo_scene_emitter_table(const spo_Scene *s, float *pdfA, float *prob,
                                 uint32_t *alias, uint32_t *pair, int32_t *base)
{
    size_t n = 0;
    for (uint32_t i = 0; i < s->objectCount; ++i)
    {
        double lum = object_emission(s, &s->objects[i]);
        base[i] = -1;
        if (lum > 0.0 && isfinite(lum))
            n += s->meshes[s->objects[i].mesh].indexCount / 3;
    }
    if (n == 0) return 0;
    double *wt = (double *)malloc(n * sizeof(double));
    double *area = (double *)malloc(n * sizeof(double));
    size_t g = 0;
    for (uint32_t i = 0; i < s->objectCount; ++i)
    {
        const spo_Object *o = &s->objects[i];
        double lum = object_emission(s, o);
        if (!(lum > 0.0) || !isfinite(lum)) continue;
        const spo_Mesh *me = &s->meshes[o->mesh];
        const float *m = o->model.m;
        base[i] = (int32_t)g;
        for (uint32_t k = 0; k < me->indexCount / 3; ++k, ++g)
        {
            double w[3][3];
            for (int v = 0; v < 3; ++v)
            {
                spo_v3 p = me->vertices[me->indices[3 * k + v]].p;
                for (int r = 0; r < 3; ++r)
                    w[v][r] = (double)m[r] * p.x + (double)m[4 + r] * p.y +
                              (double)m[8 + r] * p.z + (double)m[12 + r];
            }
            double ux = w[1][0] - w[0][0], uy = w[1][1] - w[0][1], uz = w[1][2] - w[0][2];
            double vx = w[2][0] - w[0][0], vy = w[2][1] - w[0][1], vz = w[2][2] - w[0][2];
            double cx = uy * vz - uz * vy, cy = uz * vx - ux * vz, cz = ux * vy - uy * vx;
            double a = 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
            area[g] = a;
            wt[g] = isfinite(a) ? lum * a : 0.0;
            pair[2 * g] = i;
            pair[2 * g + 1] = k;
        }
    }
    double total = 0.0;
    for (size_t j = 0; j < n; ++j) total += wt[j];
    uint32_t count = 0;
    if (total > 0.0 && isfinite(total))
    {
        double *P = (double *)malloc(n * sizeof(double));
        for (size_t j = 0; j < n; ++j)
        {
            P[j] = wt[j] / total;
            pdfA[j] = area[j] > 0.0 ? (float)(P[j] / area[j]) : 0.0f;
        }
        alias_table(P, n, prob, alias);
        free(P);
        count = (uint32_t)n;
    }
    else
        for (uint32_t i = 0; i < s->objectCount; ++i) base[i] = -1;
    free(wt); free(area);
    return count;
}

/* Per emitter instance of the table (base >= 0): world vertices in
 * double through the column-major model matrix (as the table's areas), the
 * instance's bounding box and area sum in triangle order; an alias table
 * over its triangles by area (restates build_emitter_instances of the
 * product's spg_env_table.h). */
uint32_t spo_scene_emitter_instances(const spo_Scene *s, const int32_t *base,
                                     uint32_t entries, float *rec, int32_t *slot,
                                     float *probL, uint32_t *aliasL)
{
    uint32_t ne = 0;
    for (uint32_t g = 0; g < entries; ++g) { probL[g] = 1.0f; aliasL[g] = 0u; }
    for (uint32_t i = 0; i < s->objectCount; ++i)
    {
        slot[i] = -1;
        if (base[i] < 0) continue;
        const spo_Object *o = &s->objects[i];
        const spo_Mesh *me = &s->meshes[o->mesh];
        const float *m = o->model.m;
        const uint32_t nt = me->indexCount / 3;
        double *area = (double *)malloc((nt ? nt : 1) * sizeof(double));
        double total = 0.0, lo[3] = {INFINITY, INFINITY, INFINITY},
               hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t k = 0; k < nt; ++k)
        {
            double w[3][3];
            for (int v = 0; v < 3; ++v)
            {
                spo_v3 p = me->vertices[me->indices[3 * k + v]].p;
                for (int r = 0; r < 3; ++r)
                {
                    w[v][r] = (double)m[r] * p.x + (double)m[4 + r] * p.y +
                              (double)m[8 + r] * p.z + (double)m[12 + r];
                    if (w[v][r] < lo[r]) lo[r] = w[v][r];
                    if (w[v][r] > hi[r]) hi[r] = w[v][r];
                }
            }
            double ux = w[1][0] - w[0][0], uy = w[1][1] - w[0][1], uz = w[1][2] - w[0][2];
            double vx = w[2][0] - w[0][0], vy = w[2][1] - w[0][1], vz = w[2][2] - w[0][2];
            double cx = uy * vz - uz * vy, cy = uz * vx - ux * vz, cz = ux * vy - uy * vx;
            double a = 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
            area[k] = isfinite(a) ? a : 0.0;
            total += area[k];
        }
        if (!(total > 0.0) || !isfinite(total)) { free(area); continue; }
        double *P = (double *)malloc((nt ? nt : 1) * sizeof(double));
        for (uint32_t k = 0; k < nt; ++k) P[k] = area[k] / total;
        alias_table(P, nt, probL + base[i], aliasL + base[i]);
        free(P);
        free(area);
        if (ne < SPO_MAX_SPATIAL_EMITTERS)
        {
            double dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
            uint32_t first = (uint32_t)base[i], count = nt;
            float *r8 = rec + 8 * (size_t)ne;
            r8[0] = (float)(0.5 * (lo[0] + hi[0]));
            r8[1] = (float)(0.5 * (lo[1] + hi[1]));
            r8[2] = (float)(0.5 * (lo[2] + hi[2]));
            r8[3] = (float)(0.5 * sqrt(dx * dx + dy * dy + dz * dz));
            r8[4] = (float)(object_emission(s, o) * total);
            r8[5] = (float)(1.0 / total);
            memcpy(&r8[6], &first, 4);
            memcpy(&r8[7], &count, 4);
        }
        slot[i] = (int32_t)ne;
        ++ne;
    }
    if (ne > SPO_MAX_SPATIAL_EMITTERS)
    {
        for (uint32_t i = 0; i < s->objectCount; ++i) slot[i] = -1;
        ne = 0;
    }
    return ne;
}

int spo_scene_build_env(spo_Scene *s)
{
    free(s->emitPdfA); free(s->emitProb); free(s->emitAlias);
    free(s->emitPair); free(s->emitBase);
    free(s->emitInst); free(s->emitProbL); free(s->emitSlot); free(s->emitAliasL);
    s->emitPdfA = s->emitProb = NULL;
    s->emitAlias = s->emitPair = NULL;
    s->emitBase = NULL;
    s->emitCount = 0;
    {
        size_t n = 0;
        for (uint32_t i = 0; i < s->objectCount; ++i)
            n += s->meshes[s->objects[i].mesh].indexCount / 3;
        size_t nn = n ? n : 1;
        s->emitPdfA = (float *)malloc(4 * nn);
        s->emitProb = (float *)malloc(4 * nn);
        s->emitAlias = (uint32_t *)malloc(4 * nn);
        s->emitPair = (uint32_t *)malloc(8 * nn);
        s->emitBase = (int32_t *)malloc(4 * (s->objectCount ? s->objectCount : 1));
        s->emitCount = spo_scene_emitter_table(s, s->emitPdfA, s->emitProb,
                                               s->emitAlias, s->emitPair, s->emitBase);
        s->emitInst = (float *)malloc(4 * 8 * SPO_MAX_SPATIAL_EMITTERS);
        s->emitSlot = (int32_t *)malloc(4 * (s->objectCount ? s->objectCount : 1));
        s->emitProbL = (float *)malloc(4 * nn);
        s->emitAliasL = (uint32_t *)malloc(4 * nn);
        s->emitInstCount =
            s->emitCount ? spo_scene_emitter_instances(s, s->emitBase, s->emitCount,
                                                       s->emitInst, s->emitSlot,
                                                       s->emitProbL, s->emitAliasL)
                         : 0;
    }
    s->envSelect = 0.0f;
    free(s->envPdf); free(s->envProb); free(s->envAlias);
    s->envPdf = s->envProb = NULL;
    s->envAlias = NULL;
    const spo_Material *m = find_material(s, s->backgroundMaterialId);
    const spo_Image *t = m ? find_texture(s, m->emissionTexture) : NULL;
    if (!t || !t->isCube) return 0;
    size_t n = 6 * (size_t)t->width * t->height;
    s->envPdf = (float *)malloc(4 * n);
    s->envProb = (float *)malloc(4 * n);
    s->envAlias = (uint32_t *)malloc(4 * n);
    if (spo_env_table_build(t->pixels, t->width, t->height, s->envPdf,
                            s->envProb, s->envAlias) < 0)
    {
        free(s->envPdf); free(s->envProb); free(s->envAlias);
        s->envPdf = s->envProb = NULL;
        s->envAlias = NULL;
        return 0;
    }
    s->envW = t->width;
    s->envH = t->height;
    s->envSelect = s->emitCount ? 0.5f : 1.0f;
    return 1;
}

/* The sky share of the light samples when the scene has both kinds (the
 * product's learned light split, spg_guide_set_light_split, fixed for a
 * comparison); no effect otherwise.  Call after spo_scene_build_env. */
void spo_scene_set_light_split(spo_Scene *s, float skyFraction)
{
    if (s->envPdf && s->emitCount) s->envSelect = skyFraction;
}

/* sp_EvaluateMaterial, src/sp_material_system.cpp:59-105, plus the
 * cube-map emission texture this build adds (sampled by direction).  A
 * missing material yields emission (1,0,1) and zero albedo/roughness
 * (src/simd_path_tracer.cpp:125-134).  out7 = albedo, emission, roughness */
void spo_material_eval_internal(const spo_Scene *s, uint32_t id, spo_v2 uv,
                                spo_v3 outgoingDir, int tx, float *out7)
{
    memset(out7, 0, 7 * sizeof(float));
    const spo_Material *m = find_material(s, id);
    if (!m)
    {
        out7[3] = 1.0f; out7[4] = 0.0f; out7[5] = 1.0f;
        return;
    }
    const spo_Image *albedoTex = find_texture(s, m->albedoTexture);
    const spo_Image *emissionTex = find_texture(s, m->emissionTexture);
    if (albedoTex && !albedoTex->isCube)
    {
        float t[4], uvv[2] = {uv.x, uv.y};
        spo_sample_nearest(albedoTex->pixels, albedoTex->width,
                           albedoTex->height, uvv, t);
        out7[0] = t[0]; out7[1] = t[1]; out7[2] = t[2];
    }
    else
        v3put(out7, m->albedo);
    if (emissionTex)
    {
        spo_v3 dir = v3neg(outgoingDir);
        float dd[3], t[4];
        v3put(dd, dir);
        if (emissionTex->isCube)
            spo_cubemap_lookup(emissionTex->pixels, emissionTex->width,
                               emissionTex->height, dd, t);
        else
        {
            float sc[2], e[2];
            spo_to_spherical(dd, tx, sc);
            spo_map_to_equirect(sc, tx, e);
            e[1] = 1.0f - e[1];
            spo_sample_nearest(emissionTex->pixels, emissionTex->width,
                               emissionTex->height, e, t);
        }
        out7[3] = t[0] * 1.0f; out7[4] = t[1] * 1.0f; out7[5] = t[2] * 1.0f;
    }
    else
        v3put(out7 + 3, m->emission);
    out7[6] = m->roughness;
}

void spo_evaluate_material(const spo_Scene *s, uint32_t id, const float *uv,
                           const float *outgoingDir, int tx, float *out7)
{
    spo_v2 u = {uv[0], uv[1]};
    spo_material_eval_internal(s, id, u, v3p(outgoingDir), tx, out7);
}

/* -------------------------------------------------------------- camera */
/* sp_ConfigureCamera, src/simd_path_tracer.cpp:1-36 */
void spo_camera_configure(spo_Camera *c, uint32_t w, uint32_t h,
                          const float *pos, const float *q,
                          float filmDistance)
{
    spo_q qq = qmake(q);
    memset(c, 0, sizeof *c);
    c->width = w;
    c->height = h;
    c->position = v3p(pos);
    c->right = qrotate(v3(1, 0, 0), qq);
    c->up = qrotate(v3(0, 1, 0), qq);
    c->forward = qrotate(v3(0, 0, -1), qq);
    c->filmCenter = v3add(c->position, v3mul(c->forward, filmDistance));
    c->halfPixelWidth = 0.5f / (float)w;
    c->halfPixelHeight = 0.5f / (float)h;
    float filmWidth = 1.0f, filmHeight = 1.0f;
    if (w > h) filmHeight = (float)h / (float)w;
    else if (w < h) filmWidth = (float)w / (float)h;
    c->halfFilmWidth = 0.5f * filmWidth;
    c->halfFilmHeight = 0.5f * filmHeight;
}

/* sp_CalculateFilmPositions, src/simd_path_tracer.cpp:38-63 */
void spo_film_position(const spo_Camera *c, float px, float py, float *out3)
{
    float fx = px / (float)c->width;
    float fy = py / (float)c->height;
    fy = 1.0f - fy;
    fx = fx * 2.0f - 1.0f;
    fy = fy * 2.0f - 1.0f;
    spo_v3 p = v3mul(c->right, c->halfFilmWidth * fx);
    p = v3add(p, v3mul(c->up, c->halfFilmHeight * fy));
    p = v3add(p, c->filmCenter);
    v3put(out3, p);
}

/* ------------------------------------------------------------ queries */
typedef struct { float t; spo_v2 uv; spo_v3 normal; uint32_t prim; } mesh_hit;

/* The per-triangle step of sp_RayIntersectMesh, src/sp_scene.cpp:161-214:
 * Moller-Trumbore, keep t > 0 and strictly smaller; UV and (optionally)
 * smooth normal from the barycentrics. */
static void consider_triangle(const spo_Mesh *m, uint32_t tri, const float *o,
                              const float *d, mesh_hit *best)
{
    const spo_Vertex *v0 = &m->vertices[m->indices[3 * tri]];
    const spo_Vertex *v1 = &m->vertices[m->indices[3 * tri + 1]];
    const spo_Vertex *v2 = &m->vertices[m->indices[3 * tri + 2]];
    float a[3], b[3], c[3], r[6];
    v3put(a, v0->p); v3put(b, v1->p); v3put(c, v2->p);
    spo_rt_mt(o, d, a, b, c, r);
    if (r[0] > 0.0f)
    {
        /* the reference keeps the first of equal t; walks in an order other
         * than ascending index keep the lower index instead */
        if (r[0] < best->t || best->t < 0.0f ||
            (m->bvh2Nodes && r[0] == best->t && tri < best->prim))
        {
            best->t = r[0];
            best->normal = v3(r[3], r[4], r[5]);
            best->prim = tri;
            float u = r[1], v = r[2];
            float w = 1.0f - u - v;
            spo_v2 uv;
            uv.x = (v0->uv.x * w + v1->uv.x * u) + v2->uv.x * v;
            uv.y = (v0->uv.y * w + v1->uv.y * u) + v2->uv.y * v;
            if (m->useSmoothShading)
            {
                spo_v3 n = v3add(v3add(v3mul(v0->n, w), v3mul(v1->n, u)),
                                 v3mul(v2->n, v));
                best->normal = v3norm(n);
            }
            best->uv = uv;
        }
    }
}

static int32_t node_ref(const float *f)
{
    int32_t r;
    memcpy(&r, f, 4);
    return r;
}

static int bvh2_slab(const float *lo, const float *hi, spo_v3 o, spo_v3 inv,
                     float *entry)
{
    float t0x = (lo[0] - o.x) * inv.x, t1x = (hi[0] - o.x) * inv.x;
    float t0y = (lo[1] - o.y) * inv.y, t1y = (hi[1] - o.y) * inv.y;
    float t0z = (lo[2] - o.z) * inv.z, t1z = (hi[2] - o.z) * inv.z;
    float mnx = t0x < t1x ? t0x : t1x, mxx = t0x > t1x ? t0x : t1x;
    float mny = t0y < t1y ? t0y : t1y, mxy = t0y > t1y ? t0y : t1y;
    float mnz = t0z < t1z ? t0z : t1z, mxz = t0z > t1z ? t0z : t1z;
    float e = spo_max(0.0f, spo_max(mnx, spo_max(mny, mnz)));
    *entry = e;
    return e <= spo_min(mxx, spo_min(mxy, mxz));
}

/* Closest-hit walk of an attached BVH2 (the GPU's traversal, spo_bvh2.c),
 * candidates scored by consider_triangle. */
static void mesh_query_bvh2(const spo_Mesh *m, const float *o, const float *d,
                            mesh_hit *best, uint64_t *metrics12)
{
    spo_v3 ro = v3p(o), inv = v3inv(v3p(d));
    int32_t stack[128];
    int sp = 0;
    int32_t node = 0;
    uint64_t tests = 0;
    for (;;)
    {
        const float *N = m->bvh2Nodes + 16 * (size_t)node;
        int32_t r0 = node_ref(N + 3), r1 = node_ref(N + 7);
        float tc = best->t < 0.0f ? INFINITY : best->t * (1.0f + 1.0f / 4096.0f);
        float e0 = 0, e1 = 0;
        int h0 = bvh2_slab(N, N + 4, ro, inv, &e0) && e0 <= tc;
        int h1 = r1 != 0x7FFFFFFF && bvh2_slab(N + 8, N + 12, ro, inv, &e1) && e1 <= tc;
        tests += r1 != 0x7FFFFFFF ? 2 : 1;
        for (int c = 0; c < 2; ++c)
        {
            int32_t rr = c == 0 ? r0 : r1;
            int h = c == 0 ? h0 : h1;
            if (h && rr < 0)
            {
                uint32_t v = (uint32_t)~rr, first = v >> 3, count = (v & 7u) + 1u;
                for (uint32_t k = 0; k < count; ++k)
                    consider_triangle(m, m->bvh2Order[first + k], o, d, best);
                if (c == 0) h0 = 0;
                else h1 = 0;
            }
        }
        if (h0 && h1)
        {
            int32_t nearN = r0, farN = r1;
            if (e1 < e0) { nearN = r1; farN = r0; }
            stack[sp++] = farN;
            node = nearN;
        }
        else if (h0) node = r0;
        else if (h1) node = r1;
        else
        {
            if (sp == 0) break;
            node = stack[--sp];
        }
    }
    if (metrics12) metrics12[10] += tests;
}

/* sp_RayIntersectMesh, src/sp_scene.cpp:127-227.  Returns -1 if the midphase
 * query overflows (the reference asserts, src/sp_scene.cpp:158). */
static int mesh_query(const spo_Scene *s, const spo_Mesh *m, const float *o,
                      const float *d, int hitMode, mesh_hit *best,
                      uint64_t *metrics12)
{
    best->t = -1.0f;
    best->prim = 0xFFFFFFFFu;
    best->uv.x = best->uv.y = 0.0f;
    best->normal = v3(0, 0, 0);
    if (hitMode == SPO_HIT_BRUTE)
    {
        for (uint32_t tri = 0; tri < m->indexCount / 3; ++tri)
            consider_triangle(m, tri, o, d, best);
        return 0;
    }
    if (m->bvh2Nodes)
    {
        mesh_query_bvh2(m, o, d, best, metrics12);
        return 0;
    }
    uint32_t leaves[1024];
    uint32_t cap = s->maxMidphaseHits < 1024 ? s->maxMidphaseHits : 1024;
    uint32_t tests = 0, err = 0;
    uint32_t n = spo_bvh_intersect(&m->midphase, o, d, cap, leaves, &tests, &err);
    if (metrics12) metrics12[10] += tests;
    if (err) return -1;
    for (uint32_t i = 0; i < n; ++i) consider_triangle(m, leaves[i], o, d, best);
    return 0;
}

float spo_mesh_intersect(const spo_Scene *s, uint32_t mesh, const float *o,
                         const float *d, int hitMode, float *out6,
                         uint32_t *prim, uint64_t *metrics12)
{
    mesh_hit h;
    if (mesh_query(s, &s->meshes[mesh], o, d, hitMode, &h, metrics12) < 0)
        return -2.0f;
    if (out6)
    {
        out6[0] = h.t; out6[1] = h.uv.x; out6[2] = h.uv.y;
        out6[3] = h.normal.x; out6[4] = h.normal.y; out6[5] = h.normal.z;
    }
    if (prim) *prim = h.prim;
    return h.t;
}

/* sp_RayIntersectScene, src/sp_scene.cpp:229-339.  Returns -1 on a BVH
 * overflow (the reference asserts). */
int spo_scene_intersect(const spo_Scene *s, const float *o, const float *d,
                        int hitMode, int tx, spo_SceneHit *out,
                        uint64_t *metrics12)
{
    (void)tx;
    out->t = -1.0f;
    out->materialId = 0;
    out->normal = v3(0, 0, 0);
    out->uv.x = out->uv.y = 0.0f;
    out->object = out->prim = 0xFFFFFFFFu;
    uint32_t objs[1024];
    uint32_t n;
    if (hitMode == SPO_HIT_BRUTE)
    {
        n = s->objectCount < 1024 ? s->objectCount : 1024;
        for (uint32_t i = 0; i < n; ++i) objs[i] = i;
    }
    else
    {
        uint32_t err = 0, tests = 0;
        uint32_t cap = s->maxBroadphaseHits < 1024 ? s->maxBroadphaseHits : 1024;
        n = spo_bvh_intersect(&s->broadphase, o, d, cap, objs, &tests, &err);
        if (err) return -1;
    }
    spo_v3 ro = v3p(o), rd = v3p(d);
    for (uint32_t i = 0; i < n; ++i)
    {
        const spo_Object *obj = &s->objects[objs[i]];
        spo_v3 lo = m4point(obj->invModel, ro);
        spo_v3 ld = v3norm(m4vector(obj->invModel, rd));
        float lof[3], ldf[3];
        v3put(lof, lo);
        v3put(ldf, ld);
        mesh_hit h;
        if (mesh_query(s, &s->meshes[obj->mesh], lof, ldf, hitMode, &h,
                       metrics12) < 0)
            return -1;
        if (metrics12) metrics12[11] += 1;
        if (h.t >= 0.0f)
        {
            spo_v3 localHit = v3add(lo, v3mul(ld, h.t));
            spo_v3 worldHit = m4point(obj->model, localHit);
            float t = v3dot(v3sub(worldHit, ro), rd);
            spo_v3 worldNormal = v3norm(m4vector(obj->model, h.normal));
            if (t < out->t || out->t < 0.0f)
            {
                out->t = t;
                out->materialId = obj->material;
                out->normal = worldNormal;
                out->uv = h.uv;
                out->object = objs[i];
                out->prim = h.prim;
            }
        }
    }
    return 0;
}

/* Occlusion query of a shadow ray toward the sky: 1 if some object's
 * closest mesh hit lies at a positive world t (the instance visit order
 * does not matter), 0 if none, -1 on a broadphase overflow.  The GPU's
 * occlusion walk stops at the first such instance
 * (vk_cinematic_amd/csrc/spg_traverse.h, anyHit). */
int spo_scene_occluded(const spo_Scene *s, const float *o, const float *d, int hitMode,
                       uint64_t *metrics12)
{
    uint32_t objs[1024];
    uint32_t n;
    if (hitMode == SPO_HIT_BRUTE)
    {
        n = s->objectCount < 1024 ? s->objectCount : 1024;
        for (uint32_t i = 0; i < n; ++i) objs[i] = i;
    }
    else
    {
        uint32_t err = 0, tests = 0;
        uint32_t cap = s->maxBroadphaseHits < 1024 ? s->maxBroadphaseHits : 1024;
        n = spo_bvh_intersect(&s->broadphase, o, d, cap, objs, &tests, &err);
        if (err) return -1;
    }
    spo_v3 ro = v3p(o), rd = v3p(d);
    for (uint32_t i = 0; i < n; ++i)
    {
        const spo_Object *obj = &s->objects[objs[i]];
        spo_v3 lo = m4point(obj->invModel, ro);
        spo_v3 ld = v3norm(m4vector(obj->invModel, rd));
        float lof[3], ldf[3];
        v3put(lof, lo);
        v3put(ldf, ld);
        mesh_hit h;
        if (mesh_query(s, &s->meshes[obj->mesh], lof, ldf, hitMode, &h, metrics12) < 0)
            return -1;
        if (metrics12) metrics12[11] += 1;
        if (h.t >= 0.0f)
        {
            spo_v3 worldHit = m4point(obj->model, v3add(lo, v3mul(ld, h.t)));
            if (v3dot(v3sub(worldHit, ro), rd) > 0.0f) return 1;
        }
    }
    return 0;
}
