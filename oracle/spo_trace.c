/* ORACLE -- test infrastructure only (see spo.h).
 * sp_PathTraceTile and ComputeRadianceForPath (src/simd_path_tracer.cpp),
 * the threaded tile scheduler (src/main.cpp:731-844) used as the CPU
 * baseline, and brute-force closest-hit batches (FP32 and FP64). */
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "spo_math.h"

void spo_material_eval_internal(const spo_Scene *s, uint32_t id, spo_v2 uv,
                                spo_v3 outgoingDir, int tx, float *out7);

/* ---------------------------------------------------------------- RNG */
typedef struct { int mode; uint32_t xs; uint64_t state, inc; } spo_Rng;

static uint32_t rng_next(spo_Rng *r)
{
    if (r->mode == SPO_RNG_PCG) return spo_pcg32_next(&r->state, r->inc);
    return spo_xorshift32(&r->xs);
}
/* RandomUnilateral / RandomBilateral, src/math_utils.h:200-214, fed by the
 * active generator. */
static float rng_uni(spo_Rng *r)
{
    uint32_t x = rng_next(r);
    float numerator = (float)(x >> 1);
    float denom = (float)(0xFFFFFFFFu >> 1);
    return numerator / denom;
}
static float rng_bi(spo_Rng *r) { return -1.0f + 2.0f * rng_uni(r); }
/* uniform in [0,1) with 24 random bits (unbiased estimator only) */
static float rng_u24(spo_Rng *r)
{
    return (float)(rng_next(r) >> 8) * (1.0f / 16777216.0f);
}

/* --------------------------------------------------------------- BRDF */
/* src/simd_path_tracer.cpp:65-105 */
static spo_v3 fresnel_schlick(float cosTheta, spo_v3 F0, int tx)
{
    float p = spo_pow5(1.0f - cosTheta, tx);
    return v3add(F0, v3mul(v3sub(v3s(1.0f), F0), p));
}
static float distribution_ggx(spo_v3 N, spo_v3 H, float roughness)
{
    float a = roughness * roughness;
    float a2 = a * a;
    float NdotH = spo_max(v3dot(N, H), 0.0f);
    float NdotH2 = NdotH * NdotH;
    float num = a2;
    float denom = (NdotH2 * (a2 - 1.0f) + 1.0f);
    denom = SPO_PI * denom * denom;
    return num / denom;
}
static float geometry_schlick_ggx(float NdotV, float roughness)
{
    float r = (roughness + 1.0f);
    float k = (r * r) / 8.0f;
    float num = NdotV;
    float denom = NdotV * (1.0f - k) + k;
    return num / denom;
}
static float geometry_smith(spo_v3 N, spo_v3 V, spo_v3 L, float roughness)
{
    float NdotV = spo_max(v3dot(N, V), 0.0f);
    float NdotL = spo_max(v3dot(N, L), 0.0f);
    float ggx2 = geometry_schlick_ggx(NdotV, roughness);
    float ggx1 = geometry_schlick_ggx(NdotL, roughness);
    return ggx1 * ggx2;
}

/* The bracket of ComputeRadianceForPath, src/simd_path_tracer.cpp:136-165:
 * X = kD*albedo/pi + specular; the caller forms E + (X o Lin) * cosine.
 * With noDeltaSpecular the GGX term is dropped at roughness 0, where the
 * reference's D is 0 almost everywhere and 0/0 on the measure-zero set
 * N.H = 1 (used by the unbiased estimator only). */
static spo_v3 brdf_bracket(spo_v3 N, spo_v3 V, spo_v3 L, const float *mat,
                           int tx, int noDeltaSpecular)
{
    spo_v3 H = v3norm(v3add(L, V));
    spo_v3 F0 = v3s(0.04f);
    spo_v3 F = fresnel_schlick(spo_max(v3dot(H, V), 0.0f), F0, tx);
    spo_v3 kS = F;
    spo_v3 kD = v3sub(v3s(1.0f), kS);
    float oneOverPI = 1.0f / SPO_PI;
    float roughness = mat[6];
    spo_v3 specular = v3(0, 0, 0);
    if (!(noDeltaSpecular && roughness == 0.0f))
    {
        float NDF = distribution_ggx(N, H, roughness);
        float G = geometry_smith(N, V, L, roughness);
        spo_v3 numerator = v3mul(F, NDF * G);
        float denominator =
            4.0f * spo_max(v3dot(N, V), 0.0f) * spo_max(v3dot(N, L), 0.0f) +
            0.0001f;
        specular = v3mul(numerator, 1.0f / denominator);
    }
    spo_v3 albedo = v3p(mat);
    return v3add(v3mul(v3had(kD, albedo), oneOverPI), specular);
}

typedef struct {
    uint32_t materialId;
    spo_v3 worldPosition, outgoingDir, incomingDir, normal;
    spo_v2 uv;
} spo_PathVertex; /* sp_PathVertex, src/sp_material_system.h:22-30 */

/* ComputeRadianceForPath, src/simd_path_tracer.cpp:107-175 */
static spo_v3 radiance_for_path(const spo_Scene *s, const spo_PathVertex *path,
                                uint32_t len, int tx, float clampValue)
{
    spo_v3 radiance = v3(0, 0, 0);
    for (int32_t i = (int32_t)len - 1; i >= 0; i--)
    {
        const spo_PathVertex *v = &path[i];
        float mat[7];
        spo_material_eval_internal(s, v->materialId, v->uv, v->outgoingDir, tx,
                                   mat);
        float cosine = spo_max(0.0f, v3dot(v->normal, v->incomingDir));
        spo_v3 incoming = radiance;
        if (clampValue > 0.0f) incoming = v3clamp(incoming, 0.0f, clampValue);
        spo_v3 X = brdf_bracket(v->normal, v->outgoingDir, v->incomingDir, mat,
                                tx, 0);
        radiance = v3add(v3p(mat + 3), v3mul(v3had(X, incoming), cosine));
    }
    return radiance;
}

/* path15: per vertex materialId(bits), worldPosition, outgoingDir,
 * incomingDir, normal, uv */
void spo_compute_radiance_for_path(const spo_Scene *s, const float *path15,
                                   uint32_t len, int tx, float clampValue,
                                   float *out3)
{
    spo_PathVertex v[8];
    memset(v, 0, sizeof v);
    for (uint32_t i = 0; i < len && i < 8; ++i)
    {
        const float *p = path15 + 15 * i;
        memcpy(&v[i].materialId, p, 4);
        v[i].worldPosition = v3p(p + 1);
        v[i].outgoingDir = v3p(p + 4);
        v[i].incomingDir = v3p(p + 7);
        v[i].normal = v3p(p + 10);
        v[i].uv.x = p[13];
        v[i].uv.y = p[14];
    }
    v3put(out3, radiance_for_path(s, v, len, tx, clampValue));
}

/* RandomDirectionOnHemisphere, src/math_lib.h:932-946, on the active RNG */
static spo_v3 hemisphere(spo_v3 normal, spo_Rng *rng, int tx)
{
    float theta = SPO_PI * rng_uni(rng);
    float phi = SPO_PI * rng_bi(rng);
    spo_v3 dir;
    dir.x = spo_sin(theta, tx) * spo_cos(phi, tx);
    dir.z = spo_sin(theta, tx) * spo_sin(phi, tx);
    dir.y = spo_cos(theta, tx);
    if (v3dot(dir, normal) < 0.0f) dir = v3neg(dir);
    return dir;
}

/* Cosine-weighted direction about n (unbiased estimator; no reference
 * counterpart): r = sqrt(u1), phi = 2 pi u2, z = sqrt(max(0, 1 - u1)), in the
 * orthonormal frame of Duff et al. 2017 ("Building an Orthonormal Basis,
 * Revisited"). */
static spo_v3 cosine_direction(spo_v3 n, float u1, float u2, int tx)
{
    float r = sqrtf(u1);
    float phi = 6.28318548f * u2;
    float lx = r * spo_cos(phi, tx);
    float ly = r * spo_sin(phi, tx);
    float lz = sqrtf(spo_max(0.0f, 1.0f - u1));
    float sign = copysignf(1.0f, n.z);
    float a = -1.0f / (sign + n.z);
    float b = n.x * n.y * a;
    spo_v3 t1 = v3(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
    spo_v3 t2 = v3(b, sign + n.y * n.y * a, -n.y);
    return v3add(v3add(v3mul(t1, lx), v3mul(t2, ly)), v3mul(n, lz));
}

/* Probability of the GGX lobe on rough materials (unbiased estimator). */
#define SPEC_PROB 0.25f

/* GGX half-vector sampling (Walter et al. 2007): cos^2 theta_h = (1 - u1) /
 * (1 + (a^2 - 1) u1), a = roughness^2, phi = 2 pi u2, L = 2 (V.H) H - V. */
static spo_v3 ggx_direction(spo_v3 n, spo_v3 V, float roughness, float u1,
                            float u2, int tx)
{
    float a = roughness * roughness;
    float a2 = a * a;
    float c2 = (1.0f - u1) / (1.0f + (a2 - 1.0f) * u1);
    float ct = sqrtf(c2);
    float st = sqrtf(spo_max(0.0f, 1.0f - c2));
    float phi = 6.28318548f * u2;
    float hx = st * spo_cos(phi, tx);
    float hy = st * spo_sin(phi, tx);
    float sign = copysignf(1.0f, n.z);
    float aa = -1.0f / (sign + n.z);
    float b = n.x * n.y * aa;
    spo_v3 t1 = v3(1.0f + sign * n.x * n.x * aa, sign * b, -sign * n.x);
    spo_v3 t2 = v3(b, sign + n.y * n.y * aa, -n.y);
    spo_v3 H = v3add(v3add(v3mul(t1, hx), v3mul(t2, hy)), v3mul(n, ct));
    float vh = v3dot(V, H);
    return v3sub(v3mul(H, 2.0f * vh), V);
}

/* Density of ggx_direction: D(H) (N.H) / (4 V.H), H = normalize(L + V). */
static float ggx_pdf(spo_v3 n, spo_v3 V, spo_v3 L, float roughness)
{
    spo_v3 H = v3norm(v3add(L, V));
    float a = roughness * roughness;
    float a2 = a * a;
    float NdotH = spo_max(v3dot(n, H), 0.0f);
    float NdotH2 = NdotH * NdotH;
    float denom = (NdotH2 * (a2 - 1.0f) + 1.0f);
    denom = SPO_PI * denom * denom;
    float D = a2 / denom;
    float VdotH = v3dot(V, H);
    return VdotH > 0.0f ? D * NdotH / (4.0f * VdotH) : 0.0f;
}

void spo_ggx_direction(const float *n, const float *V, float roughness, float u1,
                       float u2, int tx, float *out3)
{
    v3put(out3, ggx_direction(v3p(n), v3p(V), roughness, u1, u2, tx));
}

float spo_ggx_pdf(const float *n, const float *V, const float *L, float roughness)
{
    return ggx_pdf(v3p(n), v3p(V), v3p(L), roughness);
}

/* ------------------------------------------- guided estimator (FP32) */
/* Its vMF density and sampling as the product evaluates them
 * (vk_cinematic_amd/csrc/spg_guide.h: vmf_pdf, vmf_sample, mixture_pdf,
 * mixture_sample): kappa / (2 pi (1 - e^-2kappa)) e^(kappa (mu.w - 1));
 * cos = 1 + log(u1 + (1 - u1) e^-2kappa) / kappa (Jakob 2012) with the
 * azimuth 2 pi u2 in the Duff et al. frame about mu. */
static float vmf_pdf_f(float kappa, const float *mu, spo_v3 w, int tx)
{
    float d = mu[0] * w.x + mu[1] * w.y + mu[2] * w.z;
    return kappa / (6.28318548f * -spo_expm1(-2.0f * kappa, tx)) *
           spo_exp(kappa * (d - 1.0f), tx);
}

static float mixture_pdf_f(const float *L, spo_v3 w, int tx)
{
    float q = 0.0f;
    for (int k = 0; k < 8; ++k) q += L[k] * vmf_pdf_f(L[8 + k], L + 16 + 3 * k, w, tx);
    return q;
}

static spo_v3 vmf_sample_f(float kappa, const float *mu, float u1, float u2, int tx)
{
    float c = 1.0f + spo_log(u1 + (1.0f - u1) * spo_exp(-2.0f * kappa, tx), tx) / kappa;
    c = fminf(fmaxf(c, -1.0f), 1.0f);
    float s = sqrtf(fmaxf(0.0f, 1.0f - c * c));
    float phi = 6.28318548f * u2;
    float sp = spo_sin(phi, tx), cp = spo_cos(phi, tx);
    spo_v3 m = v3(mu[0], mu[1], mu[2]);
    float sign = copysignf(1.0f, m.z);
    float a = -1.0f / (sign + m.z);
    float b = m.x * m.y * a;
    spo_v3 t1 = v3(1.0f + sign * m.x * m.x * a, sign * b, -sign * m.x);
    spo_v3 t2 = v3(b, sign + m.y * m.y * a, -m.y);
    return v3add(v3add(v3mul(t1, s * cp), v3mul(t2, s * sp)), v3mul(m, c));
}

static spo_v3 mixture_sample_f(const float *L, float u0, float u1, float u2, int tx)
{
    int k = 7;
    float c = 0.0f;
    for (int j = 0; j < 7; ++j)
    {
        c += L[j];
        if (u0 < c) { k = j; break; }
    }
    return vmf_sample_f(L[8 + k], L + 16 + 3 * k, u1, u2, tx);
}

static spo_v3 guided_scatter(const float *L, spo_v3 N, spo_v3 V, float rough, float fb,
                             float u0, float u1, float u2, int tx, float *pdf)
{
    float ps = rough > 0.0f ? SPEC_PROB : 0.0f;
    spo_v3 wi;
    if (u0 < fb)
        wi = u0 < fb * ps ? ggx_direction(N, V, rough, u1, u2, tx)
                          : cosine_direction(N, u1, u2, tx);
    else
        wi = mixture_sample_f(L, (u0 - fb) / (1.0f - fb), u1, u2, tx);
    float cosv = v3dot(N, wi);
    float pb = 0.0f;
    if (cosv > 0.0f)
    {
        pb = (1.0f - ps) * cosv * (1.0f / SPO_PI);
        if (ps > 0.0f) pb = pb + ps * ggx_pdf(N, V, wi, rough);
    }
    *pdf = cosv > 0.0f ? fb * pb + (1.0f - fb) * mixture_pdf_f(L, wi, tx) : 0.0f;
    return wi;
}

void spo_guided_scatter(const float *lobes40, const float *n, const float *v,
                        float roughness, float fb, float u0, float u1, float u2,
                        int tx, float *wi3, float *pdf)
{
    v3put(wi3, guided_scatter(lobes40, v3p(n), v3p(v), roughness, fb, u0, u1, u2, tx,
                              pdf));
}

static spo_v3 unit_rcp(spo_v3 a)
{
    float inv = 1.0f / sqrtf(a.x * a.x + a.y * a.y + a.z * a.z);
    return v3(a.x * inv, a.y * inv, a.z * inv);
}

void spo_guide_analytic_lobes(const float *Pf, const float *nf, const float *vf,
                              float *L)
{
    spo_v3 P = v3p(Pf), N = v3p(nf), V = v3p(vf);
    float nv = N.x * V.x + N.y * V.y + N.z * V.z;
    spo_v3 mus[3] = {unit_rcp(v3(0.3f - P.x, 2.5f - P.y, -0.4f - P.z)), N,
                     unit_rcp(v3(2.0f * nv * N.x - V.x, 2.0f * nv * N.y - V.y,
                                 2.0f * nv * N.z - V.z))};
    const float pis[3] = {0.5f, 0.3f, 0.2f}, kaps[3] = {40.0f, 3.0f, 120.0f};
    for (int k = 0; k < 8; ++k)
    {
        L[k] = k < 3 ? pis[k] : 0.0f;
        L[8 + k] = k < 3 ? kaps[k] : 1.0f;
        L[16 + 3 * k] = k < 3 ? mus[k].x : 0.0f;
        L[17 + 3 * k] = k < 3 ? mus[k].y : 0.0f;
        L[18 + 3 * k] = k < 3 ? mus[k].z : 1.0f;
    }
}

static float *g_sink;
static uint32_t g_sinkCap, g_sinkCount;

void spo_guided_sink(float *rows16, uint32_t cap)
{
    g_sink = rows16;
    g_sinkCap = rows16 ? cap : 0;
    g_sinkCount = 0;
}

uint32_t spo_guided_sink_count(void) { return g_sinkCount; }

static void rng_for_pixel(spo_Rng *rng, const spo_Config *cfg, uint32_t width,
                          uint32_t x, uint32_t y, uint32_t sample,
                          uint32_t *tileState)
{
    uint32_t pixelIndex = y * width + x;
    rng->mode = cfg->rng;
    if (cfg->rng == SPO_RNG_PCG)
    {
        spo_pcg32_seed(&rng->state, &rng->inc, cfg->seed, pixelIndex);
        spo_pcg32_advance(&rng->state, rng->inc, (uint64_t)sample << 20);
    }
    else if (cfg->rng == SPO_RNG_XORSHIFT_PIXEL)
    {
        if (sample == 0) rng->xs = spo_pixel_seed(cfg->seed, pixelIndex);
    }
    else
        rng->xs = *tileState;
}

/* world-space vertices of triangle k of object i */
/* Sign of the determinant of a column-major 4x4's upper 3x3 in double,
 * col0 . (col1 x col2) (the product's model_orient, spg_api.cpp): +1 keeps
 * a mesh's winding in world space, -1 mirrors it, 0 degenerate. */
static int model_orient(const float *m)
{
    const double x = (double)m[5] * m[10] - (double)m[6] * m[9];
    const double y = (double)m[6] * m[8] - (double)m[4] * m[10];
    const double z = (double)m[4] * m[9] - (double)m[5] * m[8];
    const double det = (double)m[0] * x + (double)m[1] * y + (double)m[2] * z;
    return det > 0.0 ? 1 : det < 0.0 ? -1 : 0;
}

static void world_triangle(const spo_Scene *s, uint32_t i, uint32_t k, spo_v3 *A,
                           spo_v3 *B, spo_v3 *C)
{
    const spo_Object *o = &s->objects[i];
    const spo_Mesh *m = &s->meshes[o->mesh];
    *A = m4point(o->model, m->vertices[m->indices[3 * k]].p);
    *B = m4point(o->model, m->vertices[m->indices[3 * k + 1]].p);
    *C = m4point(o->model, m->vertices[m->indices[3 * k + 2]].p);
}

/* The spatial emitter choice (emitInstCount > 0): emitter instance k seen
 * from x with shading normal n weighs power x (1 / max(d^2, r^2)), 0 when
 * its bounding sphere lies wholly behind the tangent plane; sums in index
 * order.  (The product's emit_weight, spg_trace.hip.) */
static float emit_weight(const spo_Scene *s, uint32_t k, spo_v3 x, spo_v3 n)
{
    const float *a = s->emitInst + 8 * (size_t)k;
    spo_v3 v = {a[0] - x.x, a[1] - x.y, a[2] - x.z};
    float d2 = v3dot(v, v), r2 = a[3] * a[3];
    return v3dot(n, v) > -a[3] ? a[4] * (1.0f / fmaxf(d2, r2)) : 0.0f;
}

static float emit_weight_sum(const spo_Scene *s, spo_v3 x, spo_v3 n)
{
    float W = 0.0f;
    for (uint32_t k = 0; k < s->emitInstCount; ++k) W += emit_weight(s, k, x, n);
    return W;
}

/* density per unit area of the emitter sampler at x (normal n) on triangle
 * k of object i: P(object | x, n) / area, or the global table's entry */
static float emit_pdf_area(const spo_Scene *s, uint32_t i, uint32_t k, spo_v3 x, spo_v3 n)
{
    int32_t b = s->emitBase[i];
    if (b < 0) return 0.0f;
    if (s->emitInstCount == 0) return s->emitPdfA[b + k];
    int32_t e = s->emitSlot[i];
    if (e < 0) return 0.0f;
    float W = emit_weight_sum(s, x, n);
    if (!(W > 0.0f)) return 0.0f;
    return emit_weight(s, (uint32_t)e, x, n) / W * s->emitInst[8 * (size_t)e + 5];
}

/* density per steradian of the emitter sampler at x (normal n) reaching
 * triangle k of object i at distance t along unit direction d:
 * pdfA t^2 / |cos| */
static float emit_pdf(const spo_Scene *s, uint32_t i, uint32_t k, float t, spo_v3 d,
                      spo_v3 x, spo_v3 nx)
{
    float pa = emit_pdf_area(s, i, k, x, nx);
    if (!(pa > 0.0f)) return 0.0f;
    spo_v3 A, B, C;
    world_triangle(s, i, k, &A, &B, &C);
    spo_v3 n = v3cross(v3sub(B, A), v3sub(C, A));
    float c = fabsf(v3dot(n, d)) / sqrtf(v3dot(n, n));
    return c > 0.0f ? pa * (t * t) / c : 0.0f;
}

/* a point on an emissive triangle, uniform on the triangle (e2, e3); the
 * unit direction from x and its density per steradian (0: nothing to
 * sample).  With the spatial choice an emitter instance by its weight from
 * (x, nx) -- the first whose running sum exceeds e4 x the total, else the
 * last with a positive weight -- then a triangle of it by area (its alias
 * table, r and e1); else an entry of the global table (r and e1). */
static spo_v3 emit_direction(const spo_Scene *s, uint32_t r, float e1, float e2,
                             float e3, float e4, spo_v3 x, spo_v3 nx, float *pdf,
                             float *dist, uint32_t *inst, uint32_t *prim)
{
    uint32_t g;
    float pa;
    if (s->emitInstCount)
    {
        float W = emit_weight_sum(s, x, nx);
        if (!(W > 0.0f))
        {
            *pdf = 0.0f;
            *dist = 0.0f;
            *inst = *prim = 0xFFFFFFFFu;
            return v3(0.0f, 0.0f, 1.0f);
        }
        float target = e4 * W, acc = 0.0f, wp = 0.0f;
        uint32_t pick = 0;
        int found = 0;
        for (uint32_t k = 0; k < s->emitInstCount; ++k)
        {
            float w = emit_weight(s, k, x, nx);
            acc += w;
            if (!found && w > 0.0f) { pick = k; wp = w; }
            found = found || acc > target;
        }
        const float *b = s->emitInst + 8 * (size_t)pick;
        uint32_t first, count;
        memcpy(&first, &b[6], 4);
        memcpy(&count, &b[7], 4);
        uint32_t j = (uint32_t)(((uint64_t)r * count) >> 32);
        if (!(e1 < s->emitProbL[first + j])) j = s->emitAliasL[first + j];
        g = first + j;
        pa = wp / W * b[5];
    }
    else
    {
        g = (uint32_t)(((uint64_t)r * s->emitCount) >> 32);
        if (!(e1 < s->emitProb[g])) g = s->emitAlias[g];
        pa = s->emitPdfA[g];
    }
    *inst = s->emitPair[2 * g];
    *prim = s->emitPair[2 * g + 1];
    spo_v3 A, B, C;
    world_triangle(s, *inst, *prim, &A, &B, &C);
    float su = sqrtf(e2);
    spo_v3 P = v3add(v3add(v3mul(A, 1.0f - su), v3mul(B, su * (1.0f - e3))),
                     v3mul(C, su * e3));
    spo_v3 v = v3sub(P, x);
    float dist2 = v3dot(v, v);
    spo_v3 w = v3norm(v);
    spo_v3 n = v3cross(v3sub(B, A), v3sub(C, A));
    float fw = v3dot(n, w);
    float c = fabsf(fw) / sqrtf(v3dot(n, n));
    /* a triangle clearly facing away from x (|cos| > 1e-3 at the light) can
     * never be the closest hit -- the intersection test accepts front faces
     * only (dot(d, n) < 0 in the mesh's frame; a mirroring instance
     * transform flips the world winding) -- so its sample casts no shadow
     * ray: pdf 0 */
    int orient = model_orient(s->objects[*inst].model.m);
    int away = orient != 0 && (orient > 0 ? fw > 0.0f : fw < 0.0f) && c > 1e-3f;
    *pdf = c > 0.0f && dist2 > 0.0f && !away ? pa * dist2 / c : 0.0f;
    *dist = sqrtf(dist2);
    return w;
}

/* One light sample of next-event estimation at x (shading normal N, view V,
 * material mat; spg_trace.hip nee_light_sample): the sky with probability
 * envSelect, else an emitter (emit_direction); draws a 24-bit uniform, one
 * 32-bit value, four 24-bit uniforms.  The power heuristic weighs the light
 * density against the scattering density: the BSDF sampler's, or with
 * `lobes` the one-sample MIS fb p_bsdf + (1 - fb) q of the guided
 * estimator.  The sky is seen unless some object lies at a positive
 * distance along the shadow ray; an emitter triangle must be the closest
 * hit within 1.001 x the light's distance (the product's shadow walk starts
 * with that bound as its best t).  Returns 1 with *C = beta o f cos Le w /
 * p_light when the light is reached, 0 when not (or nothing was sampled),
 * -1 if the scene query failed. */
static int nee_sample(const spo_Scene *s, const spo_Config *cfg, spo_Rng *rng, spo_v3 x,
                      spo_v3 N, spo_v3 V, const float *mat, spo_v3 beta,
                      const float *lobes, float fb, uint64_t *metrics, spo_v3 *C)
{
    const int tx = cfg->tx;
    float e0 = rng_u24(rng);
    uint32_t r = rng_next(rng);
    float e1 = rng_u24(rng), e2 = rng_u24(rng), e3 = rng_u24(rng);
    float e4 = rng_u24(rng);
    int useEnv = e0 < s->envSelect;
    if (!(useEnv ? s->envPdf != NULL : s->emitCount > 0)) return 0;
    spo_v3 wl, Le;
    float pl, dl = 0.0f;
    uint32_t ti = 0xFFFFFFFFu, tp = 0;
    if (useEnv)
    {
        float wl3[3];
        spo_env_direction(s->envProb, s->envAlias, s->envW, s->envH, r, e1, e2, e3, wl3);
        wl = v3p(wl3);
        pl = s->envSelect * spo_env_pdf(s->envPdf, s->envW, s->envH, wl3);
        float bg[7];
        spo_v2 uv0 = {0, 0};
        spo_material_eval_internal(s, s->backgroundMaterialId, uv0, v3neg(wl), tx, bg);
        Le = v3p(bg + 3);
    }
    else
    {
        float pg;
        wl = emit_direction(s, r, e1, e2, e3, e4, x, N, &pg, &dl, &ti, &tp);
        pl = (1.0f - s->envSelect) * pg;
        Le = v3(0.0f, 0.0f, 0.0f);
        if (pg > 0.0f)  /* else nothing was sampled */
        {
            float em[7];
            spo_v2 uv0 = {0, 0};
            spo_material_eval_internal(s, s->objects[ti].material, uv0, v3neg(wl), tx, em);
            Le = v3p(em + 3);
        }
    }
    float cl = v3dot(N, wl);
    if (!(cl > 0.0f && pl > 0.0f)) return 0;
    float rough = mat[6];
    float ps = rough > 0.0f ? SPEC_PROB : 0.0f;
    spo_v3 Xl = brdf_bracket(N, V, wl, mat, tx, 1);
    float pcl = (1.0f - ps) * cl * (1.0f / SPO_PI);
    float pb = ps > 0.0f ? pcl + ps * ggx_pdf(N, V, wl, rough) : pcl;
    if (lobes) pb = fb * pb + (1.0f - fb) * mixture_pdf_f(lobes, wl, tx);
    float pl2 = pl * pl;
    float wmis = pl2 / (pl2 + pb * pb);
    *C = v3had(v3had(beta, v3mul(Xl, (cl * wmis) / pl)), Le);
    float so[3], sd[3];
    v3put(so, x); v3put(sd, wl);
    int visible;
    if (ti == 0xFFFFFFFFu)
    {
        int occ = spo_scene_occluded(s, so, sd, cfg->hit, metrics);
        if (occ < 0) return -1;
        visible = !occ;
    }
    else
    {
        spo_SceneHit sh;
        if (spo_scene_intersect(s, so, sd, cfg->hit, tx, &sh, metrics) < 0) return -1;
        visible = sh.t > 0.0f && !(sh.t > dl * 1.001f) && sh.object == ti && sh.prim == tp;
    }
    metrics[2]++;
    return visible;
}

/* Power-heuristic share of emission reached by a scattered ray (density pb
 * at the previous vertex, there shading normal lastN) against the light
 * sampler: toward object i's triangle k at distance t along d from o, or
 * (i = ~0) the sky along d (spg_trace.hip nee_hit_weight). */
static float nee_hit_weight(const spo_Scene *s, uint32_t i, uint32_t k, float t, spo_v3 d,
                            spo_v3 o, spo_v3 lastN, float pb)
{
    float pl;
    if (i == 0xFFFFFFFFu)
    {
        float dd[3];
        v3put(dd, d);
        pl = s->envPdf ? s->envSelect * spo_env_pdf(s->envPdf, s->envW, s->envH, dd) : 0.0f;
    }
    else
        pl = (1.0f - s->envSelect) * emit_pdf(s, i, k, t, d, o, lastN);
    float pb2 = pb * pb;
    return pb2 / (pb2 + pl * pl);
}

/* The guided estimator with the verification mixture (the product's
 * SPG_EST_GUIDED under SPG_GUIDE_ANALYTIC_MIXTURE): at each hit that may
 * scatter, three 24-bit uniforms and one-sample MIS of the BSDF lobe and
 * the mixture (guided_scatter); vertex records (E, X, cos/pdf); the
 * radiance by the reverse recurrence L_j = E_j + (X_j o L_j+1) cos_j/pdf_j,
 * and, for the training sink, each scattering vertex's weight
 * Y(X_j o L_j+1 cos_j/pdf_j) (luminance 0.2126, 0.7152, 0.0722).  With
 * SPO_EST_GUIDED_NEE (the product's SPG_EST_GUIDED_NEE) emission reached by
 * a scattered ray is weighted against the light sampler (nee_hit_weight at
 * the previous vertex with its MIS density), and each scattering vertex
 * draws one light sample after its scatter uniforms (nee_sample against the
 * one-sample MIS density), whose reached contribution joins E_j. */
static spo_v3 trace_guided(const spo_Scene *s, const spo_Config *cfg, spo_v3 o,
                           spo_v3 d, spo_Rng *rng, uint64_t *metrics, int *fail,
                           uint32_t pixel, uint32_t sample)
{
    const int tx = cfg->tx;
    const float fb = cfg->bsdfFraction;
    const int nee = cfg->estimator == SPO_EST_GUIDED_NEE;
    spo_v3 E[16], X[16], P[16], Nv[16], Vv[16], Wv[16];
    float cv[16];
    float lastPdf = 0.0f;
    spo_v3 lastN = v3(0.0f, 0.0f, 0.0f);
    uint32_t vc = 0;
    uint32_t bounces = cfg->bounces < 16 ? cfg->bounces : 16;
    for (uint32_t bounce = 0; bounce < bounces; ++bounce)
    {
        float of[3], df[3];
        v3put(of, o); v3put(df, d);
        spo_SceneHit h;
        if (spo_scene_intersect(s, of, df, cfg->hit, tx, &h, metrics) < 0)
        { *fail = 1; return v3(0, 0, 0); }
        metrics[2]++;
        spo_v3 V = v3neg(d);
        E[vc] = X[vc] = v3(0, 0, 0);
        cv[vc] = 0.0f;
        int push = 0;
        if (h.t > 0.0f)
        {
            metrics[3]++;
            float mat[7];
            spo_material_eval_internal(s, h.materialId, h.uv, V, tx, mat);
            E[vc] = v3p(mat + 3);
            if (nee && vc > 0 && s->emitCount && s->emitBase[h.object] >= 0)
                E[vc] = v3mul(E[vc],
                              nee_hit_weight(s, h.object, h.prim, h.t, d, o, lastN, lastPdf));
            if (vc + 1 < cfg->bounces)
            {
                float u0 = rng_u24(rng), u1 = rng_u24(rng), u2 = rng_u24(rng);
                spo_v3 Ph = v3add(o, v3mul(d, h.t));
                float L[40], Pf[3], nf[3], vf[3];
                v3put(Pf, Ph); v3put(nf, h.normal); v3put(vf, V);
                spo_guide_analytic_lobes(Pf, nf, vf, L);
                float pdf;
                spo_v3 wi = guided_scatter(L, h.normal, V, mat[6], fb, u0, u1, u2, tx, &pdf);
                float cosv = v3dot(h.normal, wi);
                if (nee)
                {
                    /* one light sample against the one-sample MIS density
                     * of BSDF and mixture; reached, it adds to this
                     * vertex's emission term */
                    spo_v3 x = v3add(Ph, v3mul(h.normal, 0.0001f));
                    spo_v3 C;
                    int vis = nee_sample(s, cfg, rng, x, h.normal, V, mat, v3s(1.0f), L, fb,
                                         metrics, &C);
                    if (vis < 0) { *fail = 1; return v3(0, 0, 0); }
                    if (vis) E[vc] = v3add(E[vc], C);
                }
                if (cosv > 0.0f && pdf > 0.0f)
                {
                    X[vc] = brdf_bracket(h.normal, V, wi, mat, tx, 1);
                    cv[vc] = cosv / pdf;
                    P[vc] = Ph; Nv[vc] = h.normal; Vv[vc] = V; Wv[vc] = wi;
                    o = v3add(Ph, v3mul(h.normal, 0.0001f));
                    d = wi;
                    lastPdf = pdf;
                    lastN = h.normal;
                    push = 1;
                }
            }
        }
        else
        {
            metrics[4]++;
            float mat[7];
            spo_v2 uv0 = {0, 0};
            spo_material_eval_internal(s, s->backgroundMaterialId, uv0, V, tx, mat);
            E[vc] = v3p(mat + 3);
            if (nee && vc > 0)
                E[vc] = v3mul(E[vc], nee_hit_weight(s, 0xFFFFFFFFu, 0, h.t, d, o, lastN,
                                                    lastPdf));
        }
        vc++;
        if (!push) break;
    }
    spo_v3 radiance = v3(0, 0, 0);
    for (int32_t j = (int32_t)vc - 1; j >= 0; --j)
    {
        spo_v3 T = v3mul(v3had(X[j], radiance), cv[j]);
        if (cv[j] > 0.0f)
        {
            float w = 0.2126f * T.x + 0.7152f * T.y + 0.0722f * T.z;
            if (w > 0.0f && g_sinkCount < g_sinkCap)
            {
                float *r = g_sink + 16 * (size_t)g_sinkCount++;
                v3put(r, P[j]); v3put(r + 3, Nv[j]); v3put(r + 6, Vv[j]);
                v3put(r + 9, Wv[j]);
                r[12] = w;
                r[13] = (float)pixel;
                r[14] = (float)sample;
                r[15] = (float)j;
            }
        }
        radiance = v3add(E[j], T);
    }
    return radiance;
}

/* One sample of one pixel.  Parity estimator: src/simd_path_tracer.cpp:
 * 216-322.  Unbiased estimator: the same material model and camera, cosine
 * sampling and the 1/pdf weight. */
static spo_v3 trace_sample(const spo_Scene *s, const spo_Camera *c,
                           const spo_Config *cfg, uint32_t x, uint32_t y,
                           spo_Rng *rng, uint64_t *metrics, int *fail,
                           uint32_t sample)
{
    int tx = cfg->tx;
    float b1 = rng_bi(rng), b2 = rng_bi(rng);
    float jx, jy;
    if (cfg->jitterYFirst) { jy = c->halfPixelHeight * b1; jx = c->halfPixelWidth * b2; }
    else { jx = c->halfPixelWidth * b1; jy = c->halfPixelHeight * b2; }
    float px = (float)x + 0.5f, py = (float)y + 0.5f;
    px = px + jx;
    py = py + jy;
    float filmP[3];
    spo_film_position(c, px, py, filmP);
    spo_v3 o = c->position;
    spo_v3 d = v3norm(v3sub(v3p(filmP), c->position));
    if (cfg->estimator == SPO_EST_GUIDED || cfg->estimator == SPO_EST_GUIDED_NEE)
        return trace_guided(s, cfg, o, d, rng, metrics, fail, y * c->width + x, sample);
    if (cfg->estimator == SPO_EST_UNBIASED || cfg->estimator == SPO_EST_NEE)
    {
        const int nee = cfg->estimator == SPO_EST_NEE;
        float lastPdf = 0.0f;
        spo_v3 lastN = v3(0.0f, 0.0f, 0.0f);  /* shading normal of the last scatter */
        spo_v3 beta = v3s(1.0f), L = v3(0, 0, 0);
        for (uint32_t bounce = 0; bounce < cfg->bounces; ++bounce)
        {
            float of[3], df[3];
            v3put(of, o); v3put(df, d);
            spo_SceneHit h;
            if (spo_scene_intersect(s, of, df, cfg->hit, tx, &h, metrics) < 0)
            { *fail = 1; return L; }
            metrics[2]++;
            spo_v3 V = v3neg(d);
            if (h.t > 0.0f)
            {
                metrics[3]++;
                float mat[7];
                spo_material_eval_internal(s, h.materialId, h.uv, V, tx, mat);
                if (nee && bounce > 0 && s->emitCount && s->emitBase[h.object] >= 0)
                {
                    /* a BSDF-sampled ray that reached an emitter: its share
                     * against the emitter sampler (power heuristic) */
                    float wm = nee_hit_weight(s, h.object, h.prim, h.t, d, o, lastN, lastPdf);
                    L = v3add(L, v3had(beta, v3mul(v3p(mat + 3), wm)));
                }
                else
                    L = v3add(L, v3had(beta, v3p(mat + 3)));
                /* cosine lobe, or GGX with probability ps on rough
                 * materials; weight X cos / pdf (X pi for ps = 0) */
                float rough = mat[6];
                float ps = rough > 0.0f ? SPEC_PROB : 0.0f;
                float u0 = ps > 0.0f ? rng_u24(rng) : 1.0f;
                float u1 = rng_u24(rng), u2 = rng_u24(rng);
                spo_v3 wi = u0 < ps ? ggx_direction(h.normal, V, rough, u1, u2, tx)
                                    : cosine_direction(h.normal, u1, u2, tx);
                float cosine = v3dot(h.normal, wi);
                if (nee && bounce + 1 < cfg->bounces)
                {
                    /* next-event estimation against the BSDF sampler's
                     * density */
                    spo_v3 x = v3add(v3add(o, v3mul(d, h.t)), v3mul(h.normal, 0.0001f));
                    spo_v3 C;
                    int vis = nee_sample(s, cfg, rng, x, h.normal, V, mat, beta, NULL, 0.0f,
                                         metrics, &C);
                    if (vis < 0) { *fail = 1; return L; }
                    if (vis) L = v3add(L, C);
                }
                if (!(cosine > 0.0f)) break;
                spo_v3 X = brdf_bracket(h.normal, V, wi, mat, tx, 1);
                float w = SPO_PI;
                float pdfc = cosine * (1.0f / SPO_PI);
                if (ps > 0.0f)
                {
                    float pc = (1.0f - ps) * cosine * (1.0f / SPO_PI);
                    pdfc = pc + ps * ggx_pdf(h.normal, V, wi, rough);
                    w = cosine / pdfc;
                }
                lastPdf = pdfc;
                lastN = h.normal;
                beta = v3had(beta, v3mul(X, w));
                spo_v3 p = v3add(o, v3mul(d, h.t));
                o = v3add(p, v3mul(h.normal, 0.0001f));
                d = wi;
            }
            else
            {
                metrics[4]++;
                float mat[7];
                spo_v2 uv0 = {0, 0};
                spo_material_eval_internal(s, s->backgroundMaterialId, uv0, V,
                                           tx, mat);
                if (nee && bounce > 0)
                {
                    /* the BSDF-sampled direction's share (power heuristic) */
                    float wm = nee_hit_weight(s, 0xFFFFFFFFu, 0, h.t, d, o, lastN, lastPdf);
                    L = v3add(L, v3had(beta, v3mul(v3p(mat + 3), wm)));
                }
                else
                    L = v3add(L, v3had(beta, v3p(mat + 3)));
                break;
            }
        }
        return L;
    }
    spo_PathVertex path[16];
    memset(path, 0, sizeof path);
    uint32_t len = 0;
    uint32_t bounces = cfg->bounces < 16 ? cfg->bounces : 16;
    for (uint32_t bounce = 0; bounce < bounces; ++bounce)
    {
        float of[3], df[3];
        v3put(of, o); v3put(df, d);
        spo_SceneHit h;
        if (spo_scene_intersect(s, of, df, cfg->hit, tx, &h, metrics) < 0)
        { *fail = 1; return v3(0, 0, 0); }
        metrics[2]++;
        spo_PathVertex *v = &path[len++];
        if (h.t > 0.0f)
        {
            v->materialId = h.materialId;
            v->worldPosition = v3add(o, v3mul(d, h.t));
            v->outgoingDir = v3neg(d);
            v->normal = h.normal;
            v->uv = h.uv;
            spo_v3 dir = hemisphere(h.normal, rng, tx);
            v->incomingDir = dir;
            o = v3add(v->worldPosition, v3mul(h.normal, 0.0001f));
            d = dir;
            metrics[3]++;
        }
        else
        {
            v->materialId = s->backgroundMaterialId;
            v->outgoingDir = v3neg(d);
            metrics[4]++;
            break;
        }
    }
    return radiance_for_path(s, path, len, tx, cfg->radianceClamp);
}

/* sp_PathTraceTile, src/simd_path_tracer.cpp:178-345.  pixels is the whole
 * RGBA32F image (width*height*4); only the tile's pixels are written. */
int spo_path_trace_tile(const spo_Scene *s, const spo_Camera *c,
                        const spo_Config *cfg, uint32_t minX, uint32_t minY,
                        uint32_t maxX, uint32_t maxY, uint32_t *rngState,
                        float *pixels, uint64_t *metrics12)
{
    uint64_t local[SPO_MAX_METRICS] = {0};
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    maxX = spo_minu(maxX, c->width);
    maxY = spo_minu(maxY, c->height);
    uint32_t tileState = rngState ? *rngState : cfg->seed;
    int fail = 0;
    for (uint32_t y = minY; y < maxY; y++)
        for (uint32_t x = minX; x < maxX; x++)
        {
            spo_v3 total = v3(0, 0, 0);
            spo_Rng rng;
            memset(&rng, 0, sizeof rng);
            for (uint32_t sample = 0; sample < cfg->samplesPerPixel; sample++)
            {
                rng_for_pixel(&rng, cfg, c->width, x, y, sample, &tileState);
                spo_v3 radiance = trace_sample(s, c, cfg, x, y, &rng, local,
                                               &fail, sample);
                if (cfg->rng == SPO_RNG_XORSHIFT_TILE) tileState = rng.xs;
                if (fail) return -1;
                total = v3add(total, v3mul(radiance,
                                           1.0f / (float)cfg->samplesPerPixel));
                local[1]++;
            }
            float *px = pixels + 4 * ((uint64_t)y * c->width + x);
            px[0] = total.x; px[1] = total.y; px[2] = total.z; px[3] = 1.0f;
        }
    if (rngState) *rngState = tileState;
    clock_gettime(CLOCK_MONOTONIC, &t1);
    local[0] = (uint64_t)((t1.tv_sec - t0.tv_sec) * 1000000000LL +
                          (t1.tv_nsec - t0.tv_nsec));
    if (metrics12)
        for (int i = 0; i < SPO_MAX_METRICS; ++i) metrics12[i] += local[i];
    return 0;
}

/* ------------------------------------------------- threaded CPU baseline */
typedef struct {
    const spo_Scene *s;
    const spo_Camera *c;
    const spo_Config *cfg;
    uint32_t *tiles;
    uint32_t tileCount;
    atomic_uint next;
    float *pixels;
    uint64_t metrics[64][SPO_MAX_METRICS];
    atomic_int fail;
} render_job;

typedef struct { render_job *job; uint32_t id; } render_arg;

/* WorkerThread, src/main.cpp:731-759: pop a tile, reset the RNG to the
 * per-tile constant, trace it. */
static void *render_worker(void *p)
{
    render_arg *a = (render_arg *)p;
    render_job *j = a->job;
    for (;;)
    {
        uint32_t i = atomic_fetch_add(&j->next, 1u);
        if (i >= j->tileCount) break;
        uint32_t *t = j->tiles + 4 * i;
        uint32_t state = j->cfg->seed;
        if (spo_path_trace_tile(j->s, j->c, j->cfg, t[0], t[1], t[2], t[3],
                                &state, j->pixels, j->metrics[a->id]) < 0)
            atomic_store(&j->fail, 1);
    }
    return NULL;
}

double spo_render_threaded(const spo_Scene *s, const spo_Camera *c,
                           const spo_Config *cfg, uint32_t tileW,
                           uint32_t tileH, uint32_t threads, float *pixels,
                           uint64_t *metrics12)
{
    if (threads < 1) threads = 1;
    if (threads > 64) threads = 64;
    uint32_t maxTiles = ((c->width + tileW - 1) / tileW) *
                        ((c->height + tileH - 1) / tileH);
    render_job *j = (render_job *)calloc(1, sizeof *j);
    j->s = s; j->c = c; j->cfg = cfg; j->pixels = pixels;
    j->tiles = (uint32_t *)malloc(16 * (maxTiles ? maxTiles : 1));
    j->tileCount = spo_compute_tiles(c->width, c->height, tileW, tileH,
                                     maxTiles, j->tiles);
    atomic_init(&j->next, 0u);
    atomic_init(&j->fail, 0);
    pthread_t th[64];
    render_arg args[64];
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (uint32_t i = 0; i < threads; ++i)
    {
        args[i].job = j;
        args[i].id = i;
        pthread_create(&th[i], NULL, render_worker, &args[i]);
    }
    for (uint32_t i = 0; i < threads; ++i) pthread_join(th[i], NULL);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    if (metrics12)
        for (uint32_t i = 0; i < threads; ++i)
            for (int k = 0; k < SPO_MAX_METRICS; ++k)
                metrics12[k] += j->metrics[i][k];
    int fail = atomic_load(&j->fail);
    free(j->tiles);
    free(j);
    double secs = (double)(t1.tv_sec - t0.tv_sec) +
                  1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    return fail ? -secs : secs;
}

/* -------------------------------------------- brute-force closest hit */
typedef struct {
    const float *tris; uint32_t ntris;
    const float *rays; uint32_t nrays;
    uint32_t *prim; float *t; double *t64;
    atomic_uint next;
} brute_job;

static void *brute_worker(void *p)
{
    brute_job *j = (brute_job *)p;
    for (;;)
    {
        uint32_t r0 = atomic_fetch_add(&j->next, 64u);
        if (r0 >= j->nrays) break;
        uint32_t r1 = r0 + 64 < j->nrays ? r0 + 64 : j->nrays;
        for (uint32_t r = r0; r < r1; ++r)
        {
            const float *o = j->rays + 6 * r, *d = o + 3;
            uint32_t best = 0xFFFFFFFFu;
            if (j->t64)
            {
                double od[3] = {o[0], o[1], o[2]}, dd[3] = {d[0], d[1], d[2]};
                double bt = -1.0;
                for (uint32_t k = 0; k < j->ntris; ++k)
                {
                    const float *tr = j->tris + 9 * k;
                    double a[3] = {tr[0], tr[1], tr[2]}, b[3] = {tr[3], tr[4], tr[5]},
                           c[3] = {tr[6], tr[7], tr[8]};
                    double t = spo_rt_mt_f64(od, dd, a, b, c);
                    if (t > 0.0 && (t < bt || bt < 0.0)) { bt = t; best = k; }
                }
                j->t64[r] = bt;
            }
            else
            {
                float bt = -1.0f;
                for (uint32_t k = 0; k < j->ntris; ++k)
                {
                    const float *tr = j->tris + 9 * k;
                    float t = spo_rt_mt(o, d, tr, tr + 3, tr + 6, NULL);
                    if (t > 0.0f && (t < bt || bt < 0.0f)) { bt = t; best = k; }
                }
                j->t[r] = bt;
            }
            j->prim[r] = best;
        }
    }
    return NULL;
}

static void brute_run(brute_job *j, uint32_t threads)
{
    if (threads < 1) threads = 1;
    if (threads > 64) threads = 64;
    pthread_t th[64];
    atomic_init(&j->next, 0u);
    for (uint32_t i = 0; i < threads; ++i)
        pthread_create(&th[i], NULL, brute_worker, j);
    for (uint32_t i = 0; i < threads; ++i) pthread_join(th[i], NULL);
}

/* The closest-hit definition the GPU kernel is held to: among all triangles
 * with a Moller-Trumbore hit at t > 0 (src/ray_intersection.cpp:156-190,
 * acceptance as src/sp_scene.cpp:189-193), the smallest t; equal t goes to
 * the lower triangle index. */
void spo_closest_hit_brute(const float *tris9, uint32_t ntris,
                           const float *rays6, uint32_t nrays,
                           uint32_t threads, uint32_t *primOut, float *tOut)
{
    brute_job j;
    memset(&j, 0, sizeof j);
    j.tris = tris9; j.ntris = ntris; j.rays = rays6; j.nrays = nrays;
    j.prim = primOut; j.t = tOut;
    brute_run(&j, threads);
}

void spo_closest_hit_brute_f64(const float *tris9, uint32_t ntris,
                               const float *rays6, uint32_t nrays,
                               uint32_t threads, uint32_t *primOut,
                               double *tOut)
{
    brute_job j;
    memset(&j, 0, sizeof j);
    j.tris = tris9; j.ntris = ntris; j.rays = rays6; j.nrays = nrays;
    j.prim = primOut; j.t64 = tOut;
    brute_run(&j, threads);
}

/* FP64 Moller-Trumbore t for given (ray, triangle) pairs. */
void spo_mt_f64_pairs(const float *tris9, const float *rays6,
                      const uint32_t *prim, uint32_t n, double *tOut)
{
    for (uint32_t r = 0; r < n; ++r)
    {
        if (prim[r] == 0xFFFFFFFFu) { tOut[r] = -1.0; continue; }
        const float *o = rays6 + 6 * r, *d = o + 3, *tr = tris9 + 9 * (uint64_t)prim[r];
        double od[3] = {o[0], o[1], o[2]}, dd[3] = {d[0], d[1], d[2]};
        double a[3] = {tr[0], tr[1], tr[2]}, b[3] = {tr[3], tr[4], tr[5]},
               c[3] = {tr[6], tr[7], tr[8]};
        tOut[r] = spo_rt_mt_f64(od, dd, a, b, c);
    }
}
