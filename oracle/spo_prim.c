/* ORACLE -- test infrastructure only (see spo.h).
 * Ray/primitive tests, RNG, spherical maps, samplers, cube maps, tiles,
 * transforms and mesh generators, each restating the reference file:line
 * named above it. */
#include <stdlib.h>
#include <string.h>
#include "spo_math.h"

/* RayIntersectTriangleMT, src/ray_intersection.cpp:156-190.
 * out6 = t, u, v, n.xyz (n unnormalized); t = -1 on a miss. */
float spo_rt_mt(const float *o, const float *d, const float *a, const float *b,
                const float *c, float *out6)
{
    spo_v3 ro = v3p(o), rd = v3p(d), A = v3p(a), B = v3p(b), C = v3p(c);
    spo_v3 T = v3sub(ro, A);
    spo_v3 e1 = v3sub(B, A);
    spo_v3 e2 = v3sub(C, A);
    spo_v3 p = v3cross(rd, e2);
    spo_v3 q = v3cross(T, e1);
    spo_v3 n = v3cross(e1, e2);
    spo_v3 m = v3(v3dot(q, e2), v3dot(p, T), v3dot(q, rd));
    float det = 1.0f / v3dot(p, e1);
    float t = det * m.x;
    float u = det * m.y;
    float v = det * m.z;
    float w = 1.0f - u - v;
    float f = v3dot(rd, n);
    float r[6] = {-1.0f, 0, 0, 0, 0, 0};
    if (u >= 0.0f && u <= 1.0f && v >= 0.0f && v <= 1.0f && w >= 0.0f &&
        w <= 1.0f && f < 0.0f)
    {
        r[0] = t; r[1] = u; r[2] = v; r[3] = n.x; r[4] = n.y; r[5] = n.z;
    }
    if (out6) memcpy(out6, r, sizeof r);
    return r[0];
}

/* RayIntersectTriangleSlow, src/ray_intersection.cpp:81-154 (tmin = F32_MAX);
 * the triangle test the reference's unit tests compile. */
float spo_rt_triangle_slow(const float *o, const float *d, const float *a,
                           const float *b, const float *c, float *out6)
{
    spo_v3 ro = v3p(o), rd = v3p(d), A = v3p(a), B = v3p(b), C = v3p(c);
    float r[6] = {-1.0f, 0, 0, 0, 0, 0};
    spo_v3 edgeCA = v3sub(C, A);
    spo_v3 edgeBA = v3sub(B, A);
    spo_v3 normal = v3norm(v3cross(edgeBA, edgeCA));
    float denom = v3dot(normal, rd);
    float t = v3dot(v3sub(A, ro), normal) / denom;
    if (t > 0.0f && t < SPO_F32_MAX)
    {
        spo_v3 p = v3add(ro, v3mul(rd, t));
        spo_v3 normalCA = v3cross(normal, edgeCA);
        float distCA = v3dot(A, normalCA);
        spo_v3 normalBA = v3cross(edgeBA, normal);
        float distBA = v3dot(A, normalBA);
        spo_v3 edgeBC = v3sub(B, C);
        spo_v3 normalBC = v3cross(normal, edgeBC);
        float distBC = v3dot(C, normalBC);
        int inCA = (v3dot(normalCA, p) - distCA) < 0.0f;
        int inBA = (v3dot(normalBA, p) - distBA) < 0.0f;
        int inBC = (v3dot(normalBC, p) - distBC) < 0.0f;
        if (inCA && inBA && inBC)
        {
            r[0] = t; r[3] = normal.x; r[4] = normal.y; r[5] = normal.z;
        }
    }
    if (out6) memcpy(out6, r, sizeof r);
    return r[0];
}

/* RayIntersectAabb, src/ray_intersection.cpp:24-77 */
float spo_rt_aabb_scalar(const float *mn, const float *mx, const float *o,
                         const float *d)
{
    float tmin = 0.0f, tmax = SPO_F32_MAX;
    for (int axis = 0; axis < 3; ++axis)
    {
        float s = o[axis], dd = d[axis], lo = mn[axis], hi = mx[axis];
        if (fabsf(dd) < SPO_EPSILON)
        {
            if (s < lo || s > hi) { tmin = -1.0f; break; }
        }
        else
        {
            float a = (lo - s) / dd;
            float b = (hi - s) / dd;
            float t0 = spo_min(a, b);
            float t1 = spo_max(a, b);
            if (t0 > tmin) tmin = t0;
            tmax = spo_min(tmax, t1);
            if (tmin > tmax) { tmin = -1.0f; break; }
        }
    }
    return tmin;
}

/* simd_RayIntersectAabb with SIMD_LANE_WIDTH 1, src/simd.h:156-195 */
float spo_rt_aabb_simd1(const float *mn, const float *mx, const float *o,
                        const float *d)
{
    float tmin = 0.0f, tmax = SPO_F32_MAX;
    uint32_t parallelMask = 0;
    for (int axis = 0; axis < 3; ++axis)
    {
        float s = o[axis], dd = d[axis], lo = mn[axis], hi = mx[axis];
        uint32_t mask = fabsf(dd) < SPO_EPSILON ? 0xFFFFFFFFu : 0u;
        uint32_t u = s < lo ? 0xFFFFFFFFu : 0u;
        uint32_t v = s > hi ? 0xFFFFFFFFu : 0u;
        parallelMask |= mask & (u | v);
        float a = (lo - s) / dd;
        float b = (hi - s) / dd;
        float t0 = (a < b) ? a : b;
        float t1 = (a > b) ? a : b;
        tmin = (tmin > t0) ? tmin : t0;
        tmax = (tmax < t1) ? tmax : t1;
    }
    uint32_t mask2 = tmin > tmax ? 0xFFFFFFFFu : 0u;
    return (parallelMask | mask2) ? -1.0f : tmin;
}

/* simd_RayIntersectAabb4, src/simd.h:198-271: per box and axis
 * t0=(min-o)*invD, t1=(max-o)*invD, tmin=MINPS(t0,t1), tmax=MAXPS(t0,t1)
 * (MINPS/MAXPS return the second operand unless the compare holds), then
 * hit <=> Max(0, MaxComponent(tmin)) <= MinComponent(tmax). */
uint32_t spo_rt_aabb4(const float *mins12, const float *maxs12, const float *o,
                      const float *invD)
{
    uint32_t mask = 0;
    for (int i = 0; i < 4; ++i)
    {
        float tmn[3], tmx[3];
        for (int axis = 0; axis < 3; ++axis)
        {
            float t0 = (mins12[3 * i + axis] - o[axis]) * invD[axis];
            float t1 = (maxs12[3 * i + axis] - o[axis]) * invD[axis];
            tmn[axis] = (t0 < t1) ? t0 : t1;
            tmx[axis] = (t0 > t1) ? t0 : t1;
        }
        float entry = spo_max(0.0f, spo_max(tmn[0], spo_max(tmn[1], tmn[2])));
        float exit_ = spo_min(tmx[0], spo_min(tmx[1], tmx[2]));
        if (entry <= exit_) mask |= 1u << i;
    }
    return mask;
}

/* Moller-Trumbore in FP64 with the reference's acceptance rule; the t-error
 * oracle for the FP32 kernels (no reference counterpart). */
double spo_rt_mt_f64(const double *o, const double *d, const double *a,
                     const double *b, const double *c)
{
    double T[3], e1[3], e2[3], p[3], q[3], n[3];
    for (int i = 0; i < 3; ++i)
    {
        T[i] = o[i] - a[i]; e1[i] = b[i] - a[i]; e2[i] = c[i] - a[i];
    }
#define CROSS(r, x, y) \
    r[0] = x[1] * y[2] - x[2] * y[1]; \
    r[1] = x[2] * y[0] - x[0] * y[2]; \
    r[2] = x[0] * y[1] - x[1] * y[0];
#define DOT(x, y) (x[0] * y[0] + x[1] * y[1] + x[2] * y[2])
    CROSS(p, d, e2);
    CROSS(q, T, e1);
    CROSS(n, e1, e2);
    double det = 1.0 / DOT(p, e1);
    double t = det * DOT(q, e2), u = det * DOT(p, T), v = det * DOT(q, d);
    double w = 1.0 - u - v;
    double f = DOT(d, n);
#undef CROSS
#undef DOT
    if (u >= 0 && u <= 1 && v >= 0 && v <= 1 && w >= 0 && w <= 1 && f < 0)
        return t;
    return -1.0;
}

/* ---------------------------------------------------------------- RNG */
/* XorShift32 (Marsaglia 13/17/5), src/math_utils.h:190-198 */
uint32_t spo_xorshift32(uint32_t *state)
{
    uint32_t x = *state;
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    *state = x;
    return x;
}

static inline float spo_unilateral_from(uint32_t x)
{
    /* src/math_utils.h:200-208 */
    float numerator = (float)(x >> 1);
    float denom = (float)(0xFFFFFFFFu >> 1);
    return numerator / denom;
}

float spo_random_unilateral(uint32_t *state)
{
    return spo_unilateral_from(spo_xorshift32(state));
}

/* src/math_utils.h:210-214 */
float spo_random_bilateral(uint32_t *state)
{
    return -1.0f + 2.0f * spo_random_unilateral(state);
}

/* Per-pixel xorshift32 seed (no reference counterpart: the reference seeds
 * every tile with one constant, src/main.cpp:738-739).  The published
 * 32-bit integer finalizer "lowbias32" (C. Wellons, hash-prospector) of
 * base + golden-ratio * index; zero is remapped since xorshift32 stays at
 * zero forever. */
uint32_t spo_pixel_seed(uint32_t base, uint32_t pixelIndex)
{
    uint32_t x = base + pixelIndex * 0x9E3779B9u;
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return x ? x : 0x6D2B79F5u;
}

/* PCG32 (XSH-RR 64/32), O'Neill 2014 "PCG: A Family of Simple Fast
 * Space-Efficient Statistically Good Algorithms for Random Number
 * Generation"; seeding as the published pcg32_srandom_r. */
#define SPO_PCG_MULT 6364136223846793005ULL
uint32_t spo_pcg32_next(uint64_t *state, uint64_t inc)
{
    uint64_t old = *state;
    *state = old * SPO_PCG_MULT + inc;
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((-rot) & 31u));
}

void spo_pcg32_seed(uint64_t *state, uint64_t *inc, uint64_t initstate,
                    uint64_t initseq)
{
    *state = 0u;
    *inc = (initseq << 1u) | 1u;
    spo_pcg32_next(state, *inc);
    *state += initstate;
    spo_pcg32_next(state, *inc);
}

/* Jump ahead by delta steps of the LCG in O(log delta) (Brown 1994,
 * "Random Number Generation with Arbitrary Strides"). */
void spo_pcg32_advance(uint64_t *state, uint64_t inc, uint64_t delta)
{
    uint64_t cur_mult = SPO_PCG_MULT, cur_plus = inc;
    uint64_t acc_mult = 1u, acc_plus = 0u;
    while (delta > 0)
    {
        if (delta & 1)
        {
            acc_mult *= cur_mult;
            acc_plus = acc_plus * cur_mult + cur_plus;
        }
        cur_plus = (cur_mult + 1) * cur_plus;
        cur_mult *= cur_mult;
        delta /= 2;
    }
    *state = acc_mult * (*state) + acc_plus;
}

/* ------------------------------------------------------- spherical maps */
/* MapSphericalToCartesianCoordinates, src/math_lib.h:864-871 */
void spo_spherical_to_cartesian(const float *sc, int tx, float *out3)
{
    out3[0] = spo_sin(sc[1], tx) * spo_cos(sc[0], tx);
    out3[2] = spo_sin(sc[1], tx) * spo_sin(sc[0], tx);
    out3[1] = spo_cos(sc[1], tx);
}

/* RandomDirectionOnHemisphere, src/math_lib.h:932-946 */
void spo_hemisphere(const float *normal, uint32_t *state, int tx, float *out)
{
    float theta = SPO_PI * spo_random_unilateral(state);
    float phi = SPO_PI * spo_random_bilateral(state);
    float sc[2] = {phi, theta};
    float d[3];
    spo_spherical_to_cartesian(sc, tx, d);
    spo_v3 dir = v3p(d);
    if (v3dot(dir, v3p(normal)) < 0.0f) dir = v3neg(dir);
    v3put(out, dir);
}

/* ToSphericalCoordinates, src/math_lib.h:849-860 -> (azimuth, inclination) */
void spo_to_spherical(const float *v, int tx, float *out2)
{
    float inc = spo_atan2(sqrtf(v[0] * v[0] + v[2] * v[2]), v[1], tx);
    float azimuth = spo_atan2(v[2], v[0], tx);
    out2[0] = azimuth;
    out2[1] = inc;
}

/* MapToEquirectangular, src/math_lib.h:873-884 */
void spo_map_to_equirect(const float *sc, int tx, float *out2)
{
    float x = sc[0];
    if (x < 0.0f) x += 2.0f * SPO_PI;
    out2[0] = x / (2.0f * SPO_PI);
    out2[1] = spo_cos(sc[1], tx) * 0.5f + 0.5f;
}

/* ------------------------------------------------------------ samplers */
/* SampleImageNearest, src/image.h:3-17.  The reference reads one texel past
 * the row (or the image) when u or v is exactly 1; the flat index is clamped
 * to the last texel here, and identically on the GPU. */
void spo_sample_nearest(const float *pixels, uint32_t w, uint32_t h,
                        const float *uv, float *out4)
{
    float fx = uv[0] * (float)w;
    float fy = uv[1] * (float)h;
    /* (u32)Floor(f) with NaN and negatives taken to 0 (undefined in C) */
    fx = floorf(fx);
    fy = floorf(fy);
    uint32_t x = fx >= 0.0f ? (fx < 4294967296.0f ? (uint32_t)fx : 0xFFFFFFFFu) : 0u;
    uint32_t y = fy >= 0.0f ? (fy < 4294967296.0f ? (uint32_t)fy : 0xFFFFFFFFu) : 0u;
    uint64_t idx = (uint64_t)y * w + x;
    uint64_t last = (uint64_t)w * h - 1;
    if (idx > last) idx = last;
    memcpy(out4, pixels + 4 * idx, 16);
}

/* Lerp(vec4), src/math_lib.h:432-436 */
static void lerp4(const float *a, const float *b, float t, float *r)
{
    for (int i = 0; i < 4; ++i) r[i] = a[i] * (1.0f - t) + b[i] * t;
}

/* SampleImageBilinear, src/image.h:34-74 */
void spo_sample_bilinear(const float *pixels, uint32_t w, uint32_t h,
                         const float *uv, float *out4)
{
    float px = (uv[0] * (float)w) - 0.5f;
    float py = (uv[1] * (float)h) - 0.5f;
    px = spo_max(px, 0.0f);
    py = spo_max(py, 0.0f);
    uint32_t x0 = (uint32_t)floorf(px), x1 = x0 + 1;
    uint32_t y0 = (uint32_t)floorf(py), y1 = y0 + 1;
    float fx = px - (float)x0;
    float fy = py - (float)y0;
    x1 = spo_minu(x1, w - 1);
    y1 = spo_minu(y1, h - 1);
    const float *s0 = pixels + 4 * ((uint64_t)y0 * w + x0);
    const float *s1 = pixels + 4 * ((uint64_t)y0 * w + x1);
    const float *s2 = pixels + 4 * ((uint64_t)y1 * w + x0);
    const float *s3 = pixels + 4 * ((uint64_t)y1 * w + x1);
    float t0[4], t1[4];
    lerp4(s0, s1, fx, t0);
    lerp4(s2, s3, fx, t1);
    lerp4(t0, t1, fy, out4);
}

/* MapCubeMapLayerIndexToBasisVectors, src/cubemap.cpp:55-103 -> fwd,up,right */
void spo_cubemap_basis(uint32_t face, float *out9)
{
    static const float B[6][9] = {
        {1, 0, 0, 0, 1, 0, 0, 0, -1},  {-1, 0, 0, 0, 1, 0, 0, 0, 1},
        {0, 1, 0, 0, 0, -1, 1, 0, 0},  {0, -1, 0, 0, 0, 1, 1, 0, 0},
        {0, 0, 1, 0, 1, 0, 1, 0, 0},   {0, 0, -1, 0, 1, 0, -1, 0, 0},
    };
    memcpy(out9, B[face % 6], sizeof B[0]);
}

/* CreateCubeMap, src/cubemap.cpp:237-291: texel (x, y) of face f holds the
 * equirect sample at dir = norm(fwd + right*(2x/W-1) + up*(2(1-y/H)-1)). */
void spo_create_cubemap(const float *equirect, uint32_t ew, uint32_t eh,
                        uint32_t w, uint32_t h, float *out)
{
    for (uint32_t f = 0; f < 6; ++f)
    {
        float b[9];
        spo_cubemap_basis(f, b);
        spo_v3 fwd = v3p(b), up = v3p(b + 3), right = v3p(b + 6);
        for (uint32_t y = 0; y < h; ++y)
            for (uint32_t x = 0; x < w; ++x)
            {
                float fx = (float)x / (float)w;
                float fy = (float)y / (float)h;
                fy = 1.0f - fy;
                fx = fx * 2.0f - 1.0f;
                fy = fy * 2.0f - 1.0f;
                spo_v3 dir = v3add(v3add(fwd, v3mul(right, fx)), v3mul(up, fy));
                dir = v3norm(dir);
                float dd[3], sc[2], uv[2];
                v3put(dd, dir);
                spo_to_spherical(dd, SPO_TX_LIBM, sc);
                spo_map_to_equirect(sc, SPO_TX_LIBM, uv);
                uv[1] = 1.0f - uv[1];
                spo_sample_bilinear(equirect, ew, eh, uv,
                                    out + 4 * (((uint64_t)f * h + y) * w + x));
            }
    }
}

/* CreateIrradianceCubeMap, src/cubemap.cpp:108-233, with the production
 * switches of src/config.h (IRRADIANCE_CUBEMAP_USE_UNIFORM_SAMPLING 1,
 * RADIANCE_CLAMP 10; clampValue <= 0 restates a build without the clamp).
 * Texel (x, y) of face f: normal n = dir as in CreateCubeMap, tangent
 * t = norm(up x n), bitangent b = norm(n x t); for phi = 0, 0.1, .. < 2 pi
 * and theta = 0, 0.1, .. < pi/2 (float loop counters, as written), the
 * direction (n*cos(theta) + t*sin(theta)cos(phi)) + b*sin(theta)sin(phi)
 * samples the equirect image bilinearly (v flipped), clamped to
 * [0, clampValue]; irradiance += L*cos(theta)*sin(theta) in loop order;
 * texel = (pi*irradiance*(1/count), 1). */
void spo_create_irradiance_cubemap(const float *equirect, uint32_t ew, uint32_t eh,
                                   uint32_t w, uint32_t h, float clampValue, int tx,
                                   float *out)
{
    const float sampleDelta = 0.1f;
    for (uint32_t f = 0; f < 6; ++f)
    {
        float b[9];
        spo_cubemap_basis(f, b);
        spo_v3 fwd = v3p(b), up = v3p(b + 3), right = v3p(b + 6);
        for (uint32_t y = 0; y < h; ++y)
            for (uint32_t x = 0; x < w; ++x)
            {
                float fx = (float)x / (float)w;
                float fy = (float)y / (float)h;
                fy = 1.0f - fy;
                fx = fx * 2.0f - 1.0f;
                fy = fy * 2.0f - 1.0f;
                spo_v3 n = v3norm(v3add(v3add(fwd, v3mul(right, fx)), v3mul(up, fy)));
                spo_v3 t = v3norm(v3cross(up, n));
                spo_v3 bt = v3norm(v3cross(n, t));
                spo_v3 irr = v3s(0.0f);
                uint32_t count = 0;
                for (float phi = 0.0f; phi < 2.0f * SPO_PI; phi += sampleDelta)
                    for (float theta = 0.0f; theta < 0.5f * SPO_PI; theta += sampleDelta)
                    {
                        float sc[2] = {phi, theta}, td[3];
                        spo_spherical_to_cartesian(sc, tx, td);
                        spo_v3 wd = v3add(v3add(v3mul(n, td[1]), v3mul(t, td[0])),
                                          v3mul(bt, td[2]));
                        float dd[3], sph[2], uv[2], smp[4];
                        v3put(dd, wd);
                        spo_to_spherical(dd, tx, sph);
                        spo_map_to_equirect(sph, tx, uv);
                        uv[1] = 1.0f - uv[1];
                        spo_sample_bilinear(equirect, ew, eh, uv, smp);
                        spo_v3 L = {smp[0], smp[1], smp[2]};
                        if (clampValue > 0.0f)
                            L = (spo_v3){spo_clamp(L.x, 0.0f, clampValue),
                                         spo_clamp(L.y, 0.0f, clampValue),
                                         spo_clamp(L.z, 0.0f, clampValue)};
                        irr = v3add(irr, v3mul(v3mul(L, spo_cos(theta, tx)),
                                               spo_sin(theta, tx)));
                        count++;
                    }
                irr = v3mul(v3mul(irr, SPO_PI), 1.0f / (float)count);
                float *o = out + 4 * (((uint64_t)f * h + y) * w + x);
                o[0] = irr.x;
                o[1] = irr.y;
                o[2] = irr.z;
                o[3] = 1.0f;
            }
    }
}

/* PerformToneMapping, src/shaders/post_processing.frag.glsl:19-26, and an
 * UNORM8 write (round to nearest even): c *= exposure, c / (1 + c),
 * pow(c, 1/2.2); NaN and negative results give 0.  tx: SPO_TX_CR pow in
 * double rounded once, else powf. */
void spo_tonemap(const float *rgba, uint32_t pixels, float exposure, int tx,
                 uint8_t *out)
{
    for (uint32_t i = 0; i < pixels; ++i)
    {
        for (int k = 0; k < 3; ++k)
        {
            float x = rgba[4 * i + k] * exposure;
            x = x / (1.0f + x);
            float g = tx == SPO_TX_CR ? (float)pow((double)x, (double)(1.0f / 2.2f))
                                      : powf(x, 1.0f / 2.2f);
            g = spo_clamp(g, 0.0f, 1.0f);
            out[4 * i + k] = (uint8_t)(uint32_t)rintf(g * 255.0f);
        }
        out[4 * i + 3] = 255;
    }
}

/* Progressive accumulation: dst = (dst * before + src * added) / total. */
void spo_accumulate(float *dst, const float *src, uint32_t pixels, uint32_t before,
                    uint32_t added)
{
    float b = (float)before, a = (float)added, t = (float)before + (float)added;
    for (uint32_t i = 0; i < pixels; ++i)
    {
        for (int k = 0; k < 3; ++k)
            dst[4 * i + k] = (dst[4 * i + k] * b + src[4 * i + k] * a) / t;
        dst[4 * i + 3] = 1.0f;
    }
}

/* Nearest cube-map lookup by direction (no reference counterpart: the
 * reference tracer samples an equirect image).  The inverse of CreateCubeMap's
 * texel-to-direction map: major axis picks the face (ties: x, then y, then
 * z), fx = dot(d,right)/|major|, fy = dot(d,up)/|major|, then the nearest
 * texel corner x = floor((fx+1)*W/2 + 1/2), y = floor((1-fy)*H/2 + 1/2),
 * clamped to the face. */
/* Face and face-plane coordinates of a direction (the inverse of
 * CreateCubeMap's texel-to-direction map, src/cubemap.cpp:237-291, with the
 * face bases of src/cubemap.cpp:55-103). */
static void cube_coord(const float *dir, uint32_t *faceOut, float *fxOut,
                       float *fyOut)
{
    float x = dir[0], y = dir[1], z = dir[2];
    float ax = fabsf(x), ay = fabsf(y), az = fabsf(z);
    uint32_t face;
    float ma, sc, tc;
    if (ax >= ay && ax >= az)
    {
        face = x >= 0.0f ? 0 : 1;
        ma = ax;
        sc = x >= 0.0f ? -z : z;
        tc = y;
    }
    else if (ay >= az)
    {
        face = y >= 0.0f ? 2 : 3;
        ma = ay;
        sc = x;
        tc = y >= 0.0f ? -z : z;
    }
    else
    {
        face = z >= 0.0f ? 4 : 5;
        ma = az;
        sc = z >= 0.0f ? x : -x;
        tc = y;
    }
    *faceOut = face;
    *fxOut = sc / ma;
    *fyOut = tc / ma;
}

/* nearest texel of face-plane point (fx, fy): face-major, rows, columns */
static uint64_t cube_index(uint32_t w, uint32_t h, uint32_t face, float fx,
                           float fy)
{
    float px = (fx + 1.0f) * (0.5f * (float)w) + 0.5f;
    float py = (1.0f - fy) * (0.5f * (float)h) + 0.5f;
    /* clamp in the float domain first so NaN and negatives map to 0 */
    px = px >= 0.0f ? px : 0.0f;
    py = py >= 0.0f ? py : 0.0f;
    px = px < (float)(w - 1) ? px : (float)(w - 1);
    py = py < (float)(h - 1) ? py : (float)(h - 1);
    uint32_t ix = (uint32_t)floorf(px);
    uint32_t iy = (uint32_t)floorf(py);
    return ((uint64_t)face * h + iy) * w + ix;
}

void spo_cubemap_lookup(const float *faces, uint32_t w, uint32_t h,
                        const float *dir, float *out4)
{
    uint32_t face;
    float fx, fy;
    cube_coord(dir, &face, &fx, &fy);
    memcpy(out4, faces + 4 * cube_index(w, h, face, fx, fy), 16);
}

/* ---- importance sampling of a cube-map environment (next-event
 * estimation; no reference counterpart).  Texel (ix, iy) of a w x h face is
 * the face-plane rectangle the lookup above maps to it: fx in
 * [(2ix-1)/w - 1, (2ix+1)/w - 1), fy in (1 - (2iy+1)/h, 1 - (2iy-1)/h],
 * clipped to [-1, 1].  Weight: mean channel value x area / (1 + xc^2 +
 * yc^2)^(3/2) at the rectangle's centre; P_i = weight / sum (double, texel
 * order); pdf_i = P_i / area_i; Vose's alias table with stacks filled in
 * ascending order (A. J. Walker 1977, M. D. Vose 1991). */
static void texel_rect(uint32_t ix, uint32_t iy, uint32_t w, uint32_t h,
                       double *x0, double *x1, double *y0, double *y1)
{
    *x0 = ix == 0 ? -1.0 : (2.0 * ix - 1.0) / w - 1.0;
    *x1 = ix + 1 == w ? 1.0 : (2.0 * ix + 1.0) / w - 1.0;
    *y1 = iy == 0 ? 1.0 : 1.0 - (2.0 * iy - 1.0) / h;
    *y0 = iy + 1 == h ? -1.0 : 1.0 - (2.0 * iy + 1.0) / h;
}

int spo_env_table_build(const float *rgba, uint32_t w, uint32_t h, float *pdf,
                        float *prob, uint32_t *alias)
{
    if (!rgba || w < 2 || h < 2) return -1;
    size_t n = 6 * (size_t)w * h;
    double *wt = (double *)malloc(n * sizeof(double));
    double *area = (double *)malloc(n * sizeof(double));
    double *q = (double *)malloc(n * sizeof(double));
    uint32_t *small = (uint32_t *)malloc(n * sizeof(uint32_t));
    uint32_t *large = (uint32_t *)malloc(n * sizeof(uint32_t));
    double total = 0.0;
    for (size_t i = 0; i < n; ++i)
    {
        uint32_t ix = (uint32_t)(i % w), iy = (uint32_t)((i / w) % h);
        double x0, x1, y0, y1;
        texel_rect(ix, iy, w, h, &x0, &x1, &y0, &y1);
        double a = (x1 - x0) * (y1 - y0);
        double xc = 0.5 * (x0 + x1), yc = 0.5 * (y0 + y1);
        double t = 1.0 + xc * xc + yc * yc;
        double lum = ((double)rgba[4 * i] + (double)rgba[4 * i + 1] +
                      (double)rgba[4 * i + 2]) / 3.0;
        if (!(lum > 0.0) || !isfinite(lum)) lum = 0.0;
        area[i] = a;
        wt[i] = lum * a / (t * sqrt(t));
        total += wt[i];
    }
    int rc = 0;
    if (!(total > 0.0) || !isfinite(total)) rc = -1;
    else
    {
        size_t ns = 0, nl = 0;
        for (size_t i = 0; i < n; ++i)
        {
            double P = wt[i] / total;
            pdf[i] = (float)(P / area[i]);
            q[i] = P * (double)n;
            prob[i] = 1.0f;
            alias[i] = (uint32_t)i;
            if (q[i] < 1.0) small[ns++] = (uint32_t)i;
            else large[nl++] = (uint32_t)i;
        }
        while (ns && nl)
        {
            uint32_t sm = small[--ns], lg = large[--nl];
            prob[sm] = (float)q[sm];
            alias[sm] = lg;
            q[lg] = (q[lg] + q[sm]) - 1.0;
            if (q[lg] < 1.0) small[ns++] = lg;
            else large[nl++] = lg;
        }
    }
    free(wt); free(area); free(q); free(small); free(large);
    return rc;
}

/* Density per steradian of the sampler below at direction dir: pdf_i of the
 * texel the lookup gives times (1 + fx^2 + fy^2)^(3/2). */
float spo_env_pdf(const float *pdfTable, uint32_t w, uint32_t h, const float *dir)
{
    uint32_t face;
    float fx, fy;
    cube_coord(dir, &face, &fx, &fy);
    float t = (1.0f + fx * fx) + fy * fy;
    return pdfTable[cube_index(w, h, face, fx, fy)] * (t * sqrtf(t));
}

/* A direction from the table: texel i = (r * n) >> 32, replaced by its
 * alias unless e1 < prob[i]; then uniform in its rectangle (e2, e3),
 * through the face basis, normalized. */
void spo_env_direction(const float *prob, const uint32_t *alias, uint32_t w,
                       uint32_t h, uint32_t r, float e1, float e2, float e3,
                       float *out3)
{
    uint32_t n = 6 * w * h;
    uint32_t i = (uint32_t)(((uint64_t)r * n) >> 32);
    if (!(e1 < prob[i])) i = alias[i];
    uint32_t ix = i % w, iy = (i / w) % h, face = i / (w * h);
    float x0 = ix == 0 ? -1.0f : (float)(2 * ix - 1) / (float)w - 1.0f;
    float x1 = ix + 1 == w ? 1.0f : (float)(2 * ix + 1) / (float)w - 1.0f;
    float y1 = iy == 0 ? 1.0f : 1.0f - (float)(2 * iy - 1) / (float)h;
    float y0 = iy + 1 == h ? -1.0f : 1.0f - (float)(2 * iy + 1) / (float)h;
    float fx = x0 + e2 * (x1 - x0), fy = y0 + e3 * (y1 - y0);
    spo_v3 v;
    switch (face)
    {
    case 0: v = v3(1.0f, fy, -fx); break;
    case 1: v = v3(-1.0f, fy, fx); break;
    case 2: v = v3(fx, 1.0f, -fy); break;
    case 3: v = v3(fx, -1.0f, fy); break;
    case 4: v = v3(fx, fy, 1.0f); break;
    default: v = v3(-fx, fy, -1.0f); break;
    }
    v3put(out3, v3norm(v));
}

/* ----------------------------------------------------------- tiles etc. */
/* ComputeTiles, src/tile.h:11-42 */
uint32_t spo_compute_tiles(uint32_t W, uint32_t H, uint32_t tw, uint32_t th,
                           uint32_t maxTiles, uint32_t *out4)
{
    uint32_t tileCountY = (uint32_t)ceilf((float)H / (float)th);
    uint32_t tileCountX = (uint32_t)ceilf((float)W / (float)tw);
    for (uint32_t ty = 0; ty < tileCountY; ++ty)
        for (uint32_t tx = 0; tx < tileCountX; ++tx)
        {
            uint32_t index = tx + ty * tileCountX;
            if (index >= maxTiles) break;
            uint32_t *t = out4 + 4 * index;
            t[0] = tx * tw;
            t[1] = ty * th;
            t[2] = spo_minu(t[0] + tw, W);
            t[3] = spo_minu(t[1] + th, H);
        }
    return spo_minu(tileCountY * tileCountX, maxTiles);
}

/* TransformAabb, src/aabb.h:29-57 */
void spo_transform_aabb(const float *mn, const float *mx, const float *pos,
                        const float *q, const float *scale, float *out6)
{
    spo_m4 model = m4mul(m4mul(m4translate(v3p(pos)), m4rotate(qmake(q))),
                         m4scale(v3p(scale)));
    spo_v3 lo = v3p(mn), hi = v3p(mx);
    spo_v3 v[8] = {
        v3(lo.x, lo.y, lo.z), v3(hi.x, lo.y, lo.z), v3(hi.x, lo.y, hi.z),
        v3(lo.x, lo.y, hi.z), v3(lo.x, hi.y, lo.z), v3(hi.x, hi.y, lo.z),
        v3(hi.x, hi.y, hi.z), v3(lo.x, hi.y, hi.z),
    };
    spo_v3 rmin = m4point(model, v[0]), rmax = rmin;
    for (int i = 1; i < 8; ++i)
    {
        spo_v3 p = m4point(model, v[i]);
        rmin = v3min(rmin, p);
        rmax = v3max(rmax, p);
    }
    v3put(out6, rmin);
    v3put(out6 + 3, rmax);
}

/* sp_AddObjectToScene's matrices, src/sp_scene.cpp:91-96 */
void spo_object_matrices(const float *pos, const float *q, const float *scale,
                         float *model16, float *inv16)
{
    spo_q qq = qmake(q);
    spo_m4 m = m4mul(m4mul(m4translate(v3p(pos)), m4rotate(qq)),
                     m4scale(v3p(scale)));
    spo_m4 mi = m4mul(m4mul(m4scale(v3inv(v3p(scale))), m4rotate(qconj(qq))),
                      m4translate(v3neg(v3p(pos))));
    memcpy(model16, m.m, 64);
    memcpy(inv16, mi.m, 64);
}

void spo_rotate_vector(const float *v, const float *q, float *out3)
{
    v3put(out3, qrotate(v3p(v), qmake(q)));
}

/* Quat(axis, angle), src/math_lib.h:556-563 */
void spo_quat_axis_angle(const float *axis, float angle, int tx, float *out4)
{
    spo_v3 a = v3mul(v3p(axis), spo_sin(angle * 0.5f, tx));
    out4[0] = a.x; out4[1] = a.y; out4[2] = a.z;
    out4[3] = spo_cos(angle * 0.5f, tx);
}

void spo_quat_mul(const float *p, const float *q, float *out4)
{
    spo_q r = qmul(qmake(p), qmake(q));
    out4[0] = r.x; out4[1] = r.y; out4[2] = r.z; out4[3] = r.w;
}

/* ------------------------------------------------------------- meshes */
/* CreateUnitVertex, src/mesh_generation.cpp:209-216 */
static void unit_vertex(float *v8, float x, float y, float z)
{
    spo_v3 p = v3norm(v3(x, y, z));
    v8[0] = p.x; v8[1] = p.y; v8[2] = p.z;
    v8[3] = p.x; v8[4] = p.y; v8[5] = p.z;
    v8[6] = p.x; v8[7] = p.y;
}

/* CreateIcosahedronMesh, src/mesh_generation.cpp:220-337: 12 base vertices,
 * 20 faces; each subdivision emits 3 new (unshared) midpoint vertices per
 * face and 4 faces in the order (i,a0,a2) (j,a1,a0) (k,a2,a1) (a0,a1,a2). */
void spo_icosphere(uint32_t level, float *verts8, uint32_t *idx, uint32_t *nv,
                   uint32_t *ni)
{
    float t = (1.0f + sqrtf(5.0f)) * 0.5f;
    const float base[12][3] = {
        {-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0},
        {0, -1, t}, {0, 1, t}, {0, -1, -t}, {0, 1, -t},
        {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1},
    };
    static const uint32_t F0[20][3] = {
        {0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11},
        {1, 5, 9},  {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
        {3, 9, 4},  {3, 4, 2},  {3, 2, 6},   {3, 6, 8},  {3, 8, 9},
        {4, 9, 5},  {2, 4, 11}, {6, 2, 10},  {8, 6, 7},  {9, 8, 1},
    };
    uint32_t vc = 0;
    for (int i = 0; i < 12; ++i, ++vc)
        unit_vertex(verts8 + 8 * vc, base[i][0], base[i][1], base[i][2]);
    uint32_t fc = 20;
    uint32_t *faces = (uint32_t *)malloc(sizeof(uint32_t) * 3 * fc);
    memcpy(faces, F0, sizeof F0);
    for (uint32_t l = 0; l < level; ++l)
    {
        uint32_t *nf = (uint32_t *)malloc(sizeof(uint32_t) * 3 * fc * 4);
        uint32_t nfc = 0;
        for (uint32_t j = 0; j < fc; ++j)
        {
            uint32_t fi = faces[3 * j], fj = faces[3 * j + 1],
                     fk = faces[3 * j + 2];
            uint32_t edges[6] = {fi, fj, fj, fk, fk, fi};
            uint32_t added[3];
            for (int k = 0; k < 3; ++k)
            {
                const float *a = verts8 + 8 * edges[2 * k];
                const float *b = verts8 + 8 * edges[2 * k + 1];
                spo_v3 u = v3add(v3p(a), v3p(b));
                unit_vertex(verts8 + 8 * vc, u.x, u.y, u.z);
                added[k] = vc++;
            }
            uint32_t quads[12] = {fi, added[0], added[2], fj, added[1], added[0],
                                  fk, added[2], added[1], added[0], added[1],
                                  added[2]};
            memcpy(nf + 3 * nfc, quads, sizeof quads);
            nfc += 4;
        }
        free(faces);
        faces = nf;
        fc = nfc;
    }
    memcpy(idx, faces, sizeof(uint32_t) * 3 * fc);
    free(faces);
    *nv = vc;
    *ni = 3 * fc;
}

/* CreatePlaneMesh, src/mesh_generation.cpp:1-21 */
void spo_plane(float *verts8, uint32_t *idx, uint32_t *nv, uint32_t *ni)
{
    static const float V[4][8] = {
        {-0.5f, -0.5f, 0, 0, 0, 1, 0, 0}, {0.5f, -0.5f, 0, 0, 0, 1, 1, 0},
        {0.5f, 0.5f, 0, 0, 0, 1, 1, 1},   {-0.5f, 0.5f, 0, 0, 0, 1, 0, 1},
    };
    static const uint32_t I[6] = {0, 1, 2, 2, 3, 0};
    memcpy(verts8, V, sizeof V);
    memcpy(idx, I, sizeof I);
    *nv = 4;
    *ni = 6;
}

/* sin and cos of each x evaluated in double by the C library and rounded
 * once to float (SPO_TX_CR) */
void spo_sincos_cr_array(const float *x, uint32_t n, float *s, float *c)
{
    for (uint32_t i = 0; i < n; ++i)
    {
        s[i] = (float)sin((double)x[i]);
        c[i] = (float)cos((double)x[i]);
    }
}
