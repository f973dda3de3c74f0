/* ORACLE -- test infrastructure only (see spo.h).
 *
 * Vector/matrix/quaternion arithmetic restated from src/math_lib.h and
 * src/math_utils.h.  The association order of every expression matches the
 * reference (e.g. Dot is (x*x + y*y) + z*z, src/math_lib.h:461-465), Min/Max
 * are compare-and-select (src/math_utils.h:9-19), and the transcendental
 * functions are routed through spo_sin/cos/atan2/pow so that the caller
 * chooses libm (the reference's own) or double-evaluated rounding.
 */
#ifndef SPO_MATH_H
#define SPO_MATH_H

#include <float.h>
#include <math.h>
#include <stdint.h>
#include "spo.h"

#define SPO_PI 3.14159265359f        /* src/math_utils.h:6 */
#define SPO_EPSILON FLT_EPSILON      /* src/math_utils.h:7 */
#define SPO_F32_MAX FLT_MAX

static inline float spo_min(float a, float b) { return a < b ? a : b; }
static inline float spo_max(float a, float b) { return a > b ? a : b; }
static inline uint32_t spo_minu(uint32_t a, uint32_t b) { return a < b ? a : b; }
/* src/math_utils.h:135-139 */
static inline float spo_clamp(float x, float lo, float hi)
{
    return spo_min(spo_max(x, lo), hi);
}

static inline float spo_sin(float x, int tx)
{
    return tx == SPO_TX_CR ? (float)sin((double)x) : sinf(x);
}
static inline float spo_cos(float x, int tx)
{
    return tx == SPO_TX_CR ? (float)cos((double)x) : cosf(x);
}
static inline float spo_atan2(float y, float x, int tx)
{
    return tx == SPO_TX_CR ? (float)atan2((double)y, (double)x) : atan2f(y, x);
}
/* exp, log, expm1 of the guide's vMF lobes (SPO_TX_CR: double, rounded
 * once, as the product's double-transcendental mode) */
static inline float spo_exp(float x, int tx)
{
    return tx == SPO_TX_CR ? (float)exp((double)x) : expf(x);
}
static inline float spo_log(float x, int tx)
{
    return tx == SPO_TX_CR ? (float)log((double)x) : logf(x);
}
static inline float spo_expm1(float x, int tx)
{
    return tx == SPO_TX_CR ? (float)expm1((double)x) : expm1f(x);
}
/* powf(x, 5.0f) (Schlick's term).  SPO_TX_CR: x^2 exact in double, then
 * x^2 * x^2 * x with one rounding each and one rounding to float (the GPU's
 * tpow5, so the two agree bit for bit; it differs from pow() in double only
 * in the last double bits) */
static inline float spo_pow5(float x, int tx)
{
    if (tx != SPO_TX_CR) return powf(x, 5.0f);
    double x2 = (double)x * (double)x;
    return (float)(x2 * x2 * (double)x);
}

static inline spo_v3 v3(float x, float y, float z)
{
    spo_v3 r = {x, y, z};
    return r;
}
static inline spo_v3 v3s(float s) { return v3(s, s, s); }
static inline spo_v3 v3p(const float *p) { return v3(p[0], p[1], p[2]); }
static inline void v3put(float *d, spo_v3 v)
{
    d[0] = v.x; d[1] = v.y; d[2] = v.z;
}
static inline spo_v3 v3add(spo_v3 a, spo_v3 b)
{
    return v3(a.x + b.x, a.y + b.y, a.z + b.z);
}
static inline spo_v3 v3sub(spo_v3 a, spo_v3 b)
{
    return v3(a.x - b.x, a.y - b.y, a.z - b.z);
}
static inline spo_v3 v3mul(spo_v3 a, float s)
{
    return v3(a.x * s, a.y * s, a.z * s);
}
static inline spo_v3 v3neg(spo_v3 a) { return v3(-a.x, -a.y, -a.z); }
static inline spo_v3 v3had(spo_v3 a, spo_v3 b)
{
    return v3(a.x * b.x, a.y * b.y, a.z * b.z);
}
static inline float v3dot(spo_v3 a, spo_v3 b)
{
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
static inline spo_v3 v3cross(spo_v3 a, spo_v3 b)
{
    spo_v3 r;
    r.x = (a.y * b.z) - (a.z * b.y);
    r.y = (a.z * b.x) - (a.x * b.z);
    r.z = (a.x * b.y) - (a.y * b.x);
    return r;
}
static inline spo_v3 v3min(spo_v3 a, spo_v3 b)
{
    return v3(spo_min(a.x, b.x), spo_min(a.y, b.y), spo_min(a.z, b.z));
}
static inline spo_v3 v3max(spo_v3 a, spo_v3 b)
{
    return v3(spo_max(a.x, b.x), spo_max(a.y, b.y), spo_max(a.z, b.z));
}
static inline spo_v3 v3clamp(spo_v3 v, float lo, float hi)
{
    return v3(spo_clamp(v.x, lo, hi), spo_clamp(v.y, lo, hi),
              spo_clamp(v.z, lo, hi));
}
/* src/math_lib.h:497-512 */
static inline float v3len(spo_v3 v) { return sqrtf(v3dot(v, v)); }
static inline spo_v3 v3norm(spo_v3 v)
{
    float length = v3len(v);
    spo_v3 r = {0, 0, 0};
    if (length > SPO_EPSILON) r = v3mul(v, 1.0f / length);
    return r;
}
/* src/math_lib.h:911-915 */
static inline spo_v3 v3inv(spo_v3 a)
{
    return v3(1.0f / a.x, 1.0f / a.y, 1.0f / a.z);
}
static inline float v3maxc(spo_v3 a) { return spo_max(a.x, spo_max(a.y, a.z)); }
static inline float v3minc(spo_v3 a) { return spo_min(a.x, spo_min(a.y, a.z)); }

/* ---- mat4, column-major (src/math_lib.h:93-97, :201-236, :250-270) ---- */
static inline spo_m4 m4zero(void)
{
    spo_m4 r;
    for (int i = 0; i < 16; ++i) r.m[i] = 0.0f;
    return r;
}
static inline spo_m4 m4identity(void)
{
    spo_m4 r = m4zero();
    r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.0f;
    return r;
}
static inline spo_m4 m4scale(spo_v3 s)
{
    spo_m4 r = m4zero();
    r.m[0] = s.x; r.m[5] = s.y; r.m[10] = s.z; r.m[15] = 1.0f;
    return r;
}
static inline spo_m4 m4translate(spo_v3 t)
{
    spo_m4 r = m4identity();
    r.m[12] = t.x; r.m[13] = t.y; r.m[14] = t.z;
    return r;
}
static inline float dot4(const float *a, const float *b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
}
static inline spo_m4 m4mul(spo_m4 a, spo_m4 b)
{
    spo_m4 r = m4zero();
    for (int i = 0; i < 4; ++i)
    {
        float row[4] = {a.m[i], a.m[4 + i], a.m[8 + i], a.m[12 + i]};
        for (int c = 0; c < 4; ++c) r.m[4 * c + i] = dot4(row, b.m + 4 * c);
    }
    return r;
}
/* src/math_lib.h:381-398 then :748-758 (w = 0 for vectors, 1 for points) */
static inline spo_v3 m4xform(spo_m4 a, spo_v3 p, float w)
{
    float b[4] = {p.x, p.y, p.z, w};
    float r[4];
    for (int i = 0; i < 3; ++i)
    {
        float row[4] = {a.m[i], a.m[4 + i], a.m[8 + i], a.m[12 + i]};
        r[i] = dot4(row, b);
    }
    return v3(r[0], r[1], r[2]);
}
static inline spo_v3 m4point(spo_m4 a, spo_v3 p) { return m4xform(a, p, 1.0f); }
static inline spo_v3 m4vector(spo_m4 a, spo_v3 d) { return m4xform(a, d, 0.0f); }

/* ---- quaternions, xyzw (src/math_lib.h:550-623) ---- */
typedef struct { float x, y, z, w; } spo_q;
static inline spo_q qmake(const float *q)
{
    spo_q r = {q[0], q[1], q[2], q[3]};
    return r;
}
static inline spo_q qmul(spo_q p, spo_q q)
{
    spo_v3 pv = v3(p.x, p.y, p.z), qv = v3(q.x, q.y, q.z);
    spo_v3 v = v3add(v3add(v3mul(qv, p.w), v3mul(pv, q.w)), v3cross(pv, qv));
    spo_q r = {v.x, v.y, v.z, p.w * q.w - v3dot(pv, qv)};
    return r;
}
static inline spo_q qconj(spo_q p)
{
    spo_q r = {-p.x, -p.y, -p.z, p.w};
    return r;
}
static inline spo_v3 qrotate(spo_v3 v, spo_q p)
{
    spo_q q = {v.x, v.y, v.z, 0.0f};
    spo_q r = qmul(qmul(p, q), qconj(p));
    return v3(r.x, r.y, r.z);
}
static inline spo_m4 m4rotate(spo_q q)
{
    spo_m4 r = m4identity();
    float x = q.x, y = q.y, z = q.z, w = q.w;
    r.m[0] = 1.0f - 2.0f * y * y - 2.0f * z * z;
    r.m[1] = 2.0f * x * y + 2.0f * z * w;
    r.m[2] = 2.0f * x * z - 2.0f * y * w;
    r.m[4] = 2.0f * x * y - 2.0f * z * w;
    r.m[5] = 1.0f - 2.0f * x * x - 2.0f * z * z;
    r.m[6] = 2.0f * y * z + 2.0f * x * w;
    r.m[8] = 2.0f * x * z + 2.0f * y * w;
    r.m[9] = 2.0f * y * z - 2.0f * x * w;
    r.m[10] = 1.0f - 2.0f * x * x - 2.0f * y * y;
    return r;
}

#endif
