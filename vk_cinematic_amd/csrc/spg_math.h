// Shared host/device FP32 arithmetic for the tracer.
//
// Every expression keeps the association order of the reference's math
// library (src/math_lib.h, src/math_utils.h) and everything is compiled with
// -ffp-contract=off and correctly rounded division/sqrt, so results are
// bit-identical between the host build, the gfx950 build and the reference's
// own x86 build.  Min/Max are compare-and-select like the reference's
// (src/math_utils.h:9-19) -- not v_min/v_max, whose NaN and signed-zero
// handling differ -- except in slab_hit_t, which is used only where no
// operand can be NaN and the two agree.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

#define SPG_HD __host__ __device__ __forceinline__

namespace spg {

constexpr float kPi = 3.14159265359f;          // src/math_utils.h:6
constexpr float kEpsilon = 1.19209290e-07f;    // FLT_EPSILON
constexpr float kF32Max = 3.40282347e+38f;

struct v3 { float x, y, z; };
struct v2 { float x, y; };

// 1 / x rounded to nearest, equal to the IEEE division.  On the device:
// the hardware reciprocal (v_rcp_f32, within one unit in the last place)
// refined by one Newton step with FMA, r = y + y (1 - x y), which rounds to
// the IEEE quotient whenever x and 1/x are normal -- the exponent field of
// x in [2, 252] (every such x is checked on the device by
// spg_check_arithmetic); lanes outside that range (zero, denormals, |x| >=
// 2^126, infinities, NaN) take the IEEE division, behind a wave vote so
// waves without such a lane skip it.  Three VALU operations instead of the
// division's eleven.  On the host: the division.
__device__ __forceinline__ float rcp_newton(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const float y = __builtin_amdgcn_rcpf(x);
#else
    const float y = 1.0f / x;
#endif
    const float e = fmaf(-x, y, 1.0f);
    return fmaf(e, y, y);
}
__device__ __forceinline__ bool rcp_fast_range(float x)
{
    const uint32_t ex = (__float_as_uint(x) >> 23) & 0xFFu;
    return ex >= 2u && ex <= 252u;
}
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __forceinline__ float rcp_cr(float x)
{
    float r = rcp_newton(x);
    const bool slow = !rcp_fast_range(x);
    if (__ballot(slow)) r = slow ? 1.0f / x : r;
    return r;
}
#else
__host__ __device__ inline float rcp_cr(float x) { return 1.0f / x; }
#endif

SPG_HD v3 mk(float x, float y, float z) { return v3{x, y, z}; }
SPG_HD v3 splat(float s) { return v3{s, s, s}; }
SPG_HD float fmin_sel(float a, float b) { return a < b ? a : b; }
SPG_HD float fmax_sel(float a, float b) { return a > b ? a : b; }
SPG_HD float clamp_sel(float x, float lo, float hi)
{
    return fmin_sel(fmax_sel(x, lo), hi);
}
SPG_HD v3 operator+(v3 a, v3 b) { return v3{a.x + b.x, a.y + b.y, a.z + b.z}; }
SPG_HD v3 operator-(v3 a, v3 b) { return v3{a.x - b.x, a.y - b.y, a.z - b.z}; }
SPG_HD v3 operator-(v3 a) { return v3{-a.x, -a.y, -a.z}; }
SPG_HD v3 operator*(v3 a, float s) { return v3{a.x * s, a.y * s, a.z * s}; }
SPG_HD v3 had(v3 a, v3 b) { return v3{a.x * b.x, a.y * b.y, a.z * b.z}; }
// Dot: (x*x + y*y) + z*z, src/math_lib.h:461-465
SPG_HD float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// Cross, src/math_lib.h:467-477
SPG_HD v3 cross(v3 a, v3 b)
{
    return v3{(a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z),
              (a.x * b.y) - (a.y * b.x)};
}
SPG_HD v3 vmin(v3 a, v3 b)
{
    return v3{fmin_sel(a.x, b.x), fmin_sel(a.y, b.y), fmin_sel(a.z, b.z)};
}
SPG_HD v3 vmax(v3 a, v3 b)
{
    return v3{fmax_sel(a.x, b.x), fmax_sel(a.y, b.y), fmax_sel(a.z, b.z)};
}
SPG_HD v3 vclamp(v3 v, float lo, float hi)
{
    return v3{clamp_sel(v.x, lo, hi), clamp_sel(v.y, lo, hi),
              clamp_sel(v.z, lo, hi)};
}
// Normalize with the EPSILON guard, src/math_lib.h:503-512
SPG_HD v3 normalize(v3 v)
{
    float length = sqrtf(dot(v, v));
    v3 r{0.0f, 0.0f, 0.0f};
    if (length > kEpsilon) r = v * rcp_cr(length);
    return r;
}
// Inverse(vec3), src/math_lib.h:911-915
SPG_HD v3 inverse(v3 a) { return v3{rcp_cr(a.x), rcp_cr(a.y), rcp_cr(a.z)}; }

// Rows 0..2 of a column-major mat4 (src/math_lib.h:93-97), i.e. what
// TransformPoint/TransformVector read (src/math_lib.h:748-758).
struct m34 { float r[3][4]; };

// mat4 * vec4 row dot, src/math_lib.h:238-248 and :381-398
SPG_HD float row_dot(const float *row, float x, float y, float z, float w)
{
    return row[0] * x + row[1] * y + row[2] * z + row[3] * w;
}
SPG_HD v3 xform(const m34 &m, v3 p, float w)
{
    return v3{row_dot(m.r[0], p.x, p.y, p.z, w), row_dot(m.r[1], p.x, p.y, p.z, w),
              row_dot(m.r[2], p.x, p.y, p.z, w)};
}
SPG_HD v3 xform_point(const m34 &m, v3 p) { return xform(m, p, 1.0f); }
SPG_HD v3 xform_vector(const m34 &m, v3 d) { return xform(m, d, 0.0f); }

// ---- the reference's 4-wide slab test, one box (src/simd.h:198-271):
// t0=(min-o)*invD, t1=(max-o)*invD, per-axis MINPS/MAXPS (second operand
// unless the compare holds), entry = Max(0, MaxComponent(tmin)),
// exit = MinComponent(tmax); hit <=> entry <= exit.
SPG_HD bool slab_hit(float lox, float loy, float loz, float hix, float hiy,
                     float hiz, v3 o, v3 invD, float *entryOut)
{
    float t0x = (lox - o.x) * invD.x, t1x = (hix - o.x) * invD.x;
    float t0y = (loy - o.y) * invD.y, t1y = (hiy - o.y) * invD.y;
    float t0z = (loz - o.z) * invD.z, t1z = (hiz - o.z) * invD.z;
    float mnx = t0x < t1x ? t0x : t1x, mxx = t0x > t1x ? t0x : t1x;
    float mny = t0y < t1y ? t0y : t1y, mxy = t0y > t1y ? t0y : t1y;
    float mnz = t0z < t1z ? t0z : t1z, mxz = t0z > t1z ? t0z : t1z;
    float entry = fmax_sel(0.0f, fmax_sel(mnx, fmax_sel(mny, mnz)));
    float exit_ = fmin_sel(mxx, fmin_sel(mxy, mxz));
    *entryOut = entry;
    return entry <= exit_;
}

SPG_HD bool finite3(v3 a)
{
    return isfinite(a.x) && isfinite(a.y) && isfinite(a.z);
}

// The slab test on plane distances already formed (the four-wide node
// visit's t = q (s invD) + (p - o) invD for a plane at p + q s): near (nx,
// ny, nz) and far (fx, fy, fz) per axis, entry clamped at 0, with IEEE
// min/max (v_max3 / v_min3 on the GPU instead of a compare and a select per
// operation).  Callers use it only when the distances cannot be NaN (finite
// origin, inverse direction and per-axis factors); there min/max and the
// reference's compare-and-select agree but for the sign of a zero, which no
// comparison sees.
SPG_HD bool slab_hit_t(float nx, float ny, float nz, float fx, float fy, float fz,
                       float *entryOut)
{
    float entry = fmaxf(fmaxf(0.0f, nx), fmaxf(ny, nz));
    float exit_ = fminf(fminf(fx, fy), fz);
    *entryOut = entry;
    return entry <= exit_;
}

// ---- Moller-Trumbore, src/ray_intersection.cpp:156-190, with e1 = b - a
// and e2 = c - a supplied (precomputed with the same subtraction).
// Returns t (or -1 on a miss) and u, v.
// n = cross(e1, e2) is an input: the GPU stores it with the triangle
// (computed with the same operations), the oracle forms it in mt_test.
SPG_HD float mt_test_n(v3 o, v3 d, v3 a, v3 e1, v3 e2, v3 n, float *uo, float *vo)
{
    v3 T = o - a;
    v3 p = cross(d, e2);
    v3 q = cross(T, e1);
    float mx = dot(q, e2), my = dot(p, T), mz = dot(q, d);
    float det = rcp_cr(dot(p, e1));
    float t = det * mx;
    float u = det * my;
    float v = det * mz;
    float w = 1.0f - u - v;
    float f = dot(d, n);
    if (u >= 0.0f && u <= 1.0f && v >= 0.0f && v <= 1.0f && w >= 0.0f &&
        w <= 1.0f && f < 0.0f)
    {
        *uo = u;
        *vo = v;
        return t;
    }
    return -1.0f;
}

SPG_HD float mt_test(v3 o, v3 d, v3 a, v3 e1, v3 e2, float *uo, float *vo)
{
    return mt_test_n(o, d, a, e1, e2, cross(e1, e2), uo, vo);
}

// ---- RNG --------------------------------------------------------------
// XorShift32, src/math_utils.h:190-198
SPG_HD uint32_t xorshift32(uint32_t &x)
{
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    return x;
}
// RandomUnilateral's conversion, src/math_utils.h:200-208
SPG_HD float unilateral(uint32_t x)
{
    float numerator = (float)(x >> 1);
    float denom = (float)(0xFFFFFFFFu >> 1);
    return numerator / denom;
}
// PCG32 XSH-RR (O'Neill 2014), seeding as pcg32_srandom_r
constexpr uint64_t kPcgMult = 6364136223846793005ULL;
SPG_HD uint32_t pcg32_next(uint64_t &state, uint64_t inc)
{
    uint64_t old = state;
    state = old * kPcgMult + inc;
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((-rot) & 31u));
}
SPG_HD void pcg32_seed(uint64_t &state, uint64_t &inc, uint64_t initstate,
                       uint64_t initseq)
{
    state = 0u;
    inc = (initseq << 1u) | 1u;
    pcg32_next(state, inc);
    state += initstate;
    pcg32_next(state, inc);
}
// Jump ahead delta LCG steps in O(log delta) (Brown 1994)
SPG_HD void pcg32_advance(uint64_t &state, uint64_t inc, uint64_t delta)
{
    uint64_t cur_mult = kPcgMult, cur_plus = inc;
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
    state = acc_mult * state + acc_plus;
}
// Per-pixel xorshift32 seed: a 32-bit integer finalizer of
// base + golden-ratio * index, zero remapped (xorshift32 stays at 0).
SPG_HD uint32_t pixel_seed(uint32_t base, uint32_t index)
{
    uint32_t x = base + index * 0x9E3779B9u;
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return x ? x : 0x6D2B79F5u;
}

// ---- transcendentals: evaluated in double, rounded once to float -------
SPG_HD float tsin(float x) { return (float)sin((double)x); }
SPG_HD float tcos(float x) { return (float)cos((double)x); }
SPG_HD float tatan2(float y, float x) { return (float)atan2((double)y, (double)x); }
SPG_HD float tpow(float x, float y) { return (float)pow((double)x, (double)y); }
SPG_HD float texp(float x) { return (float)exp((double)x); }
SPG_HD float tlog(float x) { return (float)log((double)x); }
SPG_HD float texpm1(float x) { return (float)expm1((double)x); }
// sin and cos of one argument from one double-precision evaluation, each
// rounded once to float.  For |x| < 2^20 pi/2: Cody-Waite reduction
// r = x - k pi/2, k = rint(x 2/pi), with pi/2 split into two 33-bit parts
// (k times either is exact) and a remainder (split from pi's decimal
// expansion), then the Taylor series of sin to r^17 and cos to r^18 on
// |r| <= pi/4 (the first omitted terms are below 1e-19 of the result) by
// Horner's rule with FMA, and the quadrant's signs and swap (|x| < 1e-9:
// x and 1 directly).  The double
// results are within a few units in the last place of sin and cos, as
// the library's are, so both round to the same float except when the
// exact value lies within ~2^-50 of a rounding boundary.  Larger |x|: the
// library's sincos.
SPG_HD void tsincos(float x, float *s, float *c)
{
    const double xd = (double)x;
    double sd, cd;
    if (!(fabs(xd) < 1647099.0))  // 2^20 * pi/2; also NaN and infinities
    {
        sincos(xd, &sd, &cd);
        *s = (float)sd;
        *c = (float)cd;
        return;
    }
    if (fabs(xd) < 1e-9)
    {
        // sin x = x and cos x = 1 after rounding to float (keeps the sign
        // of a zero)
        *s = x;
        *c = 1.0f;
        return;
    }
    const double kd = rint(xd * 0.6366197723675814);
    const double P1 = 1.5707963267341256, P2 = 6.077100506303966e-11,
                 P3 = 2.0222662487959506e-21;
    const double r = ((xd - kd * P1) - kd * P2) - kd * P3;
    const double r2 = r * r;
    double ps = 2.8114572543455206e-15;
    ps = fma(ps, r2, -7.647163731819816e-13);
    ps = fma(ps, r2, 1.6059043836821613e-10);
    ps = fma(ps, r2, -2.505210838544172e-08);
    ps = fma(ps, r2, 2.7557319223985893e-06);
    ps = fma(ps, r2, -0.0001984126984126984);
    ps = fma(ps, r2, 0.008333333333333333);
    ps = fma(ps, r2, -0.16666666666666666);
    const double sr = fma(r * r2, ps, r);
    double pc = -1.5619206968586225e-16;
    pc = fma(pc, r2, 4.779477332387385e-14);
    pc = fma(pc, r2, -1.1470745597729725e-11);
    pc = fma(pc, r2, 2.08767569878681e-09);
    pc = fma(pc, r2, -2.755731922398589e-07);
    pc = fma(pc, r2, 2.48015873015873e-05);
    pc = fma(pc, r2, -0.001388888888888889);
    pc = fma(pc, r2, 0.041666666666666664);
    pc = fma(pc, r2, -0.5);
    const double cr = fma(r2, pc, 1.0);
    const int q = (int)((long long)kd & 3);
    sd = q == 0 ? sr : q == 1 ? cr : q == 2 ? -sr : -cr;
    cd = q == 0 ? cr : q == 1 ? -sr : q == 2 ? -cr : sr;
    *s = (float)sd;
    *c = (float)cd;
}
// x^5 in double, rounded once: x^2 is exact in double, x^4 and x^5 carry
// one rounding each (relative error below 2^-51, as pow((double)x, 5.0)
// has below 2^-52), so the float result is the correctly rounded x^5
// unless x^5 lies within 2^-51 of a float rounding boundary
SPG_HD float tpow5(float x)
{
    double x2 = (double)x * (double)x;
    return (float)(x2 * x2 * (double)x);
}

}  // namespace spg
