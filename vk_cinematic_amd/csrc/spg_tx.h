// Device transcendentals of the tracer's two modes.
#pragma once

#include "spg_math.h"

namespace spg {

// Transcendentals: kDouble evaluates in double and rounds once (bit-
// reproducible across host and device); otherwise device float intrinsics.
template <bool kDouble> struct Tx;
template <> struct Tx<true>
{
    static __device__ __forceinline__ float sin_(float x) { return tsin(x); }
    static __device__ __forceinline__ float cos_(float x) { return tcos(x); }
    static __device__ __forceinline__ float atan2_(float y, float x)
    {
        return tatan2(y, x);
    }
    static __device__ __forceinline__ void sincos_(float x, float *s, float *c)
    {
        tsincos(x, s, c);
    }
    // Schlick's (1 - cos)^5 (the reference's powf(x, 5.0f))
    static __device__ __forceinline__ float pow5_(float x) { return tpow5(x); }
    // the guide's vMF density and sampling
    static __device__ __forceinline__ float exp_(float x) { return texp(x); }
    static __device__ __forceinline__ float log_(float x) { return tlog(x); }
    static __device__ __forceinline__ float expm1_(float x) { return texpm1(x); }
};
template <> struct Tx<false>
{
    static __device__ __forceinline__ float sin_(float x) { return __sinf(x); }
    static __device__ __forceinline__ float cos_(float x) { return __cosf(x); }
    static __device__ __forceinline__ float atan2_(float y, float x)
    {
        return atan2f(y, x);
    }
    static __device__ __forceinline__ void sincos_(float x, float *s, float *c)
    {
        __sincosf(x, s, c);
    }
    static __device__ __forceinline__ float pow5_(float x)
    {
        float x2 = x * x;
        return x2 * x2 * x;
    }
    static __device__ __forceinline__ float exp_(float x) { return expf(x); }
    static __device__ __forceinline__ float log_(float x) { return logf(x); }
    static __device__ __forceinline__ float expm1_(float x) { return expm1f(x); }
};

}  // namespace spg
