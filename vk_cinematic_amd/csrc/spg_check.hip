// Self-check of the device arithmetic the tracer relies on for bit-exact
// results: the fast reciprocal (rcp_cr, spg_math.h) against the IEEE
// division for every one of the 2^32 float bit patterns in its fast range.
#include "spg_internal.h"

namespace spg {
namespace {

__global__ void k_check_rcp(unsigned long long *out)
{
    unsigned long long tested = 0, bad = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; b < (1ull << 32);
         b += stride)
    {
        const float x = __uint_as_float((uint32_t)b);
        if (!rcp_fast_range(x)) continue;
        ++tested;
        const float want = 1.0f / x;
        const float got = rcp_newton(x);
        bad += __float_as_uint(got) != __float_as_uint(want) ? 1ull : 0ull;
    }
    atomicAdd(&out[0], tested);
    atomicAdd(&out[1], bad);
}

}  // namespace
}  // namespace spg

using namespace spg;

extern "C" int spg_check_arithmetic(spg_Device *d, uint64_t *counts2)
{
    if (!d || !counts2) return fail(SPG_E_INVALID, "spg_check_arithmetic: NULL");
    std::lock_guard<std::mutex> g(d->lock);
    HIPCHK(hipSetDevice(d->ordinal));
    unsigned long long *buf = nullptr;
    HIPCHK(hipMalloc(&buf, 16));
    int r = hip_check(hipMemsetAsync(buf, 0, 16, d->stream), "hipMemset");
    if (!r)
    {
        hipLaunchKernelGGL(k_check_rcp, dim3(8192), dim3(256), 0, d->stream, buf);
        r = hip_check(hipGetLastError(), "k_check_rcp");
    }
    unsigned long long h[2] = {0, 0};
    if (!r) r = hip_check(hipMemcpyAsync(h, buf, 16, hipMemcpyDeviceToHost, d->stream), "copy");
    if (!r) r = hip_check(hipStreamSynchronize(d->stream), "k_check_rcp");
    hipFree(buf);
    counts2[0] = h[0];
    counts2[1] = h[1];
    return r;
}
