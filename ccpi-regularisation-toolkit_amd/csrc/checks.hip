// checks.hip -- self-checks of the library's own arithmetic building blocks
// on the device they run on.
#include "runtime.hpp"
#include "../../include/cilreg_ext.h"

namespace cilreg {

namespace {

// Every 32-bit pattern [lo, hi) as a float: sqrt_rn against sqrtf, bit for
// bit (any NaN matching any NaN).  Mismatches and checked values are summed
// per workgroup, then added to the two counters.
__global__ __launch_bounds__(256) void check_sqrt_kernel(unsigned long long lo, unsigned long long hi,
                                                         unsigned long long* counts) {
    __shared__ unsigned long long red[2][4];
    unsigned long long bad = 0, seen = 0;
    for (unsigned long long u = lo + (unsigned long long)blockIdx.x * 256 + threadIdx.x; u < hi;
         u += (unsigned long long)gridDim.x * 256) {
        const float x = __uint_as_float((uint32_t)u);
        const float a = sqrt_rn(x), b = sqrtf(x);
        const bool same = __float_as_uint(a) == __float_as_uint(b) || (a != a && b != b);
        bad += same ? 0 : 1;
        seen += 1;
    }
    #pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        bad += __shfl_xor(bad, off, 64);
        seen += __shfl_xor(seen, off, 64);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { red[0][w] = bad; red[1][w] = seen; }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(&counts[0], red[0][0] + red[0][1] + red[0][2] + red[0][3]);
        atomicAdd(&counts[1], red[1][0] + red[1][1] + red[1][2] + red[1][3]);
    }
}

// Every 32-bit pattern [lo, hi) as a float s: Mixed<float>::rof_sqrt_d(s,
// eps) against (float)sqrt((double)s + eps), bit for bit (any NaN matching any
// NaN).  Negative s (which the fluxes never produce) is skipped.
__global__ __launch_bounds__(256) void check_rof_sqrt_d_kernel(unsigned long long lo, unsigned long long hi,
                                                               double eps, unsigned long long* counts) {
    __shared__ unsigned long long red[2][4];
    unsigned long long bad = 0, seen = 0;
    for (unsigned long long u = lo + (unsigned long long)blockIdx.x * 256 + threadIdx.x; u < hi;
         u += (unsigned long long)gridDim.x * 256) {
        const float x = __uint_as_float((uint32_t)u);
        const float a = Mixed<float>::rof_sqrt_d(x, eps), b = (float)sqrt((double)x + eps);
        const bool same = __float_as_uint(a) == __float_as_uint(b) || (a != a && b != b);
        bad += same ? 0 : 1;
        seen += 1;
    }
    #pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        bad += __shfl_xor(bad, off, 64);
        seen += __shfl_xor(seen, off, 64);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { red[0][w] = bad; red[1][w] = seen; }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(&counts[0], red[0][0] + red[0][1] + red[0][2] + red[0][3]);
        atomicAdd(&counts[1], red[1][0] + red[1][1] + red[1][2] + red[1][3]);
    }
}

}  // namespace

}  // namespace cilreg

using namespace cilreg;

extern "C" {

int cilreg_check_sqrt_rn(unsigned long long* mismatches, unsigned long long* checked) {
    if (!mismatches || !checked) { set_error("null pointer argument"); return -1; }
    hipStream_t s = device_stream();
    if (!s) { set_error("no stream"); return -1; }
    Arena ar;
    const size_t oc = ar.reserve(2 * sizeof(unsigned long long));
    if (ar.allocate() != 0) return -1;
    unsigned long long* d = (unsigned long long*)ar.at(oc);
    launch_zero(d, 2 * sizeof(unsigned long long), s);
    hipLaunchKernelGGL(check_sqrt_kernel, dim3(16384), dim3(256), 0, s, 0ULL, 1ULL << 32, d);
    unsigned long long h[2] = {0, 0};
    CILREG_CK(hipGetLastError());
    CILREG_CK(hipMemcpyAsync(h, d, sizeof h, hipMemcpyDeviceToHost, s));
    CILREG_CK(hipStreamSynchronize(s));
    *mismatches = h[0];
    *checked = h[1];
    return 0;
}

int cilreg_check_rof_sqrt_d(double eps, unsigned long long* mismatches, unsigned long long* checked) {
    if (!mismatches || !checked) { set_error("null pointer argument"); return -1; }
    if (!(eps > 0.0)) { set_error("eps must be positive"); return -1; }
    hipStream_t s = device_stream();
    if (!s) { set_error("no stream"); return -1; }
    Arena ar;
    const size_t oc = ar.reserve(2 * sizeof(unsigned long long));
    if (ar.allocate() != 0) return -1;
    unsigned long long* d = (unsigned long long*)ar.at(oc);
    launch_zero(d, 2 * sizeof(unsigned long long), s);
    // non-negative patterns: +0 .. +inf and the positive NaNs
    hipLaunchKernelGGL(check_rof_sqrt_d_kernel, dim3(16384), dim3(256), 0, s, 0ULL, 1ULL << 31, eps, d);
    unsigned long long h[2] = {0, 0};
    CILREG_CK(hipGetLastError());
    CILREG_CK(hipMemcpyAsync(h, d, sizeof h, hipMemcpyDeviceToHost, s));
    CILREG_CK(hipStreamSynchronize(s));
    *mismatches = h[0];
    *checked = h[1];
    return 0;
}

}  // extern "C"
