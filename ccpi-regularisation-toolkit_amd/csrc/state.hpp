// state.hpp -- a solver's iterate state in caller memory (warm start).
//
// The caller keeps the state as FP32 channels of n values each; the solver
// converts it into its working buffers (whatever the dual storage type) at the
// start and back at the end, so an outer loop (FISTA, ADMM) can resume the
// inner solver where the previous call left it.
#pragma once
#include "runtime.hpp"

namespace cilreg {

struct WarmState {
    float* d_state;   // channels x n FP32 values on the device
    float* tstate;    // host: FGP's {t_k, multiplier of the next iteration}; unused otherwise
    int warm;         // 1: start from d_state (and tstate); 0: cold start, state only written
};

// stride: element distance in the working buffer (2 for one half of a pair field)
template <typename S>
__global__ __launch_bounds__(256) void state_in_kernel(S* d, const float* s, size_t n, int stride) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        Store<S>::st(d, i * stride, s[i]);
}
template <typename S>
__global__ __launch_bounds__(256) void state_out_kernel(float* d, const S* s, size_t n, int stride) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        d[i] = (float)Store<S>::ld(s, i * stride);
}

inline dim3 state_grid(size_t n) {
    size_t b = (n + 255) / 256;
    return dim3((unsigned)(b > 16384 ? 16384 : (b ? b : 1)));
}
// FP32 state channel -> working buffer in storage type S (every stride-th element)
template <typename S>
void state_in(S* dst, const float* src, size_t n, hipStream_t s, int stride = 1) {
    hipLaunchKernelGGL(state_in_kernel<S>, state_grid(n), dim3(256), 0, s, dst, src, n, stride);
}
// working buffer in storage type S (every stride-th element) -> FP32 state channel
template <typename S>
void state_out(float* dst, const S* src, size_t n, hipStream_t s, int stride = 1) {
    hipLaunchKernelGGL(state_out_kernel<S>, state_grid(n), dim3(256), 0, s, dst, src, n, stride);
}

}  // namespace cilreg
