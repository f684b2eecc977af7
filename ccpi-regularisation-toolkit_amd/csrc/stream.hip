// stream.hip -- the achievable HBM bandwidth of this device, measured with
// streaming kernels written like the solvers' (16-byte vector loads, non-
// temporal vector stores, grid-stride over whole blocks): copy (1 read, 1
// write), read (1 read, a per-block sum stored once), write (1 write) and
// sum2 (2 reads, 1 write, the solvers' typical mix).  The ratio of a solver's
// algorithmic GB/s to these, rather than to the 8 TB/s specification, says
// how close it runs to what the memory system delivers.
#include "runtime.hpp"

namespace cilreg {

typedef float f4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void stream_copy_kernel(f4* __restrict__ d, const f4* __restrict__ s, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        __builtin_nontemporal_store(s[i], &d[i]);
}

__global__ __launch_bounds__(256) void stream_read_kernel(float* __restrict__ out, const f4* __restrict__ s, size_t n) {
    float acc = 0.f;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const f4 v = s[i];
        acc += v.x + v.y + v.z + v.w;
    }
    // one store per thread keeps the loads live (the sum itself is not used)
    out[(size_t)blockIdx.x * 256 + threadIdx.x] = acc;
}

__global__ __launch_bounds__(256) void stream_write_kernel(f4* __restrict__ d, size_t n, float v) {
    const f4 w = {v, v, v, v};
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        __builtin_nontemporal_store(w, &d[i]);
}

__global__ __launch_bounds__(256) void stream_sum2_kernel(f4* __restrict__ d, const f4* __restrict__ a,
                                                          const f4* __restrict__ b, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        __builtin_nontemporal_store(a[i] + b[i], &d[i]);
}

}  // namespace cilreg

using namespace cilreg;

extern "C" CILREG_API int cilreg_stream_bench(size_t bytes, int reps, double* gbps) {
    if (!gbps || reps <= 0 || bytes < (1u << 20)) { set_error("invalid stream bench arguments"); return -1; }
    const size_t n = bytes / 16;   // float4 elements per array
    // two launch shapes per kernel, the better one counts: one element per
    // thread, and a grid-stride loop over 1024 workgroups
    const size_t full = (n + 255) / 256;
    const size_t shapes[2] = {full, full < 1024 ? full : 1024};
    hipStream_t s = device_stream();
    Arena ar;
    const size_t oa = ar.reserve(n * 16), ob = ar.reserve(n * 16), oc = ar.reserve(n * 16),
                 osum = ar.reserve(full * 256 * sizeof(float));
    if (ar.allocate() != 0) return -1;
    f4* A = (f4*)ar.at(oa);
    f4* B = (f4*)ar.at(ob);
    f4* C = (f4*)ar.at(oc);
    float* sums = (float*)ar.at(osum);
    hipEvent_t e0, e1;
    CILREG_CK(hipEventCreate(&e0));
    CILREG_CK(hipEventCreate(&e1));
    int rc = 0;
    const dim3 blk(256);
    hipLaunchKernelGGL(stream_write_kernel, dim3((unsigned)full), blk, 0, s, A, n, 1.0f);
    hipLaunchKernelGGL(stream_write_kernel, dim3((unsigned)full), blk, 0, s, B, n, 2.0f);
    for (int t = 0; t < 4; ++t) gbps[t] = 0.0;
    for (int t = 0; t < 4 && rc == 0; ++t) {
        for (int sh = 0; sh < 2 && rc == 0; ++sh) {
            const dim3 grid((unsigned)shapes[sh]);
            for (int r = -1; r < reps; ++r) {   // r = -1: untimed warm-up
                if (r == 0) (void)hipEventRecord(e0, s);
                switch (t) {
                    case 0: hipLaunchKernelGGL(stream_copy_kernel, grid, blk, 0, s, C, (const f4*)A, n); break;
                    case 1: hipLaunchKernelGGL(stream_read_kernel, grid, blk, 0, s, sums, (const f4*)A, n); break;
                    case 2: hipLaunchKernelGGL(stream_write_kernel, grid, blk, 0, s, C, n, 3.0f); break;
                    default: hipLaunchKernelGGL(stream_sum2_kernel, grid, blk, 0, s, C, (const f4*)A, (const f4*)B, n);
                }
            }
            (void)hipEventRecord(e1, s);
            float ms = 0.f;
            if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                set_error("stream bench timing failed");
                rc = -1;
                break;
            }
            const double per = ms * 1e-3 / reps;
            const double moved[4] = {2.0 * n * 16, 1.0 * n * 16, 1.0 * n * 16, 3.0 * n * 16};
            const double v = moved[t] / per / 1e9;
            if (v > gbps[t]) gbps[t] = v;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    hipError_t e = hipGetLastError();
    if (rc == 0 && e != hipSuccess) { set_error("stream bench failed: %s", hipGetErrorString(e)); rc = -1; }
    return rc;
}
