// capi.hip -- the C ABI of libcilregcuda.so.
//
// The reference entry points (src/Core/regularisers_GPU/*_core.h) take host
// pointers, own all device memory for the duration of the call and return 0
// or -1.  Here each one uploads the input to HBM, runs the device-resident
// solver of its method and downloads the output; the *_dev extensions skip the
// host copies.
#include "solvers.hpp"
#include "tune.hpp"
#include "../../include/cilregcuda.h"
#include <string.h>

using namespace cilreg;

namespace {

// Drains the stream on an early return, before the device buffers of the
// call go back to the allocator cache (declared after the Arena, so it runs
// first); a completed call has already synchronised.
struct DrainOnError {
    hipStream_t s;
    bool done = false;
    ~DrainOnError() {
        if (!done) (void)hipStreamSynchronize(s);
    }
};

bool host_ptrs_ok(const void* in, const void* out, const void* info) {
    if (!in || !out || !info) {
        set_error("null %s pointer", !in ? "input" : !out ? "output" : "infovector");
        return false;
    }
    return true;
}

// Upload `in` (and optionally the caller's `out` contents), run `solve` on
// device buffers, download the result into `out` unless `keep_out_on_zero_iter`
// and iter == 0.
template <class Solve>
int host_call_impl(int gpu_device, const float* in, float* out, float* info, size_t n, bool upload_out,
                   bool skip_download, Solve&& solve) {
    if (!host_ptrs_ok(in, out, info)) return -1;
    if (select_device(gpu_device) != 0) return -1;
    hipStream_t s = device_stream();
    if (!s) { set_error("cannot create a stream on device %d", gpu_device); return -1; }
    Arena ar;
    const size_t oin = ar.reserve(n * sizeof(float));
    const size_t oout = ar.reserve(n * sizeof(float));
    if (ar.allocate() != 0) return -1;
    DrainOnError drain{s};
    float* d_in = (float*)ar.at(oin);
    float* d_out = (float*)ar.at(oout);
    CILREG_CK(hipMemcpyAsync(d_in, in, n * sizeof(float), hipMemcpyHostToDevice, s));
    if (upload_out) CILREG_CK(hipMemcpyAsync(d_out, out, n * sizeof(float), hipMemcpyHostToDevice, s));
    int done = 0;
    float re = 0.f;
    if (solve(d_in, d_out, s, &done, &re) != 0) return -1;
    if (!skip_download) {
        CILREG_CK(hipMemcpyAsync(out, d_out, n * sizeof(float), hipMemcpyDeviceToHost, s));
    }
    CILREG_CK(hipStreamSynchronize(s));
    drain.done = true;
    info[0] = (float)done;
    info[1] = re;
    return 0;
}

// host_call with a second input uploaded beside the first (FGP-dTV's
// reference image); solve receives both device pointers.
template <class Solve>
int host_call_ref_impl(int gpu_device, const float* in, const float* ref, float* out, float* info, size_t n,
                       bool upload_out, bool skip_download, Solve&& solve) {
    if (!host_ptrs_ok(in, out, info)) return -1;
    if (!ref) { set_error("null reference image pointer"); return -1; }
    if (select_device(gpu_device) != 0) return -1;
    hipStream_t s = device_stream();
    if (!s) { set_error("cannot create a stream on device %d", gpu_device); return -1; }
    Arena ar;
    const size_t oin = ar.reserve(n * sizeof(float));
    const size_t oref = ar.reserve(n * sizeof(float));
    const size_t oout = ar.reserve(n * sizeof(float));
    if (ar.allocate() != 0) return -1;
    DrainOnError drain{s};
    float* d_in = (float*)ar.at(oin);
    float* d_ref = (float*)ar.at(oref);
    float* d_out = (float*)ar.at(oout);
    CILREG_CK(hipMemcpyAsync(d_in, in, n * sizeof(float), hipMemcpyHostToDevice, s));
    CILREG_CK(hipMemcpyAsync(d_ref, ref, n * sizeof(float), hipMemcpyHostToDevice, s));
    if (upload_out) CILREG_CK(hipMemcpyAsync(d_out, out, n * sizeof(float), hipMemcpyHostToDevice, s));
    int done = 0;
    float re = 0.f;
    if (solve(d_in, d_ref, d_out, s, &done, &re) != 0) return -1;
    if (!skip_download) CILREG_CK(hipMemcpyAsync(out, d_out, n * sizeof(float), hipMemcpyDeviceToHost, s));
    CILREG_CK(hipStreamSynchronize(s));
    drain.done = true;
    info[0] = (float)done;
    info[1] = re;
    return 0;
}

// The drop-in entry points own no device memory between calls, like the
// reference's (which allocate and free everything inside each call): the
// workspace goes back to the device once the call's buffers are released,
// unless CILREG_KEEP_CACHE=1 keeps it cached for the next call.
void release_after_host_call() {
    if (!keep_cache_after_host_calls()) (void)pool_release_idle();
}

template <class Solve>
int host_call(int gpu_device, const float* in, float* out, float* info, size_t n, bool upload_out, bool skip_download,
              Solve&& solve) {
    const int rc = host_call_impl(gpu_device, in, out, info, n, upload_out, skip_download, solve);
    release_after_host_call();
    return rc;
}

template <class Solve>
int host_call_ref(int gpu_device, const float* in, const float* ref, float* out, float* info, size_t n,
                  bool upload_out, bool skip_download, Solve&& solve) {
    const int rc = host_call_ref_impl(gpu_device, in, ref, out, info, n, upload_out, skip_download, solve);
    release_after_host_call();
    return rc;
}

bool dims_ok(int nx, int ny, int nz) {
    if (nx <= 0 || ny <= 0 || nz < 0) {
        set_error("invalid dimensions (%d, %d, %d)", nx, ny, nz);
        return false;
    }
    if ((long long)nx * ny > (1LL << 40)) {
        set_error("plane too large (%d x %d)", nx, ny);
        return false;
    }
    return true;
}

hipStream_t pick_stream(void* stream) {
    return caller_stream(stream);
}

}  // namespace

extern "C" {

const char* cilreg_last_error(void) { return last_error(); }

int cilreg_last_profile(double* mean_ms, double* min_ms, int* count) {
    const ProfileStats& ps = thread_profile();
    if (mean_ms) *mean_ms = ps.mean_ms;
    if (min_ms) *min_ms = ps.min_ms;
    if (count) *count = ps.count;
    return ps.count > 0 ? 0 : -1;
}

int cilreg_release_cache(void) { return pool_release_idle(); }

int cilreg_pool_stats(size_t* busy_bytes, size_t* idle_bytes) {
    if (!busy_bytes || !idle_bytes) { set_error("null pointer argument"); return -1; }
    pool_stats(busy_bytes, idle_bytes);
    return 0;
}

void cilreg_inject_alloc_failure(int n) { inject_alloc_failure(n); }

// ---------------------------------------------------------------------------
// reference ABI
// ---------------------------------------------------------------------------
int TV_ROF_GPU_main(float* Input, float* Output, float* infovector, float lambdaPar, int iter, float tau, float epsil,
                    int gpu_device, int N, int M, int Z) {
    if (Z == 0) Z = 1;   // TV_ROF_GPU_core.cu:71
    if (!dims_ok(N, M, Z)) return -1;
    const size_t n = (size_t)N * M * Z;
    Opts o;
    return host_call(gpu_device, Input, Output, infovector, n, false, false,
                     [&](const float* di, float* dout, hipStream_t s, int* done, float* re) {
                         return rof_solve<float, float>(di, dout, lambdaPar, iter, tau, epsil, N, M, Z, o, s, done, re);
                     });
}

int TV_FGP_GPU_main(float* Input, float* Output, float* infovector, float lambdaPar, int iter, float epsil,
                    int methodTV, int nonneg, int gpu_device, int dimX, int dimY, int dimZ) {
    if (!dims_ok(dimX, dimY, dimZ)) return -1;
    const int nz = dimZ <= 1 ? 1 : dimZ;
    const size_t n = (size_t)dimX * dimY * nz;
    Opts o;
    // FGP leaves Output untouched when no iteration runs (FGP_TV_core.c:64)
    return host_call(gpu_device, Input, Output, infovector, n, epsil != 0.0f, iter <= 0,
                     [&](const float* di, float* dout, hipStream_t s, int* done, float* re) {
                         return fgp_solve<float, float>(di, dout, lambdaPar, iter, epsil, methodTV, nonneg, dimX, dimY,
                                                         nz, o, s, done, re);
                     });
}

int TV_PD_GPU_main(float* Input, float* Output, float* infovector, float lambdaPar, int iter, float epsil,
                   float lipschitz_const, int methodTV, int nonneg, int gpu_device, int dimX, int dimY, int dimZ) {
    if (!dims_ok(dimX, dimY, dimZ)) return -1;
    const int nz = dimZ <= 1 ? 1 : dimZ;
    const size_t n = (size_t)dimX * dimY * nz;
    Opts o;
    return host_call(gpu_device, Input, Output, infovector, n, false, false,
                     [&](const float* di, float* dout, hipStream_t s, int* done, float* re) {
                         return pd_solve<float, float>(di, dout, lambdaPar, iter, epsil, lipschitz_const, methodTV,
                                                       nonneg, dimX, dimY, nz, o, s, done, re);
                     });
}

int PDTV_GPU_main(float* Input, float* Output, float* infovector, float lambdaPar, int iter, float epsil,
                  float lipschitz_const, int methodTV, int nonneg, int gpu_device, int dimX, int dimY, int dimZ) {
    return TV_PD_GPU_main(Input, Output, infovector, lambdaPar, iter, epsil, lipschitz_const, methodTV, nonneg,
                          gpu_device, dimX, dimY, dimZ);
}

int TV_SB_GPU_main(float* Input, float* Output, float* infovector, float mu, int iter, float epsil, int methodTV,
                   int gpu_device, int dimX, int dimY, int dimZ) {
    if (!dims_ok(dimX, dimY, dimZ)) return -1;
    const int nz = dimZ <= 1 ? 1 : dimZ;
    const size_t n = (size_t)dimX * dimY * nz;
    Opts o;
    return host_call(gpu_device, Input, Output, infovector, n, false, false,
                     [&](const float* di, float* dout, hipStream_t s, int* done, float* re) {
                         return sb_solve<float, float>(di, dout, mu, iter, epsil, methodTV, dimX, dimY, nz, o, s, done,
                                                       re);
                     });
}

int SB_TV_GPU_main(float* Input, float* Output, float* infovector, float mu, int iter, float epsil, int methodTV,
                   int gpu_device, int dimX, int dimY, int dimZ) {
    return TV_SB_GPU_main(Input, Output, infovector, mu, iter, epsil, methodTV, gpu_device, dimX, dimY, dimZ);
}

int TGV_GPU_main(float* U0, float* U, float* infovector, float lambda, float alpha1, float alpha0, int iterationsNumb,
                 float L2, float epsil, int gpu_device, int dimX, int dimY, int dimZ) {
    if (dimZ == 0) dimZ = 1;
    if (!dims_ok(dimX, dimY, dimZ)) return -1;
    const size_t n = (size_t)dimX * dimY * dimZ;
    Opts o;
    return host_call(gpu_device, U0, U, infovector, n, false, false,
                     [&](const float* di, float* dout, hipStream_t s, int* done, float* re) {
                         return tgv_solve<float, float>(di, dout, lambda, alpha1, alpha0, iterationsNumb, L2, epsil,
                                                        dimX, dimY, dimZ, o, s, done, re);
                     });
}

int dTV_FGP_GPU_main(float* Input, float* InputRef, float* Output, float* infovector, float lambdaPar, int iter,
                     float epsil, float eta, int methodTV, int nonneg, int gpu_device, int dimX, int dimY, int dimZ) {
    if (!dims_ok(dimX, dimY, dimZ)) return -1;
    const int nz = dimZ <= 1 ? 1 : dimZ;
    const size_t n = (size_t)dimX * dimY * nz;
    Opts o;
    // like FGP, Output is left untouched when no iteration runs (FGP_dTV_core.c:76)
    return host_call_ref(gpu_device, Input, InputRef, Output, infovector, n, epsil != 0.0f, iter <= 0,
                         [&](const float* di, const float* dr, float* dout, hipStream_t s, int* done, float* re) {
                             return dtv_solve<float>(di, dr, dout, lambdaPar, iter, epsil, eta, methodTV, nonneg, dimX,
                                                     dimY, nz, o, s, done, re);
                         });
}

int LLT_ROF_GPU_main(float* Input, float* Output, float* infovector, float lambdaROF, float lambdaLLT,
                     int iterationsNumb, float tau, float epsil, int gpu_device, int N, int M, int Z) {
    if (Z == 0) Z = 1;   // LLT_ROF_GPU_core.cu:469
    if (!dims_ok(N, M, Z)) return -1;
    const size_t n = (size_t)N * M * Z;
    Opts o;
    return host_call(gpu_device, Input, Output, infovector, n, false, false,
                     [&](const float* di, float* dout, hipStream_t s, int* done, float* re) {
                         return llt_solve(di, dout, lambdaROF, lambdaLLT, iterationsNumb, tau, epsil, N, M, Z, o, s,
                                          done, re);
                     });
}

int NonlDiff_GPU_main(float* Input, float* Output, float* infovector, float lambdaPar, float sigmaPar,
                      int iterationsNumb, float tau, int penaltytype, float epsil, int gpu_device, int N, int M,
                      int Z) {
    if (Z == 0) Z = 1;
    if (!dims_ok(N, M, Z)) return -1;
    const size_t n = (size_t)N * M * Z;
    Opts o;
    return host_call(gpu_device, Input, Output, infovector, n, false, false,
                     [&](const float* di, float* dout, hipStream_t s, int* done, float* re) {
                         return ndf_solve(di, dout, lambdaPar, sigmaPar, iterationsNumb, tau, penaltytype, epsil, N,
                                          M, Z, o, s, done, re);
                     });
}

int Diffus4th_GPU_main(float* Input, float* Output, float* infovector, float lambdaPar, float sigmaPar,
                       int iterationsNumb, float tau, float epsil, int gpu_device, int N, int M, int Z) {
    if (Z == 0) Z = 1;
    if (!dims_ok(N, M, Z)) return -1;
    const size_t n = (size_t)N * M * Z;
    Opts o;
    return host_call(gpu_device, Input, Output, infovector, n, false, false,
                     [&](const float* di, float* dout, hipStream_t s, int* done, float* re) {
                         return d4_solve(di, dout, lambdaPar, sigmaPar, iterationsNumb, tau, epsil, N, M, Z, o, s,
                                         done, re);
                     });
}

// ---------------------------------------------------------------------------
// device-resident extensions
// ---------------------------------------------------------------------------
#define DEV_PROLOGUE(NX, NY, NZ, METHOD, ISO)                    \
    if (!d_input || !d_output) { set_error("null device pointer"); return -1; } \
    if (!dims_ok(NX, NY, NZ)) return -1;                         \
    const Opts o = tuned_opts(METHOD, NX, NY, (NZ) <= 1 ? 1 : (NZ), ISO, resolve_opts(opts)); \
    hipStream_t s = pick_stream(stream);                         \
    if (!s) { set_error("no stream"); return -1; }               \
    int done = 0;                                                \
    float re = 0.f;                                              \
    int rc = -1;

#define DEV_EPILOGUE()                        \
    if (rc == 0 && infovector) {              \
        infovector[0] = (float)done;          \
        infovector[1] = re;                   \
    }                                         \
    return rc;

int cilreg_rof_dev(const float* d_input, float* d_output, float* infovector, float lambdaPar, int iter, float tau,
                   float epsil, int dimX, int dimY, int dimZ, const cilreg_opts* opts, void* stream) {
    if (dimZ == 0) dimZ = 1;
    DEV_PROLOGUE(dimX, dimY, dimZ, CILREG_ROF, 1)
    rc = rof_solve<float, float>(d_input, d_output, lambdaPar, iter, tau, epsil, dimX, dimY, dimZ, o, s, &done, &re);
    DEV_EPILOGUE()
}

int cilreg_fgp_dev(const float* d_input, float* d_output, float* infovector, float lambdaPar, int iter, float epsil,
                   int methodTV, int nonneg, int dimX, int dimY, int dimZ, const cilreg_opts* opts, void* stream) {
    DEV_PROLOGUE(dimX, dimY, dimZ, CILREG_FGP, methodTV == 0)
    const int nz = dimZ <= 1 ? 1 : dimZ;
    if (o.precision == CILREG_BF16)
        rc = fgp_solve<float, bf16_t>(d_input, d_output, lambdaPar, iter, epsil, methodTV, nonneg, dimX, dimY, nz, o,
                                      s, &done, &re);
    else if (o.precision == CILREG_FP16)
        rc = fgp_solve<float, _Float16>(d_input, d_output, lambdaPar, iter, epsil, methodTV, nonneg, dimX, dimY, nz,
                                        o, s, &done, &re);
    else
        rc = fgp_solve<float, float>(d_input, d_output, lambdaPar, iter, epsil, methodTV, nonneg, dimX, dimY, nz, o, s,
                                     &done, &re);
    DEV_EPILOGUE()
}

int cilreg_pd_dev(const float* d_input, float* d_output, float* infovector, float lambdaPar, int iter, float epsil,
                  float lipschitz_const, int methodTV, int nonneg, int dimX, int dimY, int dimZ,
                  const cilreg_opts* opts, void* stream) {
    DEV_PROLOGUE(dimX, dimY, dimZ, CILREG_PD, methodTV == 0)
    const int nz = dimZ <= 1 ? 1 : dimZ;
    if (o.precision == CILREG_BF16)
        rc = pd_solve<float, bf16_t>(d_input, d_output, lambdaPar, iter, epsil, lipschitz_const, methodTV, nonneg,
                                     dimX, dimY, nz, o, s, &done, &re);
    else if (o.precision == CILREG_FP16)
        rc = pd_solve<float, _Float16>(d_input, d_output, lambdaPar, iter, epsil, lipschitz_const, methodTV, nonneg,
                                       dimX, dimY, nz, o, s, &done, &re);
    else
        rc = pd_solve<float, float>(d_input, d_output, lambdaPar, iter, epsil, lipschitz_const, methodTV, nonneg,
                                    dimX, dimY, nz, o, s, &done, &re);
    DEV_EPILOGUE()
}

int cilreg_sb_dev(const float* d_input, float* d_output, float* infovector, float mu, int iter, float epsil,
                  int methodTV, int dimX, int dimY, int dimZ, const cilreg_opts* opts, void* stream) {
    DEV_PROLOGUE(dimX, dimY, dimZ, -1, 1)
    const int nz = dimZ <= 1 ? 1 : dimZ;
    rc = sb_solve<float, float>(d_input, d_output, mu, iter, epsil, methodTV, dimX, dimY, nz, o, s, &done, &re);
    DEV_EPILOGUE()
}

int cilreg_tgv_dev(const float* d_input, float* d_output, float* infovector, float lambda, float alpha1, float alpha0,
                   int iter, float L2, float epsil, int dimX, int dimY, int dimZ, const cilreg_opts* opts,
                   void* stream) {
    if (dimZ == 0) dimZ = 1;
    DEV_PROLOGUE(dimX, dimY, dimZ, CILREG_TGV, 1)
    if (o.precision == CILREG_BF16)
        rc = tgv_solve<float, bf16_t>(d_input, d_output, lambda, alpha1, alpha0, iter, L2, epsil, dimX, dimY, dimZ, o,
                                      s, &done, &re);
    else if (o.precision == CILREG_FP16)
        rc = tgv_solve<float, _Float16>(d_input, d_output, lambda, alpha1, alpha0, iter, L2, epsil, dimX, dimY, dimZ,
                                        o, s, &done, &re);
    else
        rc = tgv_solve<float, float>(d_input, d_output, lambda, alpha1, alpha0, iter, L2, epsil, dimX, dimY, dimZ, o,
                                     s, &done, &re);
    DEV_EPILOGUE()
}

int cilreg_llt_dev(const float* d_input, float* d_output, float* infovector, float lambdaROF, float lambdaLLT,
                   int iter, float tau, float epsil, int dimX, int dimY, int dimZ, const cilreg_opts* opts,
                   void* stream) {
    if (dimZ == 0) dimZ = 1;
    DEV_PROLOGUE(dimX, dimY, dimZ, -1, 1)
    rc = llt_solve(d_input, d_output, lambdaROF, lambdaLLT, iter, tau, epsil, dimX, dimY, dimZ, o, s, &done, &re);
    DEV_EPILOGUE()
}

int cilreg_ndf_dev(const float* d_input, float* d_output, float* infovector, float lambdaPar, float sigmaPar, int iter,
                   float tau, int penaltytype, float epsil, int dimX, int dimY, int dimZ, const cilreg_opts* opts,
                   void* stream) {
    if (dimZ == 0) dimZ = 1;
    DEV_PROLOGUE(dimX, dimY, dimZ, -1, 1)
    rc = ndf_solve(d_input, d_output, lambdaPar, sigmaPar, iter, tau, penaltytype, epsil, dimX, dimY, dimZ, o, s,
                   &done, &re);
    DEV_EPILOGUE()
}

int cilreg_d4_dev(const float* d_input, float* d_output, float* infovector, float lambdaPar, float sigmaPar, int iter,
                  float tau, float epsil, int dimX, int dimY, int dimZ, const cilreg_opts* opts, void* stream) {
    if (dimZ == 0) dimZ = 1;
    DEV_PROLOGUE(dimX, dimY, dimZ, -1, 1)
    rc = d4_solve(d_input, d_output, lambdaPar, sigmaPar, iter, tau, epsil, dimX, dimY, dimZ, o, s, &done, &re);
    DEV_EPILOGUE()
}

int cilreg_dtv_dev(const float* d_input, const float* d_ref, float* d_output, float* infovector, float lambdaPar,
                   int iter, float epsil, float eta, int methodTV, int nonneg, int dimX, int dimY, int dimZ,
                   const cilreg_opts* opts, void* stream) {
    DEV_PROLOGUE(dimX, dimY, dimZ, -1, 1)
    if (!d_ref) { set_error("null device pointer"); return -1; }
    if (o.precision != CILREG_FP32) { set_error("FGP-dTV keeps its fields in FP32 (dual_precision must be 0)"); return -1; }
    const int nz = dimZ <= 1 ? 1 : dimZ;
    rc = dtv_solve<float>(d_input, d_ref, d_output, lambdaPar, iter, epsil, eta, methodTV, nonneg, dimX, dimY, nz, o, s,
                          &done, &re);
    DEV_EPILOGUE()
}

int cilreg_fgp_dev_f64(const double* d_input, double* d_output, double* infovector, double lambdaPar, int iter,
                       double epsil, int methodTV, int nonneg, int dimX, int dimY, int dimZ, const cilreg_opts* opts,
                       void* stream) {
    DEV_PROLOGUE(dimX, dimY, dimZ, -1, 1)
    const int nz = dimZ <= 1 ? 1 : dimZ;
    rc = fgp_solve<double, double>(d_input, d_output, lambdaPar, iter, epsil, methodTV, nonneg, dimX, dimY, nz, o, s,
                                   &done, &re);
    DEV_EPILOGUE()
}

int cilreg_pd_dev_f64(const double* d_input, double* d_output, double* infovector, double lambdaPar, int iter,
                      double epsil, double lipschitz_const, int methodTV, int nonneg, int dimX, int dimY, int dimZ,
                      const cilreg_opts* opts, void* stream) {
    DEV_PROLOGUE(dimX, dimY, dimZ, -1, 1)
    const int nz = dimZ <= 1 ? 1 : dimZ;
    rc = pd_solve<double, double>(d_input, d_output, lambdaPar, iter, epsil, lipschitz_const, methodTV, nonneg, dimX,
                                  dimY, nz, o, s, &done, &re);
    DEV_EPILOGUE()
}

int cilreg_tgv_dev_f64(const double* d_input, double* d_output, double* infovector, double lambda, double alpha1,
                       double alpha0, int iter, double L2, double epsil, int dimX, int dimY, int dimZ,
                       const cilreg_opts* opts, void* stream) {
    if (dimZ == 0) dimZ = 1;
    DEV_PROLOGUE(dimX, dimY, dimZ, -1, 1)
    rc = tgv_solve<double, double>(d_input, d_output, lambda, alpha1, alpha0, iter, L2, epsil, dimX, dimY, dimZ, o, s,
                                   &done, &re);
    DEV_EPILOGUE()
}

// ---------------------------------------------------------------------------
// warm start: the iterate state in caller memory (FP32 channels of n values)
// ---------------------------------------------------------------------------
int cilreg_fgp_dev_ws(const float* d_input, float* d_output, float* infovector, float lambdaPar, int iter, float epsil,
                      int methodTV, int nonneg, int dimX, int dimY, int dimZ, float* d_state, float* t_state, int warm,
                      const cilreg_opts* opts, void* stream) {
    if (!d_state || !t_state || iter < 1) { set_error("warm start needs d_state, t_state and iter >= 1"); return -1; }
    DEV_PROLOGUE(dimX, dimY, dimZ, CILREG_FGP, methodTV == 0)
    const int nz = dimZ <= 1 ? 1 : dimZ;
    WarmState ws{d_state, t_state, warm != 0};
    if (o.precision == CILREG_BF16)
        rc = fgp_solve<float, bf16_t>(d_input, d_output, lambdaPar, iter, epsil, methodTV, nonneg, dimX, dimY, nz, o,
                                      s, &done, &re, &ws);
    else if (o.precision == CILREG_FP16)
        rc = fgp_solve<float, _Float16>(d_input, d_output, lambdaPar, iter, epsil, methodTV, nonneg, dimX, dimY, nz,
                                        o, s, &done, &re, &ws);
    else
        rc = fgp_solve<float, float>(d_input, d_output, lambdaPar, iter, epsil, methodTV, nonneg, dimX, dimY, nz, o, s,
                                     &done, &re, &ws);
    DEV_EPILOGUE()
}

int cilreg_pd_dev_ws(const float* d_input, float* d_output, float* infovector, float lambdaPar, int iter, float epsil,
                     float lipschitz_const, int methodTV, int nonneg, int dimX, int dimY, int dimZ, float* d_state,
                     int warm, const cilreg_opts* opts, void* stream) {
    if (!d_state || iter < 1) { set_error("warm start needs d_state and iter >= 1"); return -1; }
    DEV_PROLOGUE(dimX, dimY, dimZ, CILREG_PD, methodTV == 0)
    const int nz = dimZ <= 1 ? 1 : dimZ;
    WarmState ws{d_state, nullptr, warm != 0};
    if (o.precision == CILREG_BF16)
        rc = pd_solve<float, bf16_t>(d_input, d_output, lambdaPar, iter, epsil, lipschitz_const, methodTV, nonneg,
                                     dimX, dimY, nz, o, s, &done, &re, &ws);
    else if (o.precision == CILREG_FP16)
        rc = pd_solve<float, _Float16>(d_input, d_output, lambdaPar, iter, epsil, lipschitz_const, methodTV, nonneg,
                                       dimX, dimY, nz, o, s, &done, &re, &ws);
    else
        rc = pd_solve<float, float>(d_input, d_output, lambdaPar, iter, epsil, lipschitz_const, methodTV, nonneg,
                                    dimX, dimY, nz, o, s, &done, &re, &ws);
    DEV_EPILOGUE()
}

int cilreg_sb_dev_ws(const float* d_input, float* d_output, float* infovector, float mu, int iter, float epsil,
                     int methodTV, int dimX, int dimY, int dimZ, float* d_state, int warm, const cilreg_opts* opts,
                     void* stream) {
    if (!d_state || iter < 1) { set_error("warm start needs d_state and iter >= 1"); return -1; }
    DEV_PROLOGUE(dimX, dimY, dimZ, -1, 1)
    const int nz = dimZ <= 1 ? 1 : dimZ;
    WarmState ws{d_state, nullptr, warm != 0};
    rc = sb_solve<float, float>(d_input, d_output, mu, iter, epsil, methodTV, dimX, dimY, nz, o, s, &done, &re, &ws);
    DEV_EPILOGUE()
}

int cilreg_tgv_dev_ws(const float* d_input, float* d_output, float* infovector, float lambda, float alpha1,
                      float alpha0, int iter, float L2, float epsil, int dimX, int dimY, int dimZ, float* d_state,
                      int warm, const cilreg_opts* opts, void* stream) {
    if (!d_state || iter < 1) { set_error("warm start needs d_state and iter >= 1"); return -1; }
    if (dimZ == 0) dimZ = 1;
    DEV_PROLOGUE(dimX, dimY, dimZ, CILREG_TGV, 1)
    WarmState ws{d_state, nullptr, warm != 0};
    if (o.precision == CILREG_BF16)
        rc = tgv_solve<float, bf16_t>(d_input, d_output, lambda, alpha1, alpha0, iter, L2, epsil, dimX, dimY, dimZ, o,
                                      s, &done, &re, &ws);
    else if (o.precision == CILREG_FP16)
        rc = tgv_solve<float, _Float16>(d_input, d_output, lambda, alpha1, alpha0, iter, L2, epsil, dimX, dimY, dimZ,
                                        o, s, &done, &re, &ws);
    else
        rc = tgv_solve<float, float>(d_input, d_output, lambda, alpha1, alpha0, iter, L2, epsil, dimX, dimY, dimZ, o,
                                     s, &done, &re, &ws);
    DEV_EPILOGUE()
}

}  // extern "C"
