// noalloc.hip -- the allocation-free FGP-TV entry point.
//
// cilreg_fgp_dev_noalloc is cilreg_fgp_dev for callers that cannot allocate
// device memory (code that must not call hipMalloc, tight serving loops,
// harnesses that forbid it): the solve's workspace -- the dual fields P (three
// or four sets), the two objective buffers D, the stopping-rule control block
// and the per-workgroup residual partials -- is carved from `scratch`, a
// file-scope device array of kScratchBytes, instead of the library's pool.
// Each GPU the library runs on holds its own copy of the array from the
// moment the code object is loaded until the process ends.
//
// On the host the array's name is only a stand-in for its device address:
// the address is obtained once per device with hipGetSymbolAddress (with that
// device current) and only that pointer is handed to kernels.  Calls on one
// device share the array, so the library serialises them (a per-device lock
// held for the whole, synchronous call); the array keeps the last call's
// contents, and every field a solve reads before writing is initialised by
// the solve itself (zero-fill and copy kernels, no memset).
#include "solvers.hpp"
#include "tune.hpp"
#include <mutex>

namespace cilreg {

constexpr size_t kScratchBytes = (size_t)64 << 20;   // 64 MiB per device

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu[kMaxDevices];
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of this device's scratch array (current device `dev`),
// resolved on the first call on that device; the caller holds g_scratch_mu[dev].
void* scratch_on(int dev) {
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("hipGetSymbolAddress of the scratch buffer failed on device %d: %s", dev, hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

}  // namespace

}  // namespace cilreg

using namespace cilreg;

extern "C" {

size_t cilreg_noalloc_scratch_bytes(void) { return kScratchBytes; }

int cilreg_fgp_dev_noalloc(const float* d_input, float* d_output, float* infovector, float lambdaPar, int iter,
                           float epsil, int methodTV, int nonneg, int dimX, int dimY, int dimZ,
                           const cilreg_opts* opts, void* stream) {
    if (!d_input || !d_output) { set_error("null device pointer"); return -1; }
    if (dimX <= 0 || dimY <= 0 || dimZ < 0) { set_error("invalid dimensions (%d, %d, %d)", dimX, dimY, dimZ); return -1; }
    int dev = 0;
    CILREG_CK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) { set_error("device %d out of range", dev); return -1; }
    Opts o = resolve_opts(opts);
    o.autotune = 0;   // timing candidates would allocate
    o.profile = 0;    // per-iteration events would be created
    hipStream_t s = caller_stream(stream);
    if (!s) { set_error("no stream"); return -1; }
    std::lock_guard<std::mutex> lk(g_scratch_mu[dev]);
    void* p = scratch_on(dev);
    if (!p) return -1;
    o.scratch = p;
    o.scratch_bytes = kScratchBytes;
    const int nz = dimZ <= 1 ? 1 : dimZ;
    int done = 0;
    float re = 0.f;
    int rc;
    if (o.precision == CILREG_BF16)
        rc = fgp_solve<float, bf16_t>(d_input, d_output, lambdaPar, iter, epsil, methodTV, nonneg, dimX, dimY, nz, o,
                                      s, &done, &re);
    else if (o.precision == CILREG_FP16)
        rc = fgp_solve<float, _Float16>(d_input, d_output, lambdaPar, iter, epsil, methodTV, nonneg, dimX, dimY, nz,
                                        o, s, &done, &re);
    else
        rc = fgp_solve<float, float>(d_input, d_output, lambdaPar, iter, epsil, methodTV, nonneg, dimX, dimY, nz, o, s,
                                     &done, &re);
    if (rc != 0) {
        // kernels of a failed solve may still be queued on the shared array
        (void)hipStreamSynchronize(s);
        return -1;
    }
    if (infovector) {
        infovector[0] = (float)done;
        infovector[1] = re;
    }
    return 0;
}

}  // extern "C"
