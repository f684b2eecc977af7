// runtime.hip -- host runtime and the small service kernels of the library.
#include "runtime.hpp"
#include <stdarg.h>
#include <string.h>
#include <mutex>
#include <stdlib.h>

namespace cilreg {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    fprintf(stderr, "cilregcuda: %s\n", g_err);
}

const char* last_error() { return g_err; }

int select_device(int dev) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        set_error("no HIP devices found (%s)", e == hipSuccess ? "count 0" : hipGetErrorString(e));
        return -1;
    }
    if (dev < 0 || dev >= n) {
        set_error("invalid gpu_device %d (%d device(s) present)", dev, n);
        return -1;
    }
    CILREG_CK(hipSetDevice(dev));
    return 0;
}

// Streams are per host thread and device, so calls from different threads
// run concurrently as the reference's default-stream calls from different
// threads do.  A thread's streams go back to a free list when it exits (for
// the next thread to take); they are never destroyed, which keeps thread and
// process teardown free of HIP calls.
namespace {
std::mutex g_stream_mu;
std::vector<std::pair<int, hipStream_t>> g_free_streams;

struct ThreadStreams {
    hipStream_t s[64] = {};
    ~ThreadStreams() {
        std::lock_guard<std::mutex> lk(g_stream_mu);
        for (int d = 0; d < 64; ++d)
            if (s[d]) g_free_streams.push_back({d, s[d]});
    }
};
}  // namespace

hipStream_t device_stream() {
    static thread_local ThreadStreams ts;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    if (!ts.s[dev]) {
        {
            std::lock_guard<std::mutex> lk(g_stream_mu);
            for (size_t i = 0; i < g_free_streams.size(); ++i) {
                if (g_free_streams[i].first == dev) {
                    ts.s[dev] = g_free_streams[i].second;
                    g_free_streams[i] = g_free_streams.back();
                    g_free_streams.pop_back();
                    break;
                }
            }
        }
        if (!ts.s[dev] && hipStreamCreateWithFlags(&ts.s[dev], hipStreamNonBlocking) != hipSuccess) ts.s[dev] = nullptr;
    }
    return ts.s[dev];
}

hipStream_t caller_stream(void* stream) {
    if (stream) return (hipStream_t)stream;
    hipStream_t s = device_stream();
    if (!s) return nullptr;
    // The library stream does not synchronise with the legacy default stream:
    // order it after everything already queued there (e.g. the kernel that
    // produced the input).
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
    const bool ok = hipEventRecord(e, nullptr) == hipSuccess && hipStreamWaitEvent(s, e, 0) == hipSuccess;
    (void)hipEventDestroy(e);
    return ok ? s : nullptr;
}

// ---------------------------------------------------------------------------
// caching device allocator
// ---------------------------------------------------------------------------
namespace {
struct PoolBlock {
    void* p;
    size_t bytes;
    int dev;
    bool busy;
};
std::mutex g_pool_mu;
std::vector<PoolBlock> g_pool;

bool cache_disabled() {
    static int v = -1;
    if (v < 0) {
        const char* e = getenv("CILREG_NO_CACHE");
        v = (e && e[0] && e[0] != '0') ? 1 : 0;
    }
    return v == 1;
}

int free_idle_locked(int dev) {
    int n = 0;
    for (size_t i = 0; i < g_pool.size();) {
        if (!g_pool[i].busy && g_pool[i].dev == dev) {
            (void)hipFree(g_pool[i].p);
            g_pool[i] = g_pool.back();
            g_pool.pop_back();
            ++n;
        } else {
            ++i;
        }
    }
    return n;
}
}  // namespace

// Failure injection for tests: the n-th next pool_alloc on this thread fails
// as an exhausted device would (0 = off).
static thread_local int g_fail_alloc_in = 0;

void inject_alloc_failure(int n) { g_fail_alloc_in = n > 0 ? n : 0; }

void* pool_alloc(size_t bytes) {
    if (g_fail_alloc_in > 0 && --g_fail_alloc_in == 0) return nullptr;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lk(g_pool_mu);
    if (!cache_disabled()) {
        int best = -1;
        for (size_t i = 0; i < g_pool.size(); ++i) {
            const PoolBlock& b = g_pool[i];
            if (b.busy || b.dev != dev || b.bytes < bytes || b.bytes > 2 * bytes) continue;
            if (best < 0 || b.bytes < g_pool[best].bytes) best = (int)i;
        }
        if (best >= 0) {
            g_pool[best].busy = true;
            return g_pool[best].p;
        }
    }
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        if (free_idle_locked(dev) == 0 || hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
    }
    g_pool.push_back(PoolBlock{p, bytes, dev, true});
    return p;
}

void pool_free(void* p) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(g_pool_mu);
    for (size_t i = 0; i < g_pool.size(); ++i) {
        if (g_pool[i].p == p) {
            if (cache_disabled()) {
                (void)hipFree(p);
                g_pool[i] = g_pool.back();
                g_pool.pop_back();
            } else {
                g_pool[i].busy = false;
            }
            return;
        }
    }
}

int pool_release_idle() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return -1;
    std::lock_guard<std::mutex> lk(g_pool_mu);
    free_idle_locked(dev);
    return 0;
}

void pool_stats(size_t* busy, size_t* idle) {
    int dev = 0;
    *busy = *idle = 0;
    if (hipGetDevice(&dev) != hipSuccess) return;
    std::lock_guard<std::mutex> lk(g_pool_mu);
    for (const PoolBlock& b : g_pool) {
        if (b.dev != dev) continue;
        (b.busy ? *busy : *idle) += b.bytes;
    }
}

bool keep_cache_after_host_calls() {
    static int v = -1;
    if (v < 0) {
        const char* e = getenv("CILREG_KEEP_CACHE");
        v = (e && e[0] && e[0] != '0') ? 1 : 0;
    }
    return v == 1;
}

ProfileStats& thread_profile() {
    static thread_local ProfileStats ps;
    return ps;
}

Arena::~Arena() {
    if (base_ && !ext_) pool_free(base_);
}

size_t Arena::reserve(size_t bytes) {
    size_t off = total_;
    total_ += (bytes + 255) & ~(size_t)255;
    return off;
}

int Arena::allocate() {
    if (total_ == 0) total_ = 256;
    if (ext_) {
        if (total_ > ext_cap_) {
            set_error("the solve needs %zu bytes of workspace, above the %zu-byte scratch buffer", total_, ext_cap_);
            return -1;
        }
        base_ = ext_;
        return 0;
    }
    base_ = pool_alloc(total_);
    if (!base_) {
        set_error("device allocation of %zu bytes failed", total_);
        return -1;
    }
    return 0;
}

namespace {
struct PinnedSlot {
    Ctrl* h = nullptr;
    ~PinnedSlot() { if (h) (void)hipHostFree(h); }
};
}  // namespace

int PinnedCtrl::allocate() {
    static thread_local PinnedSlot slot;
    if (!slot.h) CILREG_CK(hipHostMalloc((void**)&slot.h, 2 * sizeof(Ctrl), hipHostMallocDefault));
    h = slot.h;
    memset(h, 0, 2 * sizeof(Ctrl));
    return 0;
}

// ---------------------------------------------------------------------------
// service kernels
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void finalize_kernel(const double* partials, int nblocks, Ctrl* ctrl, float epsil,
                                                       int iter) {
    finalize_body(partials, nblocks, ctrl, epsil, iter);
}

// Both checks of a two-iteration pass in one launch.
__global__ __launch_bounds__(256) void finalize2_kernel(const double* p0, const double* p1, int nblocks, Ctrl* ctrl,
                                                        float epsil, int iter) {
    finalize2_body(p0, p1, nblocks, ctrl, epsil, iter);
}

__global__ void ctrl_reset_kernel(Ctrl* ctrl) {
    if (threadIdx.x == 0) {
        ctrl->stopped = 0;
        ctrl->stop_iter = 0;
        ctrl->count = 0;
        ctrl->re = 0.f;
        ctrl->sum_d = 0;
        ctrl->sum_u = 0;
    }
}

__global__ __launch_bounds__(256) void zero_kernel(uint32_t* p, size_t nwords) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    size_t stride = (size_t)gridDim.x * 256;
    for (; i < nwords; i += stride) p[i] = 0u;
}

__global__ __launch_bounds__(256) void copy_kernel(uint32_t* d, const uint32_t* s, size_t nwords) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    size_t stride = (size_t)gridDim.x * 256;
    for (; i < nwords; i += stride) d[i] = s[i];
}

__global__ void tail_bytes_kernel(unsigned char* d, const unsigned char* s, size_t n) {
    // n < 4: zero (s == null) or copy the last bytes
    for (size_t i = threadIdx.x; i < n; i += blockDim.x) d[i] = s ? s[i] : 0;
}

void launch_finalize(const double* partials, int nblocks, Ctrl* ctrl, float epsil, int iter, hipStream_t s) {
    hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(256), 0, s, partials, nblocks, ctrl, epsil, iter);
}

void launch_finalize2(const double* p0, const double* p1, int nblocks, Ctrl* ctrl, float epsil, int iter,
                      hipStream_t s) {
    hipLaunchKernelGGL(finalize2_kernel, dim3(1), dim3(256), 0, s, p0, p1, nblocks, ctrl, epsil, iter);
}

void launch_ctrl_reset(Ctrl* ctrl, hipStream_t s) {
    hipLaunchKernelGGL(ctrl_reset_kernel, dim3(1), dim3(64), 0, s, ctrl);
}

static unsigned grid_for(size_t nwords) {
    size_t b = (nwords + 255) / 256;
    if (b > 8192) b = 8192;
    if (b == 0) b = 1;
    return (unsigned)b;
}

void launch_zero(void* p, size_t bytes, hipStream_t s) {
    size_t nw = bytes / 4;
    if (nw) hipLaunchKernelGGL(zero_kernel, dim3(grid_for(nw)), dim3(256), 0, s, (uint32_t*)p, nw);
    if (bytes & 3)
        hipLaunchKernelGGL(tail_bytes_kernel, dim3(1), dim3(64), 0, s, (unsigned char*)p + nw * 4,
                           (const unsigned char*)nullptr, bytes & 3);
}

void launch_copy(void* dst, const void* src, size_t bytes, hipStream_t s) {
    size_t nw = bytes / 4;
    if (nw) hipLaunchKernelGGL(copy_kernel, dim3(grid_for(nw)), dim3(256), 0, s, (uint32_t*)dst, (const uint32_t*)src, nw);
    if (bytes & 3)
        hipLaunchKernelGGL(tail_bytes_kernel, dim3(1), dim3(64), 0, s, (unsigned char*)dst + nw * 4,
                           (const unsigned char*)src + nw * 4, bytes & 3);
}

Opts resolve_opts(const cilreg_opts* o) {
    Opts r;
    if (o) {
        if (o->check_every > 0) r.check_every = o->check_every;
        r.precision = o->dual_precision;
        r.tile_y = o->tile_y;
        r.zchunk = o->zchunk;
        r.profile = o->profile;
        r.wide_index = o->wide_index;
        r.stream_launches = o->stream_launches;
        r.depth = o->temporal_depth;
        r.autotune = o->autotune;
        r.persistent = o->persistent;
    }
    return r;
}

int plan_zchunk(long long tiles, int nz, long long resident, double prologue) {
    // prologue: the extra planes a chunk evaluates before its first output
    // plane (about one for the one-iteration kernels)
    int best = 1;
    double best_cost = 1e300;
    for (int c = 1; c <= nz; ++c) {
        const int zc = (nz + c - 1) / c;
        if (c > 1 && (nz + zc - 1) / zc != c) continue;   // same zc as a smaller c
        const long long blocks = tiles * (long long)((nz + zc - 1) / zc);
        const long long rounds = (blocks + resident - 1) / resident;
        const double cost = (double)rounds * (zc + prologue);
        if (cost < best_cost - 1e-9) { best_cost = cost; best = (nz + zc - 1) / zc; }
    }
    return best;
}

void auto_tiling(int nx, int ny, int nz, int* ty, int* zc) {
    int t = (*ty == 4 || *ty == 8) ? *ty : 4;
    *ty = t;
    if (*zc > 0) return;
    if (nz <= 1) { *zc = 1; return; }
    long long tiles = (long long)((nx + TX - 1) / TX) * ((ny + t - 1) / t);
    // aim for >= 2048 workgroups (8 per CU) while marching as deep as possible
    long long want_chunks = (2048 + tiles - 1) / tiles;
    if (want_chunks < 1) want_chunks = 1;
    int c = (int)((nz + want_chunks - 1) / want_chunks);
    if (c < 4) c = 4;
    if (c > 64) c = 64;
    if (c > nz) c = nz;
    *zc = c;
}

}  // namespace cilreg
