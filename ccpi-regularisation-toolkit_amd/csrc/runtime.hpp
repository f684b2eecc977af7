// runtime.hpp -- host-side runtime of the regulariser library: device
// selection, error reporting, device buffers, the stopping-rule finaliser and
// the iteration driver shared by every solver.
#pragma once
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdint.h>
#include <vector>
#include <math.h>
#include "common.hpp"
#include "../../include/cilreg_ext.h"

namespace cilreg {

// Records the message as the thread's last error and prints it to stderr, as
// the reference's CHECK / checkCudaErrors macros do (regularisers_GPU/shared.h:15-38).
void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
const char* last_error();

#define CILREG_CK(call)                                                                   \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) {                                                           \
            ::cilreg::set_error("%s:%d: %s failed: %s", __FILE__, __LINE__, #call,          \
                                hipGetErrorString(e_));                                   \
            return -1;                                                                    \
        }                                                                                 \
    } while (0)

// hipSetDevice(dev) after checking the device count; -1 with a message otherwise.
int select_device(int dev);

// The library's own stream for this host thread on the current device
// (default priority, created once per thread and device).
hipStream_t device_stream();
// The stream a device-API call runs on: the caller's, or for NULL the
// library's stream ordered after the work queued on the legacy default stream.
hipStream_t caller_stream(void* stream);

// Caching device allocator: blocks freed by a finished solve are kept and
// handed to the next solve that fits (at most 2x its request), so repeated
// solves do not pay hipMalloc/hipFree.  CILREG_NO_CACHE=1 disables caching.
void* pool_alloc(size_t bytes);
void pool_free(void* p);
int pool_release_idle();
// Bytes of the current device's pool held by running solves / kept idle.
void pool_stats(size_t* busy, size_t* idle);
// Test hook: the n-th next pool_alloc on this thread fails (0 = off).
void inject_alloc_failure(int n);
// The drop-in host-pointer entry points free their device memory before
// returning, as the reference's do, unless CILREG_KEEP_CACHE=1.
bool keep_cache_after_host_calls();

// One device allocation carved into aligned pieces; returned to the pool on
// destruction, so every exit path of a solve releases its memory.
class Arena {
public:
    Arena() = default;
    ~Arena();
    Arena(const Arena&) = delete;
    Arena& operator=(const Arena&) = delete;
    // Reserve bytes (256-byte aligned) before allocate(); returns the offset.
    size_t reserve(size_t bytes);
    int allocate();                       // one hipMalloc for everything reserved
    // Carve the reservations from caller-owned device memory of `cap` bytes
    // instead: allocate() then fails if they do not fit and allocates nothing.
    void use_external(void* base, size_t cap) { ext_ = base; ext_cap_ = cap; }
    void* at(size_t off) const { return (char*)base_ + off; }
    size_t bytes() const { return total_; }
private:
    void* base_ = nullptr;
    size_t total_ = 0;
    void* ext_ = nullptr;
    size_t ext_cap_ = 0;
};

// Pinned host mirror of the control block (for stop polling); one cached
// buffer per host thread.
class PinnedCtrl {
public:
    int allocate();
    Ctrl* h = nullptr;
};

// Launch the fixed-order reduction of per-block partials that applies the
// stopping rule for iteration `iter` (one block; deterministic).
// The stopping rule's final reduction for one solve, run by one 256-thread
// workgroup: sums the per-block partials in a fixed order, stores the
// residual and applies the rule.  Shared by the single-solve finaliser and the
// grouped one, so both reach the same decision from the same partials.
__device__ inline void finalize_body(const double* partials, int nblocks, Ctrl* ctrl, float epsil, int iter) {
    __shared__ double red[2 * 256];
    if (ctrl->stopped) return;
    double a = 0, b = 0;
    for (int i = threadIdx.x; i < nblocks; i += 256) {
        a += partials[2 * i];
        b += partials[2 * i + 1];
    }
    red[threadIdx.x] = a;
    red[256 + threadIdx.x] = b;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            red[threadIdx.x] += red[threadIdx.x + s];
            red[256 + threadIdx.x] += red[256 + threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double sd = red[0], su = red[256];
        // re = sqrt(sum d^2) / sqrt(sum u^2), as ROF_TV_core.c:86 (0/0 -> NaN never stops)
        float re = (float)(sqrt(sd) / sqrt(su));
        ctrl->re = re;
        ctrl->sum_d = sd;
        ctrl->sum_u = su;
        int count = ctrl->count;
        if (re < epsil) count++;
        ctrl->count = count;
        if (count > 3) {
            ctrl->stop_iter = iter;
            ctrl->stopped = 1;
        }
    }
}

// The stopping rule for iterations `iter` and `iter` + 1 of a two-iteration
// pass in one workgroup: the same fixed-order sums as finalize_body over each
// partials array, then the rule for `iter` and, unless that stopped the solve,
// for `iter` + 1 -- the state two finalize_body calls in a row leave.
__device__ inline void finalize2_body(const double* p0, const double* p1, int nblocks, Ctrl* ctrl, float epsil,
                                      int iter) {
    __shared__ double red[4 * 256];
    // the flag load overlaps the partials' loads (after a stop they hold
    // stale sums, which are then discarded)
    const int stopped = ctrl->stopped;
    double a0 = 0, b0 = 0, a1 = 0, b1 = 0;
    for (int i = threadIdx.x; i < nblocks; i += 256) {
        a0 += p0[2 * i];
        b0 += p0[2 * i + 1];
        a1 += p1[2 * i];
        b1 += p1[2 * i + 1];
    }
    if (stopped) return;
    red[threadIdx.x] = a0;
    red[256 + threadIdx.x] = b0;
    red[512 + threadIdx.x] = a1;
    red[768 + threadIdx.x] = b1;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            #pragma unroll
            for (int q = 0; q < 4; ++q) red[256 * q + threadIdx.x] += red[256 * q + threadIdx.x + st];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        int count = ctrl->count;
        for (int h = 0; h < 2; ++h) {
            const double sd = red[512 * h], su = red[512 * h + 256];
            const float re = (float)(sqrt(sd) / sqrt(su));
            ctrl->re = re;
            ctrl->sum_d = sd;
            ctrl->sum_u = su;
            if (re < epsil) count++;
            ctrl->count = count;
            if (count > 3) {
                ctrl->stop_iter = iter + h;
                ctrl->stopped = 1;
                break;
            }
        }
    }
}

void launch_finalize(const double* partials, int nblocks, Ctrl* ctrl, float epsil, int iter, hipStream_t s);
// Both iterations of a two-iteration pass in one launch (iter and iter + 1),
// with the result of two launch_finalize calls in a row.
void launch_finalize2(const double* p0, const double* p1, int nblocks, Ctrl* ctrl, float epsil, int iter,
                      hipStream_t s);
// Zero the control block (a small fill kernel, safe under graph capture).
void launch_ctrl_reset(Ctrl* ctrl, hipStream_t s);
// Fill a float/double/16-bit array with zeros, or copy one array to another.
void launch_zero(void* p, size_t bytes, hipStream_t s);
void launch_copy(void* dst, const void* src, size_t bytes, hipStream_t s);

inline float dsqrt_host(float x) { return sqrtf(x); }
inline double dsqrt_host(double x) { return sqrt(x); }

// Per-solve options resolved from the public cilreg_opts.
struct Opts {
    int check_every = 5;   // the reference checks every 5th iteration
    int precision = CILREG_FP32;
    int tile_y = 0;        // 0: auto
    int zchunk = 0;        // 0: auto
    int poll = 16;         // host polls the stop flag every `poll` iterations
    int profile = 0;       // time each iteration with HIP events
    int wide_index = 0;    // force 64-bit element indexing
    int stream_launches = 0;  // batch: launch directly instead of one hipGraph
    int depth = 0;         // iterations per pass where supported (FGP, PD, TGV 3D; FGP 2D below 2^20 pixels: 2 by default; 1 = off)
    int autotune = 0;      // choose tile_y / zchunk / depth by timing (tune.hip) when none is given
    int persistent = 0;    // batch: one persistent work-queue launch per group instead of per-level launches
    // Allocation-free solves: the workspace is carved from this device memory
    // (scratch_bytes long) and the stop flag is not polled through pinned
    // host memory (after a stop the remaining launches return at once).
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
};

// 2D FGP runs two iterations per pass below this many pixels: there the
// one-iteration kernel is launch/latency-bound (256^2: 6.5 -> 3.8 us per
// iteration); from 1024^2 on the pass's ring evaluation costs more than the
// traffic it saves (2048^2: 30.5 vs 32.3 us).
constexpr size_t kFgp2dPassMaxPixels = (size_t)1 << 20;

// z planes per chunk of a grouped (batched / served) 3D job.
constexpr int kGroupZChunk = 32;

// Per-iteration kernel timing of the last profiled solve on this thread.
struct ProfileStats { double mean_ms = 0, min_ms = 0; int count = 0, passes = 0; };
ProfileStats& thread_profile();
Opts resolve_opts(const cilreg_opts* o);

// Default tile height and z-chunk for a volume so the grid has enough blocks
// to cover 256 CUs several times.
void auto_tiling(int nx, int ny, int nz, int* ty, int* zc);

// Number of z chunks that fills whole rounds of resident workgroups:
// minimises rounds * (planes per chunk + prologue cost) for `tiles` xy tiles
// and `resident` concurrently resident workgroups on the device.
int plan_zchunk(long long tiles, int nz, long long resident, double prologue = 1.0);

// Workgroups of `kernel` (with `threads` threads) resident on the whole
// current device at once.
template <class K>
long long resident_blocks(K kernel, int threads) {
    int per_cu = 0, cus = 0, dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess) per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 0;
    (void)hipGetLastError();
    return (long long)per_cu * cus;
}

// Geometry for a marching kernel: explicit zchunk if given, else planned
// from the kernel's residency.
template <class K>
Geom plan_geom(K kernel, int threads, int nx, int ny, int nz, int ty, int zc_hint, double prologue = 1.0) {
    if (zc_hint > 0 || nz <= 1) return make_geom(nx, ny, nz, ty, zc_hint > 0 ? zc_hint : 1);
    const long long tiles = (long long)((nx + TX - 1) / TX) * ((ny + ty - 1) / ty);
    const long long res = resident_blocks(kernel, threads);
    if (res <= 0) {
        int t = ty, zc = 0;
        auto_tiling(nx, ny, nz, &t, &zc);
        return make_geom(nx, ny, nz, ty, zc);
    }
    const int chunks = plan_zchunk(tiles, nz, res, prologue);
    return make_geom(nx, ny, nz, ty, (nz + chunks - 1) / chunks);
}

// Drives `iter` iterations on stream s in passes.  step(k, nk, check0,
// check1) launches the pass that advances iterations k .. k+nk-1 (nk is 1 or,
// when the caller's pass width is 2 and two iterations remain, 2); check0 /
// check1 say whether iteration k / k+1 evaluates the stopping rule, and the
// pass leaves those residual sums in partials0 / partials1.  drive_passes then
// launches the finalisers in iteration order, polls the stop flag every
// opts.poll iterations without draining the queue, and on return has
// synchronised the stream and filled *iters_done and *re.  With opts.profile
// every pass is timed with HIP events; the profile reports time per
// iteration (pass time / nk).
template <class StepFn>
int drive_passes(int iter, int width, float epsil, const Opts& opts, Ctrl* d_ctrl, double* d_partials0,
                 double* d_partials1, int nblocks, hipStream_t s, StepFn&& step, int* iters_done, float* re) {
    PinnedCtrl pin;
    const bool use_tol = epsil != 0.0f;
    const bool polls = use_tol && !opts.scratch;   // an allocation-free solve does not poll
    if (polls && pin.allocate() != 0) return -1;
    launch_ctrl_reset(d_ctrl, s);
    hipEvent_t ev[2] = {nullptr, nullptr};
    int flag_slot_pending[2] = {0, 0};
    if (polls) {
        CILREG_CK(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
        CILREG_CK(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    }
    int poll_idx = 0, since_poll = 0;
    int rc = 0;
    std::vector<hipEvent_t> pev;
    std::vector<int> pass_k, pass_n;
    if (opts.profile) {
        pev.resize(2 * (size_t)iter, nullptr);
        for (auto& e : pev)
            if (hipEventCreate(&e) != hipSuccess) { set_error("hipEventCreate failed"); rc = -1; break; }
    }
    int npass = 0;
    for (int k = 0; k < iter && rc == 0;) {
        const int nk = (width >= 2 && k + 1 < iter) ? 2 : 1;
        const bool c0 = use_tol && (k % opts.check_every == 0);
        const bool c1 = nk == 2 && use_tol && ((k + 1) % opts.check_every == 0);
        if (opts.profile) (void)hipEventRecord(pev[2 * npass], s);
        step(k, nk, c0, c1);
        if (opts.profile) {
            (void)hipEventRecord(pev[2 * npass + 1], s);
            pass_k.push_back(k);
            pass_n.push_back(nk);
        }
        ++npass;
        if (c0 && c1) {
            launch_finalize2(d_partials0, d_partials1, nblocks, d_ctrl, epsil, k, s);
        } else {
            if (c0) launch_finalize(d_partials0, nblocks, d_ctrl, epsil, k, s);
            if (c1) launch_finalize(d_partials1, nblocks, d_ctrl, epsil, k + 1, s);
        }
        k += nk;
        since_poll += nk;
        if (polls && since_poll >= opts.poll && k < iter) {
            since_poll = 0;
            const int slot = poll_idx & 1;
            if (hipMemcpyAsync(&pin.h[slot].stopped, &d_ctrl->stopped, sizeof(int), hipMemcpyDeviceToHost, s) !=
                    hipSuccess ||
                hipEventRecord(ev[slot], s) != hipSuccess) {
                set_error("stop-flag poll failed");
                rc = -1;
                break;
            }
            flag_slot_pending[slot] = 1;
            const int other = slot ^ 1;
            if (flag_slot_pending[other]) {
                if (hipEventSynchronize(ev[other]) != hipSuccess) { set_error("event sync failed"); rc = -1; break; }
                flag_slot_pending[other] = 0;
                if (pin.h[other].stopped) break;   // the remaining launches would be no-ops
            }
            ++poll_idx;
        }
    }
    hipError_t e = hipGetLastError();
    if (rc == 0 && e != hipSuccess) { set_error("kernel launch failed: %s", hipGetErrorString(e)); rc = -1; }
    e = hipStreamSynchronize(s);
    if (rc == 0 && e != hipSuccess) { set_error("solve failed: %s", hipGetErrorString(e)); rc = -1; }
    for (int i = 0; i < 2; ++i) if (ev[i]) (void)hipEventDestroy(ev[i]);
    Ctrl c;
    if (rc == 0) {
        hipError_t ce = hipMemcpy(&c, d_ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost);
        if (ce != hipSuccess) { set_error("control block read failed: %s", hipGetErrorString(ce)); rc = -1; }
    }
    if (opts.profile) {
        ProfileStats& ps = thread_profile();
        ps = ProfileStats();
        // passes that ran (a stop at iteration j leaves later passes as no-ops)
        const int ran = rc == 0 ? (c.stopped ? c.stop_iter + 1 : iter) : 0;
        double sum = 0, mn = 1e300;
        int cnt = 0, its = 0;
        for (size_t q = 0; q < pass_k.size(); ++q) {
            if (pass_k[q] >= ran) break;
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, pev[2 * q], pev[2 * q + 1]) == hipSuccess) {
                sum += ms;
                its += pass_n[q];
                const double per = ms / pass_n[q];
                mn = per < mn ? per : mn;
                ++cnt;
            }
        }
        if (cnt) { ps.mean_ms = sum / its; ps.min_ms = mn; ps.count = its; ps.passes = cnt; }
        for (auto& ev2 : pev) if (ev2) (void)hipEventDestroy(ev2);
    }
    if (rc != 0) return rc;
    *iters_done = c.stopped ? c.stop_iter : iter;
    *re = c.re;
    return 0;
}

// One iteration per pass: step(k, check) launches iteration k.
template <class StepFn>
int drive_iterations(int iter, float epsil, const Opts& opts, Ctrl* d_ctrl, double* d_partials,
                     int nblocks, hipStream_t s, StepFn&& step, int* iters_done, float* re) {
    return drive_passes(iter, 1, epsil, opts, d_ctrl, d_partials, d_partials, nblocks, s,
                        [&](int k, int, bool c0, bool) { step(k, c0); }, iters_done, re);
}

}  // namespace cilreg
