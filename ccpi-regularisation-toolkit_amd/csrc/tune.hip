// tune.hip -- the autotuner: chooses the tile height, the z-chunk depth and
// the temporal depth (iterations per pass) of a solve by timing them.
//
// For a (method, dual storage, TV type, volume extent) it runs a short
// fixed-count solve (four iterations) of a synthetic volume of that extent
// for each candidate -- tile height 8 or 4, z chunk planned by occupancy or
// 8 / 16 / 32 planes (grouped jobs: the engine's 32-plane chunks after an untimed
// warm-up solve, eight timed iterations), one or two iterations per pass
// where the method has both -- takes the per-iteration kernel time from HIP
// events (the solves' own profiling), and keeps the fastest.  Results are cached per process,
// keyed by the extent rounded up to 64 x 32 x 32, so a batch or a server pays
// for a size class once.  Every candidate computes the same values bit for bit (the results
// do not depend on the tiling or the pass width), so tuning never changes an
// answer -- only, with a tolerance, the order of the FP64 residual sums.
//
// The choices persist in a text file when the caller asks for it
// (cilreg_tune_save / cilreg_tune_load, or CILREG_TUNE_CACHE=<path>: loaded
// before the first tuning of the process and rewritten after every new
// choice).  Each entry names the device it was measured on (architecture and
// CU count) and is used only on such a device.
#include "solvers.hpp"
#include "tune.hpp"
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

namespace cilreg {

namespace {

struct Key {
    std::string device;   // architecture name and CU count, e.g. "gfx950:256"
    int method, precision, iso, fixed_ty, nx, ny, nz;
    bool operator<(const Key& o) const {
        return std::tie(device, method, precision, iso, fixed_ty, nx, ny, nz) <
               std::tie(o.device, o.method, o.precision, o.iso, o.fixed_ty, o.nx, o.ny, o.nz);
    }
};
std::mutex g_mu;
std::map<Key, TuneChoice> g_cache;
bool g_env_loaded = false;   // CILREG_TUNE_CACHE read (guarded by g_mu)

constexpr const char* kCacheHeader = "# cilreg autotune cache v1";

// Identity of the current device for the cache: its architecture (the first
// field of gcnArchName) and its CU count; looked up once per device.
int device_identity(std::string* id) {
    static std::mutex mu;
    static std::map<int, std::string> known;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        set_error("autotune: no current device");
        return -1;
    }
    std::lock_guard<std::mutex> lk(mu);
    auto it = known.find(dev);
    if (it != known.end()) {
        *id = it->second;
        return 0;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        set_error("autotune: no device properties");
        return -1;
    }
    std::string arch = prop.gcnArchName;
    const size_t colon = arch.find(':');
    if (colon != std::string::npos) arch.resize(colon);
    if (arch.empty()) arch = "unknown";
    for (char& ch : arch)
        if (ch == ' ') ch = '_';   // one whitespace-free field of the file
    *id = known[dev] = arch + ":" + std::to_string(prop.multiProcessorCount);
    return 0;
}

// Writes every cached choice to path (through path.tmp and a rename, so a
// reader never sees half a file); the number written, or -1.  Caller holds g_mu.
int save_locked(const char* path) {
    const std::string tmp = std::string(path) + ".tmp";
    FILE* f = std::fopen(tmp.c_str(), "w");
    if (!f) {
        set_error("autotune cache: cannot write %s", tmp.c_str());
        return -1;
    }
    std::fprintf(f, "%s\n# device method precision iso grouped nx ny nz tile_y zchunk depth ms_per_iter\n", kCacheHeader);
    int n = 0;
    for (const auto& kv : g_cache) {
        const Key& k = kv.first;
        const TuneChoice& c = kv.second;
        std::fprintf(f, "%s %d %d %d %d %d %d %d %d %d %d %.17g\n", k.device.c_str(), k.method, k.precision, k.iso,
                     k.fixed_ty, k.nx, k.ny, k.nz, c.tile_y, c.zchunk, c.depth, c.ms_per_iter);
        ++n;
    }
    const bool ok = std::fflush(f) == 0 && !std::ferror(f);
    if (std::fclose(f) != 0 || !ok || std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        set_error("autotune cache: cannot write %s", path);
        return -1;
    }
    return n;
}

// Reads path into the cache; entries already cached are replaced.  The whole
// file is checked before anything is merged: a malformed line rejects it.
// missing_ok: a file that does not exist loads nothing.  Caller holds g_mu.
int load_locked(const char* path, bool missing_ok) {
    FILE* f = std::fopen(path, "r");
    if (!f) {
        if (missing_ok) return 0;
        set_error("autotune cache: cannot read %s", path);
        return -1;
    }
    std::vector<std::pair<Key, TuneChoice>> got;
    char line[512];
    int lineno = 0;
    bool header = false, bad = false;
    while (std::fgets(line, sizeof line, f)) {
        ++lineno;
        if (!std::strchr(line, '\n') && !std::feof(f)) { bad = true; break; }   // line too long
        if (lineno == 1) {
            header = std::strncmp(line, kCacheHeader, std::strlen(kCacheHeader)) == 0;
            if (!header) break;
            continue;
        }
        if (line[0] == '#' || line[0] == '\n') continue;
        char dev[128];
        Key k;
        TuneChoice c;
        int used = 0;
        const int m = std::sscanf(line, "%127s %d %d %d %d %d %d %d %d %d %d %lf %n", dev, &k.method, &k.precision,
                                  &k.iso, &k.fixed_ty, &k.nx, &k.ny, &k.nz, &c.tile_y, &c.zchunk, &c.depth,
                                  &c.ms_per_iter, &used);
        const bool valid = m == 12 && line[used] == '\0' && k.method >= CILREG_ROF && k.method <= CILREG_TGV &&
                           k.precision >= CILREG_FP32 && k.precision <= CILREG_FP16 && (k.iso == 0 || k.iso == 1) &&
                           (k.fixed_ty == 0 || k.fixed_ty == 1) && k.nx > 0 && k.ny > 0 && k.nz > 0 &&
                           (c.tile_y == 4 || c.tile_y == 8) && c.zchunk >= 0 && (c.depth == 1 || c.depth == 2) &&
                           c.ms_per_iter >= 0;
        if (!valid) { bad = true; break; }
        k.device = dev;
        got.emplace_back(k, c);
    }
    std::fclose(f);
    if (!header) {
        set_error("autotune cache %s: not a cilreg autotune cache (first line must be '%s')", path, kCacheHeader);
        return -1;
    }
    if (bad) {
        set_error("autotune cache %s: malformed entry on line %d", path, lineno);
        return -1;
    }
    for (const auto& e : got) g_cache[e.first] = e.second;
    return (int)got.size();
}

// Loads CILREG_TUNE_CACHE once per process (a missing file is an empty
// cache).  Caller holds g_mu.
void env_load_locked() {
    if (g_env_loaded) return;
    g_env_loaded = true;
    const char* path = std::getenv("CILREG_TUNE_CACHE");
    if (path && *path) (void)load_locked(path, true);   // an unreadable cache is reported and ignored
}

// deterministic pseudo-random volume in [0, 1) (hash of the index)
__global__ __launch_bounds__(256) void probe_fill_kernel(float* p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        uint32_t h = (uint32_t)i * 2654435761u;
        h ^= h >> 15;
        h *= 2246822519u;
        h ^= h >> 13;
        __builtin_nontemporal_store((float)(h >> 8) * (1.0f / 16777216.0f), &p[i]);
    }
}

template <typename S>
int run_method(int method, const float* in, float* out, int iso, int nx, int ny, int nz, int iters, const Opts& o,
               hipStream_t s) {
    int done = 0;
    float re = 0.f;
    switch (method) {
        case CILREG_ROF: return rof_solve<float, float>(in, out, 0.02f, iters, 0.001f, 0.0f, nx, ny, nz, o, s, &done, &re);
        case CILREG_FGP:
            return fgp_solve<float, S>(in, out, 0.02f, iters, 0.0f, iso ? 0 : 1, 0, nx, ny, nz, o, s, &done, &re);
        case CILREG_PD:
            return pd_solve<float, S>(in, out, 0.02f, iters, 0.0f, 12.0f, iso ? 0 : 1, 0, nx, ny, nz, o, s, &done, &re);
        default:
            return tgv_solve<float, S>(in, out, 0.02f, 1.0f, 2.0f, iters, 12.0f, 0.0f, nx, ny, nz, o, s, &done, &re);
    }
}

}  // namespace

int autotune(int method, int nx, int ny, int nz, int iso, const Opts& base, bool fixed_ty, TuneChoice* out) {
    *out = TuneChoice{};
    if (method < CILREG_ROF || method > CILREG_TGV) {
        set_error("autotune: unknown method %d", method);
        return -1;
    }
    if (nz <= 1) {   // 2D: one plane per workgroup, 64 x 4 tiles; small FGP images run two iterations per pass
        const bool two = method == CILREG_FGP && (size_t)nx * ny < kFgp2dPassMaxPixels;
        *out = TuneChoice{4, 1, two ? 2 : 1, 0.0};
        return 0;
    }
    // ROF has no dual field: one choice serves every storage precision
    const int kprec = method == CILREG_ROF ? CILREG_FP32 : base.precision;
    std::string dev_id;
    if (device_identity(&dev_id) != 0) return -1;
    const Key key{dev_id, method, kprec, iso ? 1 : 0, fixed_ty ? 1 : 0, (nx + TX - 1) / TX * TX,
                  (ny + 31) / 32 * 32, (nz + 31) / 32 * 32};
    {
        std::lock_guard<std::mutex> lk(g_mu);
        env_load_locked();
        auto it = g_cache.find(key);
        if (it != g_cache.end()) {
            *out = it->second;
            return 0;
        }
    }
    // Grouped execution shares the device with other jobs, so there the
    // choice that matters is the cost per voxel with the device full: the
    // probe then stacks copies of the extent along z up to ~32 M voxels.
    int nzp = nz;
    if (fixed_ty) {
        const size_t per = (size_t)nx * ny * nz, target = (size_t)32 << 20;
        size_t reps = per >= target ? 1 : (target + per - 1) / per;
        while (reps > 1 && per * reps > ((size_t)64 << 20)) --reps;
        nzp = (int)(nz * reps);
    }
    const size_t n = (size_t)nx * ny * nzp;
    hipStream_t s = device_stream();
    if (!s) { set_error("autotune: no stream"); return -1; }
    Arena ar;
    const size_t oi = ar.reserve(n * sizeof(float)), oo = ar.reserve(n * sizeof(float));
    if (ar.allocate() != 0) return -1;
    float* in = (float*)ar.at(oi);
    float* outv = (float*)ar.at(oo);
    {
        size_t blocks = (n + 255) / 256;
        if (blocks > 65536) blocks = 65536;
        hipLaunchKernelGGL(probe_fill_kernel, dim3((unsigned)blocks), dim3(256), 0, s, in, n);
    }
    std::vector<int> depths = {1};
    if (method != CILREG_ROF) depths = {2, 1};
    std::vector<int> tys = fixed_ty ? std::vector<int>{8} : std::vector<int>{8, 4};
    std::vector<int> zcs;
    if (!fixed_ty) zcs.push_back(0);   // the occupancy-planned depth
    // grouped: the engine's fixed chunk depth (only the pass width is tuned)
    if (fixed_ty) zcs.push_back(kGroupZChunk < nz ? kGroupZChunk : nz);
    else
        for (int z : {8, 16, 32})
            if (z <= nz) zcs.push_back(z);
    if (zcs.empty()) zcs.push_back(nz);
    TuneChoice best{8, fixed_ty ? zcs.front() : 0, depths.front(), 1e300};
    const int iters = fixed_ty ? 8 : 4;
    auto run = [&](const Opts& o, int it) {
        if (base.precision == CILREG_BF16) return run_method<bf16_t>(method, in, outv, iso, nx, ny, nzp, it, o, s);
        if (base.precision == CILREG_FP16) return run_method<_Float16>(method, in, outv, iso, nx, ny, nzp, it, o, s);
        return run_method<float>(method, in, outv, iso, nx, ny, nzp, it, o, s);
    };
    for (int d : depths) {
        for (int ty : tys) {
            for (int zc : zcs) {
                Opts o = base;
                o.tile_y = ty;
                o.zchunk = zc;
                o.depth = d;
                o.autotune = 0;
                o.profile = 0;
                // a candidate's first launches run on cold caches and code
                // objects: an untimed solve first, then the timed one
                if (fixed_ty && run(o, 2) != 0) return -1;
                o.profile = 1;
                const int rc = run(o, iters);
                if (rc != 0) return -1;
                const ProfileStats& ps = thread_profile();
                if (ps.count > 0 && ps.mean_ms < best.ms_per_iter) best = TuneChoice{ty, zc, d, ps.mean_ms};
            }
        }
    }
    std::lock_guard<std::mutex> lk(g_mu);
    g_cache[key] = best;
    *out = best;
    const char* path = std::getenv("CILREG_TUNE_CACHE");
    if (path && *path) (void)save_locked(path);   // a failed record is reported; the choice stands
    return 0;
}

Opts tuned_opts(int method, int nx, int ny, int nz, int iso, const Opts& o) {
    if (!o.autotune || o.tile_y != 0 || o.zchunk != 0 || o.depth != 0 || nz <= 1) return o;
    if (method < CILREG_ROF || method > CILREG_TGV) return o;   // methods without tuning candidates
    TuneChoice c;
    if (autotune(method, nx, ny, nz, iso, o, false, &c) != 0) return o;   // keep the defaults
    Opts r = o;
    r.tile_y = c.tile_y;
    r.zchunk = c.zchunk;
    r.depth = c.depth;
    return r;
}

}  // namespace cilreg

using namespace cilreg;

extern "C" CILREG_API int cilreg_autotune(int method, int dimX, int dimY, int dimZ, int methodTV,
                                          const cilreg_opts* base, cilreg_opts* best, double* ms_per_iter) {
    if (dimX <= 0 || dimY <= 0 || dimZ < 0) { set_error("invalid dimensions"); return -1; }
    const Opts o = resolve_opts(base);
    TuneChoice c;
    if (autotune(method, dimX, dimY, dimZ <= 1 ? 1 : dimZ, methodTV == 0, o, false, &c) != 0) return -1;
    if (best) {
        if (base) *best = *base;
        else *best = cilreg_opts{};
        best->tile_y = c.tile_y;
        best->zchunk = c.zchunk;
        best->temporal_depth = c.depth;
        best->autotune = 0;
    }
    if (ms_per_iter) *ms_per_iter = c.ms_per_iter;
    return 0;
}

extern "C" CILREG_API int cilreg_tune_save(const char* path) {
    if (!path || !*path) { set_error("cilreg_tune_save: no path"); return -1; }
    std::lock_guard<std::mutex> lk(g_mu);
    return save_locked(path);
}

extern "C" CILREG_API int cilreg_tune_load(const char* path) {
    if (!path || !*path) { set_error("cilreg_tune_load: no path"); return -1; }
    std::lock_guard<std::mutex> lk(g_mu);
    return load_locked(path, false);
}

extern "C" CILREG_API int cilreg_tune_entries(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    return (int)g_cache.size();
}

extern "C" CILREG_API void cilreg_tune_clear(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_cache.clear();
}
