/*
 * cilreg_ext.h -- entry points of libcilregcuda.so beyond the reference ABI.
 *
 * None of these has a counterpart in the reference; each is for a caller the
 * reference cannot serve:
 *  - *_dev: device-resident solves on caller-owned device buffers and a
 *    caller stream (proximal reconstruction loops, the benchmark, torch
 *    tensors), with options for the residual-check cadence, reduced-precision
 *    storage of the dual fields and the tiling;
 *  - *_dev_f64: the same solvers computing and storing in FP64 (the oracle of
 *    the precision study);
 *  - cilreg_batch_*: many independent volumes in one launch sequence;
 *  - cilreg_last_error: the thread's last error message.
 * All pointers named d_* are device pointers on the current device; `stream`
 * is a hipStream_t (NULL = the library's own stream).  Every call returns 0 on
 * success and -1 on error, and is synchronous at return.
 */
#ifndef CILREG_EXT_H
#define CILREG_EXT_H

#ifndef CILREG_API
#if defined(__GNUC__) || defined(__clang__)
#define CILREG_API __attribute__((visibility("default")))
#else
#define CILREG_API
#endif
#endif

#ifdef __cplusplus
extern "C" {
#endif

enum { CILREG_FP32 = 0, CILREG_BF16 = 1, CILREG_FP16 = 2 };

typedef struct cilreg_opts {
    int check_every;     /* residual-check cadence in iterations; 0 = 5 (the reference's) */
    int dual_precision;  /* CILREG_FP32 / CILREG_BF16 / CILREG_FP16 storage of dual and auxiliary fields */
    int tile_y;          /* tile height in y (4 or 8); 0 = auto */
    int zchunk;          /* z planes marched per workgroup; 0 = auto */
    int profile;         /* 1: time every iteration's kernels with HIP events (see cilreg_last_profile) */
    int wide_index;      /* 1: force 64-bit element indexing (chosen automatically for fields >= 2 GiB) */
    int reserved[2];
} cilreg_opts;

/* Last error recorded on this host thread ("" if none). */
CILREG_API const char *cilreg_last_error(void);

/* Per-iteration kernel time of the last solve on this host thread that ran
 * with opts.profile = 1: mean and minimum in milliseconds over `count`
 * iterations, measured with HIP events on the solve's stream around the
 * iteration's main kernel(s) (the stopping-rule finaliser excluded). */
CILREG_API int cilreg_last_profile(double *mean_ms, double *min_ms, int *count);

/* Frees the device memory the library keeps cached between calls on the
 * current device (solves reuse freed workspaces instead of calling hipMalloc;
 * set CILREG_NO_CACHE=1 to allocate and free per call like the reference). */
CILREG_API int cilreg_release_cache(void);

/* Device-resident solves.  d_input and d_output are device arrays of
 * dimX*dimY*dimZ floats (2D: dimZ = 1); infovector is a host float[2]
 * receiving [iterations, last residual] as in the reference.  For FGP with
 * epsil != 0, d_output's contents seed the first residual like the reference
 * C-OMP Output does. */
CILREG_API int cilreg_rof_dev(const float *d_input, float *d_output, float *infovector, float lambdaPar, int iter, float tau,
                   float epsil, int dimX, int dimY, int dimZ, const cilreg_opts *opts, void *stream);
CILREG_API int cilreg_fgp_dev(const float *d_input, float *d_output, float *infovector, float lambdaPar, int iter, float epsil,
                   int methodTV, int nonneg, int dimX, int dimY, int dimZ, const cilreg_opts *opts, void *stream);
CILREG_API int cilreg_pd_dev(const float *d_input, float *d_output, float *infovector, float lambdaPar, int iter, float epsil,
                  float lipschitz_const, int methodTV, int nonneg, int dimX, int dimY, int dimZ,
                  const cilreg_opts *opts, void *stream);
CILREG_API int cilreg_sb_dev(const float *d_input, float *d_output, float *infovector, float mu, int iter, float epsil,
                  int methodTV, int dimX, int dimY, int dimZ, const cilreg_opts *opts, void *stream);
CILREG_API int cilreg_tgv_dev(const float *d_input, float *d_output, float *infovector, float lambda, float alpha1,
                   float alpha0, int iter, float L2, float epsil, int dimX, int dimY, int dimZ,
                   const cilreg_opts *opts, void *stream);

/* FP64 solves (compute and storage in double; dual_precision is ignored). */
CILREG_API int cilreg_fgp_dev_f64(const double *d_input, double *d_output, double *infovector, double lambdaPar, int iter,
                       double epsil, int methodTV, int nonneg, int dimX, int dimY, int dimZ,
                       const cilreg_opts *opts, void *stream);
CILREG_API int cilreg_tgv_dev_f64(const double *d_input, double *d_output, double *infovector, double lambda, double alpha1,
                       double alpha0, int iter, double L2, double epsil, int dimX, int dimY, int dimZ,
                       const cilreg_opts *opts, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* CILREG_EXT_H */
