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
 * is a hipStream_t (NULL = the library's own stream, ordered after the work
 * already queued on the legacy default stream).  Every call returns 0 on
 * success and -1 on error, and is synchronous at return.
 */
#ifndef CILREG_EXT_H
#define CILREG_EXT_H

#include <stddef.h>

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
    int stream_launches; /* batch: 1 = launch directly instead of capturing one hipGraph */
    int temporal_depth;  /* iterations fused per pass where supported (FGP, PD, TGV 3D): 0 = auto (2), 1 = off */
    int autotune;        /* 1: tile_y / zchunk / temporal_depth left at 0 are chosen by cilreg_autotune */
    int persistent;      /* batch: 1 = one persistent work-queue kernel per method group; disabled in this
                            build (cilreg_batch_run returns -1 with the reason) */
} cilreg_opts;

/* Methods of a batch or served job. */
enum { CILREG_ROF = 0, CILREG_FGP = 1, CILREG_PD = 2, CILREG_TGV = 3 };

/* One volume of a batch or one request of the server.  Parameters not used
 * by the method are ignored.  The input is either a device array (d_input) or,
 * when d_input is NULL and phantom_shapes > 0, a seeded ellipsoid phantom
 * generated on the device (cilreg_phantom_dev). */
typedef struct cilreg_job {
    int method;                    /* CILREG_ROF / _FGP / _PD / _TGV */
    int dimX, dimY, dimZ;          /* dimX fastest; dimZ <= 1: 2D */
    int iter;                      /* iteration count */
    float epsil;                   /* tolerance, 0 = fixed iteration count */
    float lambda;                  /* regularisation parameter (all methods) */
    float tau;                     /* ROF marching step */
    float lipschitz;               /* PD Lipschitz constant (<= 0: 8, as in demos/demo_gpu_regularisers3D.py:197) */
    float alpha1, alpha0, L2;      /* TGV (L2 <= 0: 12, as in demos/demo_cpu_regularisers3D.py:338) */
    int methodTV;                  /* FGP, PD: 0 isotropic, 1 anisotropic */
    int nonneg;                    /* FGP, PD */
    const float *d_input;          /* device input, or NULL for a generated phantom */
    float *d_output;               /* device output (may be NULL for server jobs: see cilreg_server_poll) */
    int phantom_shapes;            /* generated input: number of ellipsoids */
    float phantom_sigma;           /* generated input: Gaussian noise sigma */
    unsigned long long phantom_seed;
    float infovector[2];           /* out: [iterations, last residual] */
} cilreg_job;

/* Last error recorded on this host thread ("" if none). */
CILREG_API const char *cilreg_last_error(void);

/* Per-iteration kernel time of the last solve on this host thread that ran
 * with opts.profile = 1: mean and minimum in milliseconds over `count`
 * iterations, measured with HIP events on the solve's stream around each
 * launch of the iteration kernel(s) (the stopping-rule finaliser excluded);
 * a launch that advances several iterations counts its time divided by them. */
CILREG_API int cilreg_last_profile(double *mean_ms, double *min_ms, int *count);

/* Frees the device memory the library keeps cached between calls on the
 * current device (solves reuse freed workspaces instead of calling hipMalloc;
 * set CILREG_NO_CACHE=1 to allocate and free per call like the reference). */
CILREG_API int cilreg_release_cache(void);

/* Device memory of the library's pool on the current device: bytes held by
 * running solves (busy) and kept cached for reuse (idle).  The drop-in
 * *_GPU_main entry points return their workspace to the device before they
 * return, as the reference's do (CILREG_KEEP_CACHE=1 keeps it cached instead),
 * so after a completed or failed drop-in call both are 0 unless other calls
 * of the process hold memory. */
CILREG_API int cilreg_pool_stats(size_t *busy_bytes, size_t *idle_bytes);

/* Test hook for failure handling: the n-th next device allocation the
 * library requests on this host thread fails as an exhausted device would
 * (n = 0 turns it off).  Every entry point then returns -1 with the
 * allocation error and releases what it had allocated. */
CILREG_API void cilreg_inject_alloc_failure(int n);

/* Device-resident solves.  d_input and d_output are device arrays of
 * dimX*dimY*dimZ floats (2D: dimZ = 1); infovector is a host float[2]
 * receiving [iterations, last residual] as in the reference (NULL skips it).
 * A NULL d_input or d_output returns -1 before any device work.  For FGP with
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
/* The explicit-diffusion family on device buffers (FP32). */
CILREG_API int cilreg_llt_dev(const float *d_input, float *d_output, float *infovector, float lambdaROF,
                   float lambdaLLT, int iter, float tau, float epsil, int dimX, int dimY, int dimZ,
                   const cilreg_opts *opts, void *stream);
CILREG_API int cilreg_ndf_dev(const float *d_input, float *d_output, float *infovector, float lambdaPar, float sigmaPar,
                   int iter, float tau, int penaltytype, float epsil, int dimX, int dimY, int dimZ,
                   const cilreg_opts *opts, void *stream);
CILREG_API int cilreg_d4_dev(const float *d_input, float *d_output, float *infovector, float lambdaPar, float sigmaPar,
                  int iter, float tau, float epsil, int dimX, int dimY, int dimZ, const cilreg_opts *opts,
                  void *stream);
/* FGP-dTV on device buffers (d_ref: the reference image, same dimensions);
 * FP32 storage only (dual_precision must be 0). */
CILREG_API int cilreg_dtv_dev(const float *d_input, const float *d_ref, float *d_output, float *infovector,
                   float lambdaPar, int iter, float epsil, float eta, int methodTV, int nonneg, int dimX, int dimY,
                   int dimZ, const cilreg_opts *opts, void *stream);

/* FGP-TV without any device allocation: cilreg_fgp_dev with the same
 * arguments, results and stopping rule (bit-identical to it), but the
 * solve's workspace (dual fields, objective buffers, control block, residual
 * partials) is carved from a fixed device array of
 * cilreg_noalloc_scratch_bytes() bytes (64 MiB) that each GPU holds from
 * library load until the process ends; no hipMalloc, hipFree or other
 * allocating call is made (opts.autotune and opts.profile are ignored, the
 * stop flag is not polled through pinned memory).  No counterpart in the
 * reference: for callers that cannot allocate device memory (code that must
 * not call hipMalloc, serving loops, harnesses that forbid it).  Volumes
 * whose workspace exceeds the array (FP32 3D: 56 B per voxel, about 1.19 M
 * voxels; 2D: 40 B per pixel; BF16/FP16 duals: 32 and 24 B) return -1 with
 * that reason.  Calls on one device share the array and are serialised by
 * the library; the array keeps the last call's contents. */
CILREG_API int cilreg_fgp_dev_noalloc(const float *d_input, float *d_output, float *infovector, float lambdaPar,
                                      int iter, float epsil, int methodTV, int nonneg, int dimX, int dimY, int dimZ,
                                      const cilreg_opts *opts, void *stream);
CILREG_API size_t cilreg_noalloc_scratch_bytes(void);

/* FP64 solves (compute and storage in double; dual_precision is ignored). */
CILREG_API int cilreg_fgp_dev_f64(const double *d_input, double *d_output, double *infovector, double lambdaPar, int iter,
                       double epsil, int methodTV, int nonneg, int dimX, int dimY, int dimZ,
                       const cilreg_opts *opts, void *stream);
CILREG_API int cilreg_pd_dev_f64(const double *d_input, double *d_output, double *infovector, double lambdaPar, int iter,
                      double epsil, double lipschitz_const, int methodTV, int nonneg, int dimX, int dimY, int dimZ,
                      const cilreg_opts *opts, void *stream);
CILREG_API int cilreg_tgv_dev_f64(const double *d_input, double *d_output, double *infovector, double lambda, double alpha1,
                       double alpha0, int iter, double L2, double epsil, int dimX, int dimY, int dimZ,
                       const cilreg_opts *opts, void *stream);

/* The ROF objective in FP64 (the "final objective" of a precision study):
 * TV_energy2D/3D of src/Core/regularisers_CPU/utils.c:64-117, i.e.
 * type 1: sum 2*lambda*|grad+ u| + sum (u - u0)^2, type 2: the gradient term
 * alone, forward differences with a replicate boundary; every term formed and
 * summed in double in a fixed order (deterministic), where the reference
 * sums serially in float.  d_U, d_U0: device arrays of dimX*dimY*dimZ values
 * (2D: dimZ = 1); *energy is a host double. */
CILREG_API int cilreg_tv_energy_dev(const float *d_U, const float *d_U0, double lambda, int type, int dimX, int dimY,
                                    int dimZ, double *energy, void *stream);
CILREG_API int cilreg_tv_energy_dev_f64(const double *d_U, const double *d_U0, double lambda, int type, int dimX,
                                        int dimY, int dimZ, double *energy, void *stream);

/* Self-check of the kernels' correctly rounded float square root on the
 * current device: compares it with sqrtf on every 32-bit pattern read as a
 * float and returns the number of differing results (0 expected) and of
 * values checked (2^32). */
CILREG_API int cilreg_check_sqrt_rn(unsigned long long *mismatches, unsigned long long *checked);

/* Self-check of the kernels' double square root of the 3D ROF / LLT-ROF flux
 * normaliser, (float)sqrt((double)s + eps), which skips sqrt's scaling and
 * special-value patches for in-range arguments: compares it with
 * (float)sqrt((double)s + eps) on the device for every non-negative float bit
 * pattern s (2^31 values, infinities and NaNs included) and returns the number
 * of differing results (0 expected) and of values checked.  No counterpart in
 * the reference. */
CILREG_API int cilreg_check_rof_sqrt_d(double eps, unsigned long long *mismatches, unsigned long long *checked);

/* Achievable HBM bandwidth of the current device, from streaming kernels
 * over arrays of `bytes` each (>= 1 MiB), the mean of `reps` timed launches:
 * gbps[0] copy (1 read + 1 write), gbps[1] read, gbps[2] write, gbps[3]
 * 2 reads + 1 write.  Bytes moved / kernel time, in GB/s. */
CILREG_API int cilreg_stream_bench(size_t bytes, int reps, double *gbps);

/* Warm start: the solver's iterate state in caller memory, so an outer loop
 * (FISTA, ADMM) can resume the inner solver.  d_state holds FP32 channels of
 * dimX*dimY*dimZ values each, whatever dual_precision is:
 *   FGP: P1, P2, P3, P1_prev, P2_prev, P3_prev (2D: channels 0, 1, 3, 4);
 *        t_state (host float[2]) = {t_k, multiplier of the next iteration};
 *   PD:  u_bar, P1, P2, P3 (2D: 0, 1, 2);
 *   TGV: u_bar, V1..V3, P1..P3, Q11, Q22, Q33, Q12, Q13, Q23
 *        (2D: u_bar, V1, V2, P1, P2 in channels 0, 1, 2, 4, 5 and Q11, Q22, Q12 in 7, 8, 9);
 *   SB:  U, D1..D3, B1..B3 (the split variables and Bregman terms of
 *        SB_TV_core.c:61-133; 2D: channels 0, 1, 2, 4, 5).
 * warm = 1 starts from the state (and t_state); warm = 0 starts cold like
 * the plain solve.  On return the state is that after the executed
 * iterations: all `iter`, or, with a tolerance, up to and including the
 * iteration at which the stopping rule fired.  N iterations followed by a
 * warm call of M equal one call of N + M bit for bit (fixed counts; the
 * stopping rule's count restarts per call).  iter >= 1. */
CILREG_API int cilreg_fgp_dev_ws(const float *d_input, float *d_output, float *infovector, float lambdaPar, int iter,
                                 float epsil, int methodTV, int nonneg, int dimX, int dimY, int dimZ, float *d_state,
                                 float *t_state, int warm, const cilreg_opts *opts, void *stream);
CILREG_API int cilreg_pd_dev_ws(const float *d_input, float *d_output, float *infovector, float lambdaPar, int iter,
                                float epsil, float lipschitz_const, int methodTV, int nonneg, int dimX, int dimY,
                                int dimZ, float *d_state, int warm, const cilreg_opts *opts, void *stream);
CILREG_API int cilreg_sb_dev_ws(const float *d_input, float *d_output, float *infovector, float mu, int iter,
                                float epsil, int methodTV, int dimX, int dimY, int dimZ, float *d_state, int warm,
                                const cilreg_opts *opts, void *stream);
CILREG_API int cilreg_tgv_dev_ws(const float *d_input, float *d_output, float *infovector, float lambda, float alpha1,
                                 float alpha0, int iter, float L2, float epsil, int dimX, int dimY, int dimZ,
                                 float *d_state, int warm, const cilreg_opts *opts, void *stream);

/* The autotuner: times short fixed-count solves of a synthetic volume of the
 * given extent for each candidate tiling (tile_y 8/4, zchunk planned or
 * 8/16/32) and temporal depth (1/2 for FGP, PD, TGV in 3D) with base's dual
 * storage, and writes the fastest into *best (base's other fields copied);
 * *ms_per_iter receives its kernel time per iteration.  Cached per process
 * and size class.  method: CILREG_ROF / _FGP / _PD / _TGV; methodTV as for
 * FGP / PD.  The choice never changes results (tiling and depth are
 * bit-neutral). */
CILREG_API int cilreg_autotune(int method, int dimX, int dimY, int dimZ, int methodTV, const cilreg_opts *base,
                               cilreg_opts *best, double *ms_per_iter);

/* The autotuner's choices as a text file (no counterpart in the reference,
 * whose tile sizes are compile-time constants).  First line
 * "# cilreg autotune cache v1", then one entry per line:
 *   device method precision iso grouped nx ny nz tile_y zchunk depth ms_per_iter
 * device = architecture and CU count ("gfx950:256"); an entry is used only on
 * such a device.  cilreg_tune_save writes every cached choice (through
 * path.tmp and a rename) and returns the number written; cilreg_tune_load
 * merges a file into the process's cache (replacing entries with the same
 * key) and returns the number read -- a file with a malformed line is
 * rejected whole (-1, cilreg_last_error names the line).  With the
 * environment variable CILREG_TUNE_CACHE=<path> the file is loaded before the
 * process's first tuning (a missing file is an empty cache) and rewritten
 * after every new choice.  These calls touch no device. */
CILREG_API int cilreg_tune_save(const char *path);
CILREG_API int cilreg_tune_load(const char *path);
CILREG_API int cilreg_tune_entries(void);  /* choices cached in this process */
CILREG_API void cilreg_tune_clear(void);   /* forget them (files are untouched) */

/* Runs every job of the batch to completion with grouped kernels: one launch
 * per iteration level and method group covers all volumes of that group, so
 * many small or uneven volumes fill the device together.  Unless
 * opts->stream_launches is set the whole level sequence is captured into one
 * hipGraph (one parallel branch per method group) and launched once.
 * Each volume's result is bit-identical to solving it alone. */
CILREG_API int cilreg_batch_run(cilreg_job *jobs, int njobs, const cilreg_opts *opts, void *stream);

/* Seeded synthetic input on the device: n_shapes random axis-aligned
 * ellipsoids (3D) or rectangles/ellipses (2D), max-combined, plus N(0,
 * sigma^2) noise from a counter-based generator (deterministic for a seed). */
CILREG_API int cilreg_phantom_dev(float *d_out, int dimX, int dimY, int dimZ, int n_shapes, float sigma,
                                  unsigned long long seed, void *stream);

/* Continuous-batching server.  Jobs are admitted while their device memory
 * fits the HBM budget; all admitted jobs advance together through grouped
 * kernels (jobs at different iteration counts share the same launches) and
 * their state stays on the device between steps.  cilreg_server_step queues
 * `levels` more iterations of every admitted job behind the step already
 * queued, then waits only for that earlier step and retires the jobs it
 * finished: the device always has the newest step to run while the caller
 * admits, polls and calls again (a job is reported done one step after its
 * last iteration was queued).  levels <= 0 waits for all queued work and
 * retires what it finished.  A job's result is bit-identical to solving it
 * alone. */
typedef struct cilreg_server cilreg_server;
CILREG_API cilreg_server *cilreg_server_create(int gpu_device, double hbm_budget_bytes, const cilreg_opts *opts);
CILREG_API int cilreg_server_submit(cilreg_server *srv, const cilreg_job *job, long long *ticket);
CILREG_API int cilreg_server_step(cilreg_server *srv, int levels, int *n_active, int *n_pending);
/* 1 when done (infovector filled; the output copied to d_dst if not NULL, then
 * the job's memory is released), 0 when pending or running, -1 if unknown. */
CILREG_API int cilreg_server_poll(cilreg_server *srv, long long ticket, float *infovector, float *d_dst);
/* 0 waiting for admission, 1 running, 2 done (collect with cilreg_server_poll),
 * -1 failed (e.g. larger than the budget; poll reports the error) or unknown. */
CILREG_API int cilreg_server_status(cilreg_server *srv, long long ticket);
/* Withdraws a job in any state (waiting, running, or done and not yet
 * polled): it is removed from the server and its device memory released;
 * the other jobs are unaffected.  0 on success, -1 for an unknown ticket. */
CILREG_API int cilreg_server_cancel(cilreg_server *srv, long long ticket);
CILREG_API void cilreg_server_destroy(cilreg_server *srv);

#ifdef __cplusplus
}
#endif

#endif /* CILREG_EXT_H */
