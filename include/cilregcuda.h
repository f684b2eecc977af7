/*
 * cilregcuda.h -- drop-in C ABI of libcilregcuda.so, the MI355X build of the
 * toolkit's GPU regularisers.
 *
 * Each entry point keeps the reference's signature, argument meaning and
 * calling convention: host pointers owned by the caller, dims with dimX the
 * fastest axis, synchronous return, 0 on success and -1 on a device error
 * with a message on stderr.  The library allocates and frees all device
 * memory inside each call.
 *
 * The per-method headers of the reference (TV_ROF_GPU_core.h, ...) are also
 * provided in this directory and include this file.
 */
#ifndef CILREGCUDA_H
#define CILREGCUDA_H

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

/* ROF-TV, explicit scheme.  Replaces
 * src/Core/regularisers_GPU/headers/TV_ROF_GPU_core.h:6.
 * 2D when Z == 1 (Z == 0 is read as 1), 3D otherwise. */
CILREG_API int TV_ROF_GPU_main(float *Input, float *Output, float *infovector, float lambdaPar, int iter, float tau,
                    float epsil, int gpu_device, int N, int M, int Z);

/* FGP-TV.  Replaces src/Core/regularisers_GPU/TV_FGP_GPU_core.h:7.
 * 2D when dimZ <= 1.  With epsil != 0 the first residual compares against the
 * caller's Output contents, as the C-OMP solver does. */
CILREG_API int TV_FGP_GPU_main(float *Input, float *Output, float *infovector, float lambdaPar, int iter, float epsil,
                    int methodTV, int nonneg, int gpu_device, int dimX, int dimY, int dimZ);

/* PD-TV (Chambolle-Pock).  Replaces
 * src/Core/regularisers_GPU/headers/TV_PD_GPU_core.h:7.  2D when dimZ <= 1. */
CILREG_API int TV_PD_GPU_main(float *Input, float *Output, float *infovector, float lambdaPar, int iter, float epsil,
                   float lipschitz_const, int methodTV, int nonneg, int gpu_device, int dimX, int dimY, int dimZ);
/* Alias of TV_PD_GPU_main under the name the Python layer uses (TV.py:567). */
CILREG_API int PDTV_GPU_main(float *Input, float *Output, float *infovector, float lambdaPar, int iter, float epsil,
                  float lipschitz_const, int methodTV, int nonneg, int gpu_device, int dimX, int dimY, int dimZ);

/* SB-TV (split Bregman).  Replaces src/Core/regularisers_GPU/TV_SB_GPU_core.h:8.
 * 2D when dimZ <= 1. */
CILREG_API int TV_SB_GPU_main(float *Input, float *Output, float *infovector, float mu, int iter, float epsil, int methodTV,
                   int gpu_device, int dimX, int dimY, int dimZ);
/* Alias of TV_SB_GPU_main under the name the Python layer uses (TV.py:609). */
CILREG_API int SB_TV_GPU_main(float *Input, float *Output, float *infovector, float mu, int iter, float epsil, int methodTV,
                   int gpu_device, int dimX, int dimY, int dimZ);

/* TGV^2-L2 primal-dual.  Replaces src/Core/regularisers_GPU/TGV_GPU_core.h:8.
 * 2D when dimZ == 1.  dimX is the fastest axis, as in every other entry point
 * (the reference CUDA kernels index the first two axes the other way round,
 * TGV_GPU_core.cu:290 vs :515; this library follows the header). */
CILREG_API int TGV_GPU_main(float *U0, float *U, float *infovector, float lambda, float alpha1, float alpha0,
                 int iterationsNumb, float L2, float epsil, int gpu_device, int dimX, int dimY, int dimZ);

/* FGP-dTV: FGP-TV guided by the normalised gradient of a reference image
 * InputRef of the same dimensions (eta smooths that gradient's magnitude).
 * Replaces src/Core/regularisers_GPU/dTV_FGP_GPU_core.h:7.  2D when
 * dimZ <= 1.  With epsil != 0 the first residual compares against the
 * caller's Output contents, as the C-OMP solver does; Output is left
 * untouched when iter <= 0. */
CILREG_API int dTV_FGP_GPU_main(float *Input, float *InputRef, float *Output, float *infovector, float lambdaPar,
                     int iter, float epsil, float eta, int methodTV, int nonneg, int gpu_device, int dimX, int dimY,
                     int dimZ);

/* LLT-ROF: ROF-TV plus the Lysaker-Lundervold-Tai second-order term,
 * explicit scheme.  Replaces src/Core/regularisers_GPU/LLT_ROF_GPU_core.h:6.
 * 2D when Z == 1 (Z == 0 is read as 1), 3D otherwise. */
CILREG_API int LLT_ROF_GPU_main(float *Input, float *Output, float *infovector, float lambdaROF, float lambdaLLT,
                     int iterationsNumb, float tau, float epsil, int gpu_device, int N, int M, int Z);

/* NDF: linear (sigmaPar == 0) or nonlinear diffusion, penaltytype 1 Huber,
 * 2 Perona-Malik, 3 Tukey biweight, 4 threshold-constrained linear,
 * 5 threshold-constrained Huber.  Replaces
 * src/Core/regularisers_GPU/NonlDiff_GPU_core.h:6.  Each step is the
 * synchronous explicit update (every neighbour from the previous iterate);
 * an unknown penalty type returns -1. */
CILREG_API int NonlDiff_GPU_main(float *Input, float *Output, float *infovector, float lambdaPar, float sigmaPar,
                      int iterationsNumb, float tau, int penaltytype, float epsil, int gpu_device, int N, int M,
                      int Z);

/* Diff4th: anisotropic fourth-order diffusion, explicit scheme.  Replaces
 * src/Core/regularisers_GPU/Diffus_4thO_GPU_core.h:6. */
CILREG_API int Diffus4th_GPU_main(float *Input, float *Output, float *infovector, float lambdaPar, float sigmaPar,
                       int iterationsNumb, float tau, float epsil, int gpu_device, int N, int M, int Z);

#ifdef __cplusplus
}
#endif

#endif /* CILREGCUDA_H */
