// solvers.hpp -- device-resident solver entry points shared by the C ABI, the
// extension API, the batch launcher and the serving loop.
//
// Every solver takes device pointers and a stream, runs `iter` iterations of
// its method with the reference's stopping rule, writes the reference's
// output into d_out and reports the iteration count and last residual.  FGP,
// PD and TGV (T = float) optionally start from and return their iterate state
// (WarmState): the state after the executed iterations -- all `iter`, or up
// to and including the iteration at which the stopping rule fired (one
// iteration per pass then).
#pragma once
#include "runtime.hpp"
#include "state.hpp"

namespace cilreg {

template <typename T, typename S>
int rof_solve(const T* d_f, T* d_out, T lambda, int iter, T tau, T epsil, int nx, int ny, int nz, const Opts& opts,
              hipStream_t s, int* iters_done, float* re);
template <typename T, typename S>
int fgp_solve(const T* d_f, T* d_out, T lambda, int iter, T epsil, int method, int nonneg, int nx, int ny, int nz,
              const Opts& opts, hipStream_t s, int* iters_done, float* re, const WarmState* ws = nullptr);
template <typename T, typename S>
int pd_solve(const T* d_f, T* d_out, T lambda, int iter, T epsil, T lipschitz, int method, int nonneg, int nx, int ny,
             int nz, const Opts& opts, hipStream_t s, int* iters_done, float* re, const WarmState* ws = nullptr);
template <typename T, typename S>
int sb_solve(const T* d_f, T* d_out, T mu, int iter, T epsil, int method, int nx, int ny, int nz, const Opts& opts,
             hipStream_t s, int* iters_done, float* re, const WarmState* ws = nullptr);
template <typename T, typename S>
int tgv_solve(const T* d_f, T* d_out, T lambda, T alpha1, T alpha0, int iter, T L2, T epsil, int nx, int ny, int nz,
              const Opts& opts, hipStream_t s, int* iters_done, float* re, const WarmState* ws = nullptr);
int llt_solve(const float* d_f, float* d_out, float lambda_rof, float lambda_llt, int iter, float tau, float epsil,
              int nx, int ny, int nz, const Opts& opts, hipStream_t s, int* iters_done, float* re);
int ndf_solve(const float* d_f, float* d_out, float lambda, float sigma, int iter, float tau, int penalty, float epsil,
              int nx, int ny, int nz, const Opts& opts, hipStream_t s, int* iters_done, float* re);
int d4_solve(const float* d_f, float* d_out, float lambda, float sigma, int iter, float tau, float epsil, int nx,
             int ny, int nz, const Opts& opts, hipStream_t s, int* iters_done, float* re);
template <typename T>
int dtv_solve(const T* d_f, const T* d_ref, T* d_out, T lambda, int iter, T epsil, T eta, int method, int nonneg,
              int nx, int ny, int nz, const Opts& opts, hipStream_t s, int* iters_done, float* re);

}  // namespace cilreg
