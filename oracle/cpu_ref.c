/*
 * cpu_ref.c -- CPU oracle for the iterative TV/TGV regularisers.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke
 * entry point and the CPU-baseline leg of bench.py load it through
 * oracle/oracle.py.  Nothing in the product library (libcilregcuda.so) links,
 * loads or calls it, and no product code path falls back to it.
 *
 * It restates the reference toolkit's C-OMP solvers operation by operation so
 * that, compiled with the reference's flags (-O3 -funsigned-char -fopenmp, no
 * -march, hence no FMA contraction on x86-64), it reproduces the reference
 * output bit for bit.  Where the reference promotes to double (an EPS that is
 * a double literal, a double sqrt, a 1.0/x), the restatement promotes at the
 * same place.  Pinning: tests/test_oracle_golden.py compares every entry
 * point bit-exactly against fixtures written by the reference library built
 * from its own sources (oracle/Makefile -> oracle/_ref/libcilreg.so), and
 * tests/test_oracle_invariants.py checks the closed forms the mathematics
 * fixes (constant-image fixpoints, exact scale equivariance, the TGV singleton
 * axis reduction, dual feasibility).
 *
 * Compiled twice: REAL=float gives the bit-exact FP32 oracle (symbols
 * cr_*_f32), REAL=double gives the FP64 oracle of the precision study
 * (symbols cr_*_f64).  In the FP64 build every promotion is a no-op.
 *
 * Memory layout (as the reference): index = (dimX*dimY)*k + j*dimX + i.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef REAL
#define REAL float
#endif

#ifdef CR_DOUBLE
#define SQRT_R(x) sqrt(x)
#define FN(name) name##_f64
#else
#define SQRT_R(x) sqrtf(x)
#define FN(name) name##_f32
#endif

#define ROF_EPS 1.0e-8 /* a double literal: ROF_TV_core.c:22 */

typedef long idx_t;

/* ------------------------------------------------------------------------ */
/* shared pieces                                                             */
/* ------------------------------------------------------------------------ */

static void copy_field(const REAL *src, REAL *dst, idx_t n)
{
    idx_t t;
#pragma omp parallel for
    for (t = 0; t < n; t++) dst[t] = src[t];
}

/* Stopping-rule residual, restating the serial loops at ROF_TV_core.c:80-86
 * (identical at FGP_TV_core.c:143-149, PD_TV_core.c:132-138,
 * SB_TV_core.c:124-130, TGV_core.c:170-176): a serial sum in the working
 * precision, in index order.  powf(x,2) is folded to x*x by gcc. */
static REAL rel_change(const REAL *u, const REAL *u_prev, idx_t n)
{
    REAL s_d = 0, s_u = 0;
    idx_t t;
    for (t = 0; t < n; t++) {
        REAL d = u[t] - u_prev[t];
        s_d += d * d;
        s_u += u[t] * u[t];
    }
    return SQRT_R(s_d) / SQRT_R(s_u);
}

/* Stopping bookkeeping shared by every solver: count is cumulative
 * (not consecutive) and the loop breaks once it exceeds 3. */
typedef struct { int count; REAL re; } stop_state;

/* Test hook: with forced_stop >= 0 the rule fires at the check of iteration
 * forced_stop and at no other (the residual and count are still computed).
 * Parity tests use it to obtain the output of a solve that stopped at a given
 * iteration, whatever the residual sums decide.  -1 (the default) is the
 * reference's rule. */
static int forced_stop = -1;

void FN(cr_set_forced_stop)(int it) { forced_stop = it; }

static int stop_check(stop_state *st, const REAL *u, const REAL *u_prev, idx_t n, REAL epsil, int it)
{
    st->re = rel_change(u, u_prev, n);
    if (st->re < epsil) st->count++;
    if (forced_stop >= 0) return it == forced_stop;
    return st->count > 3;
}

/* Unit-ball projection of the dual field, utils.c:149-213.
 * iso: scale by 1/sqrt(|p|^2) when |p|^2 > 1; aniso: p_d / max(|p_d|, 1). */
static void project_unit(REAL *p1, REAL *p2, REAL *p3, int ndim, int aniso, idx_t n)
{
    idx_t t;
#pragma omp parallel for
    for (t = 0; t < n; t++) {
        if (!aniso) {
            REAL nrm2 = p1[t] * p1[t] + p2[t] * p2[t];
            if (ndim == 3) nrm2 = nrm2 + p3[t] * p3[t];
            if (nrm2 > 1) {
                REAL s = (REAL)1 / SQRT_R(nrm2);
                p1[t] = p1[t] * s;
                p2[t] = p2[t] * s;
                if (ndim == 3) p3[t] = p3[t] * s;
            }
        } else {
            REAL a1 = (REAL)fabs(p1[t]), a2 = (REAL)fabs(p2[t]);
            if (a1 < 1) a1 = 1;
            if (a2 < 1) a2 = 1;
            p1[t] = p1[t] / a1;
            p2[t] = p2[t] / a2;
            if (ndim == 3) {
                REAL a3 = (REAL)fabs(p3[t]);
                if (a3 < 1) a3 = 1;
                p3[t] = p3[t] / a3;
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* ROF-TV: explicit gradient descent on the smoothed TV functional.          */
/* ROF_TV_core.c:50-317                                                      */
/* ------------------------------------------------------------------------ */

static int sgn_of(REAL x) { return (x > 0) - (x < 0); }

/* Squared minmod of two one-sided differences (ROF_TV_core.c:132-133):
 * fabs promotes to double, the half-sum of signs is a float. */
static REAL minmod_sq(REAL a, REAL b)
{
    double fa = fabs(a), fb = fabs(b);
    REAL m = (REAL)((REAL)0.5 * (sgn_of(a) + sgn_of(b)) * (fa < fb ? fa : fb));
    return m * m;
}

/* Normalised forward difference along one axis (ROF_TV_core.c:102-266).
 * 'fwd' is the axis of the forward difference, the other two axes enter
 * through minmod.  Mirror boundary: +1 past the end maps to -1 and -1 before
 * the start maps to +1.  Only the 3D D1 takes a double sqrt of the double
 * sum (ROF_TV_core.c:136); every other variant rounds the double sum to
 * float and calls sqrtf (:160,199,223,262). */
/* Mirrored neighbours: +1, or -1 at the last index; -1, or +1 at index 0.
 * An axis of extent 1 has no mirror image (the reference reads index -1
 * there); it maps to the point itself. */
static idx_t mir_p(idx_t i, idx_t n) { return i + 1 < n ? i + 1 : (i - 1 >= 0 ? i - 1 : 0); }
static idx_t mir_m(idx_t i, idx_t n) { return i - 1 >= 0 ? i - 1 : (i + 1 < n ? i + 1 : 0); }

static void rof_dir(const REAL *u, REAL *D, idx_t nx, idx_t ny, idx_t nz, int fwd, int ndim, double eps)
{
    idx_t k;
    idx_t sxy = nx * ny;
#pragma omp parallel for
    for (k = 0; k < nz; k++) {
        idx_t j, i;
        idx_t kp = mir_p(k, nz);
        idx_t km = mir_m(k, nz);
        for (j = 0; j < ny; j++) {
            idx_t jp = mir_p(j, ny);
            idx_t jm = mir_m(j, ny);
            for (i = 0; i < nx; i++) {
                idx_t ip = mir_p(i, nx);
                idx_t im = mir_m(i, nx);
                idx_t c = sxy * k + nx * j + i;
                REAL uc = u[c];
                /* forward / backward one-sided differences per axis */
                REAL fj = u[sxy * k + nx * jp + i] - uc, bj = uc - u[sxy * k + nx * jm + i];
                REAL fi = u[sxy * k + nx * j + ip] - uc, bi = uc - u[sxy * k + nx * j + im];
                REAL num, s;
                double sd;
                if (fwd == 0) {          /* D1: forward along j */
                    num = fj;
                    s = num * num + minmod_sq(fi, bi);
                } else if (fwd == 1) {   /* D2: forward along i */
                    num = fi;
                    s = num * num + minmod_sq(fj, bj);
                } else {                 /* D3: forward along k */
                    num = u[sxy * kp + nx * j + i] - uc;
                    s = num * num + minmod_sq(fj, bj);
                    s = s + minmod_sq(fi, bi);
                }
                if (ndim == 3 && fwd != 2) {
                    REAL fk = u[sxy * kp + nx * j + i] - uc, bk = uc - u[sxy * km + nx * j + i];
                    s = s + minmod_sq(fk, bk);
                }
                sd = (double)s + eps;
                if (ndim == 3 && fwd == 0)
                    D[c] = num / (REAL)sqrt(sd);
                else
                    D[c] = num / SQRT_R((REAL)sd);
            }
        }
    }
}

/* u += tau*(lambda*div D - (u - f)), backward differences with mirror
 * boundary (ROF_TV_core.c:269-317). */
static void rof_update(const REAL *D1, const REAL *D2, const REAL *D3, REAL *u, const REAL *f,
                       const REAL *lam, int lam_is_arr, REAL tau, idx_t nx, idx_t ny, idx_t nz, int ndim)
{
    idx_t k;
    idx_t sxy = nx * ny;
#pragma omp parallel for
    for (k = 0; k < nz; k++) {
        idx_t j, i;
        idx_t km = mir_m(k, nz);
        for (j = 0; j < ny; j++) {
            idx_t jm = mir_m(j, ny);
            for (i = 0; i < nx; i++) {
                idx_t im = mir_m(i, nx);
                idx_t c = sxy * k + nx * j + i;
                REAL lv = lam[c * lam_is_arr];
                REAL dv = (D1[c] - D1[sxy * k + nx * jm + i]) + (D2[c] - D2[sxy * k + nx * j + im]);
                if (ndim == 3) dv = dv + (D3[c] - D3[sxy * km + nx * j + i]);
                u[c] += tau * (lv * dv - (u[c] - f[c]));
            }
        }
    }
}

REAL FN(cr_rof)(const REAL *Input, REAL *Output, REAL *infovector, const REAL *lambdaPar, int lambda_is_arr,
                int iterationsNumb, REAL tau, REAL epsil, int dimX, int dimY, int dimZ)
{
    idx_t n = (idx_t)dimX * dimY * dimZ;
    int ndim = dimZ > 1 ? 3 : 2;
    REAL *D1 = calloc(n ? n : 1, sizeof(REAL)), *D2 = calloc(n ? n : 1, sizeof(REAL));
    REAL *D3 = calloc(n ? n : 1, sizeof(REAL)), *prev = NULL;
    stop_state st = {0, 0};
    int it;

    copy_field(Input, Output, n);
    if (epsil != 0) prev = calloc(n ? n : 1, sizeof(REAL));
    for (it = 0; it < iterationsNumb; it++) {
        int check = (epsil != 0) && (it % 5 == 0);
        if (check) copy_field(Output, prev, n);
        rof_dir(Output, D1, dimX, dimY, dimZ, 0, ndim, ROF_EPS);
        rof_dir(Output, D2, dimX, dimY, dimZ, 1, ndim, ROF_EPS);
        if (ndim == 3) rof_dir(Output, D3, dimX, dimY, dimZ, 2, ndim, ROF_EPS);
        rof_update(D1, D2, D3, Output, Input, lambdaPar, lambda_is_arr, tau, dimX, dimY, dimZ, ndim);
        if (check && stop_check(&st, Output, prev, n, epsil, it)) break;
    }
    free(D1); free(D2); free(D3); free(prev);
    infovector[0] = (REAL)it;
    infovector[1] = st.re;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* FGP-TV: FISTA on the dual of the ROF problem (Beck-Teboulle).             */
/* FGP_TV_core.c:40-269                                                      */
/* ------------------------------------------------------------------------ */

/* D = f - lambda*div R with R_d(x - e_d) = 0 before the start
 * (FGP_TV_core.c:173-187, :220-236). */
static void fgp_objective(const REAL *f, REAL *D, const REAL *R1, const REAL *R2, const REAL *R3,
                          REAL lambda, idx_t nx, idx_t ny, idx_t nz, int ndim)
{
    idx_t k;
    idx_t sxy = nx * ny;
#pragma omp parallel for
    for (k = 0; k < nz; k++) {
        idx_t j, i;
        for (j = 0; j < ny; j++)
            for (i = 0; i < nx; i++) {
                idx_t c = sxy * k + nx * j + i;
                REAL b1 = i ? R1[c - 1] : 0, b2 = j ? R2[c - nx] : 0;
                if (ndim == 3) {
                    REAL b3 = k ? R3[c - sxy] : 0;
                    D[c] = f[c] - lambda * (R1[c] + R2[c] + R3[c] - b1 - b2 - b3);
                } else {
                    D[c] = f[c] - lambda * (R1[c] + R2[c] - b1 - b2);
                }
            }
    }
}

/* P = R + step * (D(x) - D(x + e_d)), zero at the last index
 * (FGP_TV_core.c:188-204, :237-256). */
static void fgp_gradient(REAL *P1, REAL *P2, REAL *P3, const REAL *D, const REAL *R1, const REAL *R2,
                         const REAL *R3, REAL step, idx_t nx, idx_t ny, idx_t nz, int ndim)
{
    idx_t k;
    idx_t sxy = nx * ny;
#pragma omp parallel for
    for (k = 0; k < nz; k++) {
        idx_t j, i;
        for (j = 0; j < ny; j++)
            for (i = 0; i < nx; i++) {
                idx_t c = sxy * k + nx * j + i;
                REAL g1 = (i == nx - 1) ? 0 : D[c] - D[c + 1];
                REAL g2 = (j == ny - 1) ? 0 : D[c] - D[c + nx];
                P1[c] = R1[c] + step * g1;
                P2[c] = R2[c] + step * g2;
                if (ndim == 3) {
                    REAL g3 = (k == nz - 1) ? 0 : D[c] - D[c + sxy];
                    P3[c] = R3[c] + step * g3;
                }
            }
    }
}

REAL FN(cr_fgp)(const REAL *Input, REAL *Output, REAL *infovector, REAL lambdaPar, int iterationsNumb,
                REAL epsil, int methodTV, int nonneg, int dimX, int dimY, int dimZ)
{
    int ndim = dimZ <= 1 ? 2 : 3;
    idx_t nz = ndim == 2 ? 1 : dimZ;
    idx_t n = (idx_t)dimX * dimY * nz, t;
    idx_t alloc = n ? n : 1;
    REAL *prev = NULL;
    REAL *P[3], *Pold[3], *R[3];
    REAL tk = 1, tkp1 = 1;
    REAL step = (REAL)1 / ((ndim == 3 ? (REAL)12 : (REAL)8) * lambdaPar);
    stop_state st = {0, 0};
    int it, d;

    for (d = 0; d < 3; d++) {
        P[d] = calloc(alloc, sizeof(REAL));
        Pold[d] = calloc(alloc, sizeof(REAL));
        R[d] = calloc(alloc, sizeof(REAL));
    }
    if (epsil != 0) prev = calloc(alloc, sizeof(REAL));

    for (it = 0; it < iterationsNumb; it++) {
        int check = (epsil != 0) && (it % 5 == 0);
        REAL m;
        if (check) copy_field(Output, prev, n);
        fgp_objective(Input, Output, R[0], R[1], R[2], lambdaPar, dimX, dimY, nz, ndim);
        if (nonneg == 1)
            for (t = 0; t < n; t++) if (Output[t] < 0) Output[t] = 0;
        fgp_gradient(P[0], P[1], P[2], Output, R[0], R[1], R[2], step, dimX, dimY, nz, ndim);
        project_unit(P[0], P[1], P[2], ndim, methodTV != 0, n);
        tkp1 = ((REAL)1 + SQRT_R((REAL)1 + (REAL)4 * tk * tk)) * (REAL)0.5;
        m = (tk - (REAL)1) / tkp1;
#pragma omp parallel for private(d)
        for (t = 0; t < n; t++)
            for (d = 0; d < ndim; d++) R[d][t] = P[d][t] + m * (P[d][t] - Pold[d][t]);
        /* The 2D branch stores P_old before the residual check, the 3D branch
         * after it (FGP_TV_core.c:84-98 vs :142-159); neither order changes
         * the output, since a break returns D. */
        if (check && stop_check(&st, Output, prev, n, epsil, it)) break;
        for (d = 0; d < ndim; d++) copy_field(P[d], Pold[d], n);
        tk = tkp1;
    }
    for (d = 0; d < 3; d++) { free(P[d]); free(Pold[d]); free(R[d]); }
    free(prev);
    infovector[0] = (REAL)it;
    infovector[1] = st.re;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* FGP-dTV: FGP whose dual field and objective gradient are projected onto  */
/* the orthogonal complement of the reference image's eta-normalised        */
/* gradient (directional TV).  FGP_dTV_core.c:46-380                         */
/* ------------------------------------------------------------------------ */

/* B = grad(ref) / sqrt(|grad(ref)|^2 + eta^2), forward differences that take
 * the value past the last index as 0 (so -ref there, not 0); the sum is in the
 * working precision, its square root in double (FGP_dTV_core.c:188-208,
 * :279-306). */
static void dtv_refgrad(const REAL *ref, REAL *B1, REAL *B2, REAL *B3, REAL eta, idx_t nx, idx_t ny, idx_t nz,
                        int ndim)
{
    REAL eta_sq = eta * eta;
    idx_t k;
    idx_t sxy = nx * ny;
#pragma omp parallel for
    for (k = 0; k < nz; k++) {
        idx_t j, i;
        for (j = 0; j < ny; j++)
            for (i = 0; i < nx; i++) {
                idx_t c = sxy * k + nx * j + i;
                REAL v1 = (i == nx - 1) ? 0 : ref[c + 1];
                REAL v2 = (j == ny - 1) ? 0 : ref[c + nx];
                REAL g1 = v1 - ref[c], g2 = v2 - ref[c], g3 = 0;
                REAL m = g1 * g1 + g2 * g2;
                if (ndim == 3) {
                    REAL v3 = (k == nz - 1) ? 0 : ref[c + sxy];
                    g3 = v3 - ref[c];
                    m = m + g3 * g3;
                }
                m = (REAL)sqrt((double)(m + eta_sq));
                B1[c] = g1 / m;
                B2[c] = g2 / m;
                if (ndim == 3) B3[c] = g3 / m;
            }
    }
}

/* v <- v - <v, B> B pointwise (ProjectVect, FGP_dTV_core.c:210-223, :308-323). */
static void dtv_project_dual(REAL *R1, REAL *R2, REAL *R3, const REAL *B1, const REAL *B2, const REAL *B3, int ndim,
                             idx_t n)
{
    idx_t t;
#pragma omp parallel for
    for (t = 0; t < n; t++) {
        REAL ip = R1[t] * B1[t] + R2[t] * B2[t];
        if (ndim == 3) ip = ip + R3[t] * B3[t];
        R1[t] = R1[t] - ip * B1[t];
        R2[t] = R2[t] - ip * B2[t];
        if (ndim == 3) R3[t] = R3[t] - ip * B3[t];
    }
}

/* P = R + step * (g - <g, B> B) with g_d = D(x) - D(x + e_d), 0 at the last
 * index (Grad_dfunc, FGP_dTV_core.c:240-262, :342-367). */
static void dtv_gradient(REAL *P1, REAL *P2, REAL *P3, const REAL *D, const REAL *R1, const REAL *R2, const REAL *R3,
                         const REAL *B1, const REAL *B2, const REAL *B3, REAL step, idx_t nx, idx_t ny, idx_t nz,
                         int ndim)
{
    idx_t k;
    idx_t sxy = nx * ny;
#pragma omp parallel for
    for (k = 0; k < nz; k++) {
        idx_t j, i;
        for (j = 0; j < ny; j++)
            for (i = 0; i < nx; i++) {
                idx_t c = sxy * k + nx * j + i;
                REAL g1 = (i == nx - 1) ? 0 : D[c] - D[c + 1];
                REAL g2 = (j == ny - 1) ? 0 : D[c] - D[c + nx];
                REAL g3 = 0, ip;
                if (ndim == 3) g3 = (k == nz - 1) ? 0 : D[c] - D[c + sxy];
                ip = g1 * B1[c] + g2 * B2[c];
                if (ndim == 3) ip = ip + g3 * B3[c];
                g1 = g1 - ip * B1[c];
                g2 = g2 - ip * B2[c];
                P1[c] = R1[c] + step * g1;
                P2[c] = R2[c] + step * g2;
                if (ndim == 3) {
                    g3 = g3 - ip * B3[c];
                    P3[c] = R3[c] + step * g3;
                }
            }
    }
}

REAL FN(cr_dtv)(const REAL *Input, const REAL *InputRef, REAL *Output, REAL *infovector, REAL lambdaPar,
                int iterationsNumb, REAL epsil, REAL eta, int methodTV, int nonneg, int dimX, int dimY, int dimZ)
{
    int ndim = dimZ <= 1 ? 2 : 3;
    idx_t nz = ndim == 2 ? 1 : dimZ;
    idx_t n = (idx_t)dimX * dimY * nz, t;
    idx_t alloc = n ? n : 1;
    REAL *prev = NULL;
    REAL *P[3], *Pold[3], *R[3], *B[3];
    REAL tk = 1, tkp1 = 1;
    /* 1/(8 lambda) in 2D, 1/(26 lambda) in 3D (FGP_dTV_core.c:244, :346) */
    REAL step = (REAL)1 / ((ndim == 3 ? (REAL)26 : (REAL)8) * lambdaPar);
    stop_state st = {0, 0};
    int it, d;

    for (d = 0; d < 3; d++) {
        P[d] = calloc(alloc, sizeof(REAL));
        Pold[d] = calloc(alloc, sizeof(REAL));
        R[d] = calloc(alloc, sizeof(REAL));
        B[d] = calloc(alloc, sizeof(REAL));
    }
    if (epsil != 0) prev = calloc(alloc, sizeof(REAL));
    dtv_refgrad(InputRef, B[0], B[1], B[2], eta, dimX, dimY, nz, ndim);

    /* iteration order of FGP_dTV_core.c:76-113 (2D) and :129-169 (3D) */
    for (it = 0; it < iterationsNumb; it++) {
        int check = (epsil != 0) && (it % 5 == 0);
        REAL m;
        if (check) copy_field(Output, prev, n);
        dtv_project_dual(R[0], R[1], R[2], B[0], B[1], B[2], ndim, n);
        fgp_objective(Input, Output, R[0], R[1], R[2], lambdaPar, dimX, dimY, nz, ndim);
        if (nonneg == 1)
            for (t = 0; t < n; t++) if (Output[t] < 0) Output[t] = 0;
        dtv_gradient(P[0], P[1], P[2], Output, R[0], R[1], R[2], B[0], B[1], B[2], step, dimX, dimY, nz, ndim);
        project_unit(P[0], P[1], P[2], ndim, methodTV != 0, n);
        /* t_{k+1} with a double square root and a double sum, rounded once
         * (FGP_dTV_core.c:95, :149) */
        tkp1 = (REAL)(((double)1 + sqrt((double)((REAL)1 + (REAL)4 * tk * tk))) * 0.5);
        m = (tk - (REAL)1) / tkp1;
#pragma omp parallel for private(d)
        for (t = 0; t < n; t++)
            for (d = 0; d < ndim; d++) R[d][t] = P[d][t] + m * (P[d][t] - Pold[d][t]);
        for (d = 0; d < ndim; d++) copy_field(P[d], Pold[d], n);
        tk = tkp1;
        if (check && stop_check(&st, Output, prev, n, epsil, it)) break;
    }
    for (d = 0; d < 3; d++) { free(P[d]); free(Pold[d]); free(R[d]); free(B[d]); }
    free(prev);
    infovector[0] = (REAL)it;
    infovector[1] = st.re;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* PD-TV: Chambolle-Pock primal-dual.  PD_TV_core.c:40-252                   */
/* ------------------------------------------------------------------------ */

REAL FN(cr_pd)(const REAL *Input, REAL *U, REAL *infovector, REAL lambdaPar, int iterationsNumb, REAL epsil,
               REAL lipschitz_const, int methodTV, int nonneg, int dimX, int dimY, int dimZ)
{
    int ndim = dimZ <= 1 ? 2 : 3;
    idx_t nz = ndim == 2 ? 1 : dimZ, sxy = (idx_t)dimX * dimY;
    idx_t n = (idx_t)dimX * dimY * dimZ, nc = sxy * nz;
    idx_t alloc = n ? n : 1, t;
    REAL tau = lambdaPar * (REAL)0.1;
    REAL sigma = (REAL)(1.0 / (lipschitz_const * tau)); /* double division, PD_TV_core.c:51 */
    REAL theta = 1;
    REAL lt = tau / lambdaPar;
    REAL *Uold = calloc(alloc, sizeof(REAL));
    REAL *P1 = calloc(alloc, sizeof(REAL)), *P2 = calloc(alloc, sizeof(REAL)), *P3 = calloc(alloc, sizeof(REAL));
    stop_state st = {0, 0};
    int it;

    copy_field(Input, U, n);
    for (it = 0; it < iterationsNumb; it++) {
        idx_t k;
        /* dual ascent, forward differences mirrored at the last index
         * (PD_TV_core.c:160-174, :212-229) */
#pragma omp parallel for
        for (k = 0; k < nz; k++) {
            idx_t i, j;
            for (j = 0; j < dimY; j++)
                for (i = 0; i < dimX; i++) {
                    idx_t c = sxy * k + (idx_t)dimX * j + i;
                    /* the mirrored neighbour of the last index; on an axis of
                     * extent 1 (where the reference reads out of bounds) the
                     * point itself */
                    idx_t cx = (i == dimX - 1) ? (dimX > 1 ? c - 1 : c) : c + 1;
                    idx_t cy = (j == dimY - 1) ? (dimY > 1 ? c - dimX : c) : c + dimX;
                    P1[c] += sigma * (U[cx] - U[c]);
                    P2[c] += sigma * (U[cy] - U[c]);
                    if (ndim == 3)
                        P3[c] += sigma * ((k == nz - 1) ? U[c - sxy] - U[c] : U[c + sxy] - U[c]);
                }
        }
        if (nonneg == 1)
            for (t = 0; t < n; t++) if (U[t] < 0) U[t] = 0;
        /* the reference projects and checks over dimX*dimY*dimZ voxels but
         * copies U_old over the 2D or 3D grid (PD_TV_core.c:79-82, :123-126) */
        project_unit(P1, P2, P3, ndim, methodTV != 0, n);
        copy_field(U, Uold, nc);
        /* primal prox step with the negative backward divergence
         * (PD_TV_core.c:177-194, :232-252); the denominator is double */
#pragma omp parallel for
        for (k = 0; k < nz; k++) {
            idx_t i, j;
            for (j = 0; j < dimY; j++)
                for (i = 0; i < dimX; i++) {
                    idx_t c = sxy * k + (idx_t)dimX * j + i;
                    REAL v1 = i == 0 ? -P1[c] : -(P1[c] - P1[c - 1]);
                    REAL v2 = j == 0 ? -P2[c] : -(P2[c] - P2[c - dimX]);
                    REAL dv = v1 + v2;
                    if (ndim == 3) dv = dv + (k == 0 ? -P3[c] : -(P3[c] - P3[c - sxy]));
                    U[c] = (REAL)((U[c] - tau * dv + lt * Input[c]) / (1.0 + lt));
                }
        }
        if ((epsil != 0) && (it % 5 == 0) && stop_check(&st, U, Uold, n, epsil, it)) break;
        /* over-relaxation, PD_TV_core.c:197-205 */
#pragma omp parallel for
        for (t = 0; t < n; t++) U[t] += theta * (U[t] - Uold[t]);
    }
    free(Uold); free(P1); free(P2); free(P3);
    infovector[0] = (REAL)it;
    infovector[1] = st.re;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* SB-TV: split Bregman with two Jacobi sweeps.  SB_TV_core.c:38-371         */
/* ------------------------------------------------------------------------ */

/* One Jacobi sweep (named gauss_seidel in the reference; it reads only
 * U_prev).  Mirror boundary on every axis; the summation order of the 2D and
 * 3D forms differs and both are kept (SB_TV_core.c:166-170, :273-279). */
static void sb_sweep(REAL *U, const REAL *f, const REAL *Up, REAL *const Dd[3], REAL *const Bb[3],
                     idx_t nx, idx_t ny, idx_t nz, int ndim, REAL lam, REAL mu)
{
    idx_t k;
    idx_t sxy = nx * ny;
    REAL nc = (REAL)1 / (mu + (ndim == 3 ? (REAL)6 : (REAL)4) * lam);
#pragma omp parallel for
    for (k = 0; k < nz; k++) {
        idx_t j, i;
        /* mirror: +1, or -1 at the last index (-1: +1 at index 0); on an axis
         * of extent 1, where the reference reads out of bounds, the point itself */
        idx_t kp = mir_p(k, nz), km = mir_m(k, nz);
        for (j = 0; j < ny; j++) {
            idx_t jp = mir_p(j, ny), jm = mir_m(j, ny);
            for (i = 0; i < nx; i++) {
                idx_t ip = mir_p(i, nx), im = mir_m(i, nx);
                idx_t c = sxy * k + nx * j + i;
                idx_t cim = sxy * k + nx * j + im, cjm = sxy * k + nx * jm + i;
                REAL acc;
                if (ndim == 2) {
                    acc = Dd[0][cim] - Dd[0][c] + Dd[1][cjm] - Dd[1][c] - Bb[0][cim] + Bb[0][c] - Bb[1][cjm] + Bb[1][c];
                    acc += Up[nx * j + ip] + Up[nx * j + im] + Up[nx * jp + i] + Up[nx * jm + i];
                } else {
                    idx_t ckm = sxy * km + nx * j + i;
                    REAL dsum = Dd[0][cim] - Dd[0][c] + Dd[1][cjm] - Dd[1][c] + Dd[2][ckm] - Dd[2][c];
                    REAL bsum = -Bb[0][cim] + Bb[0][c] - Bb[1][cjm] + Bb[1][c] - Bb[2][ckm] + Bb[2][c];
                    acc = dsum + bsum;
                    acc += Up[sxy * k + nx * j + ip] + Up[cim] + Up[sxy * k + nx * jp + i] + Up[cjm]
                           + Up[sxy * kp + nx * j + i] + Up[ckm];
                }
                acc *= lam;
                acc += mu * f[c];
                U[c] = nc * acc;
            }
        }
    }
}

/* Shrinkage of grad U + B followed by the Bregman update
 * (SB_TV_core.c:175-248, :284-371).  Forward differences mirrored at the
 * last index. */
static void sb_shrink_bregman(const REAL *U, REAL *const Dd[3], REAL *const Bb[3], idx_t nx, idx_t ny, idx_t nz,
                              int ndim, int aniso, REAL lam)
{
    idx_t k;
    idx_t sxy = nx * ny;
    REAL thr = (REAL)1 / lam;
#pragma omp parallel for
    for (k = 0; k < nz; k++) {
        idx_t j, i;
        idx_t kp = mir_p(k, nz);   /* extent 1: the point itself (see sb_sweep) */
        for (j = 0; j < ny; j++) {
            idx_t jp = mir_p(j, ny);
            for (i = 0; i < nx; i++) {
                idx_t ip = mir_p(i, nx);
                idx_t c = sxy * k + nx * j + i;
                REAL g[3], v[3] = {0, 0, 0};
                int d;
                g[0] = U[sxy * k + nx * j + ip] - U[c];
                g[1] = U[sxy * k + nx * jp + i] - U[c];
                g[2] = ndim == 3 ? U[sxy * kp + nx * j + i] - U[c] : 0;
                for (d = 0; d < ndim; d++) v[d] = g[d] + Bb[d][c];
                if (aniso) {
                    for (d = 0; d < ndim; d++) {
                        REAL s = (REAL)fabs(v[d]) - thr;
                        if (s < 0) s = 0;
                        Dd[d][c] = (v[d] != 0) ? (v[d] / (REAL)fabs(v[d])) * s : 0;
                    }
                } else {
                    REAL nrm2 = v[0] * v[0] + v[1] * v[1];
                    REAL nrm, s;
                    if (ndim == 3) nrm2 = nrm2 + v[2] * v[2];
                    nrm = SQRT_R(nrm2);
                    s = nrm - thr;
                    if (s < 0) s = 0;
                    for (d = 0; d < ndim; d++) Dd[d][c] = (nrm != 0) ? s * (v[d] / nrm) : 0;
                }
                /* The reference runs the Bregman update as a separate sweep
                 * after the shrink; it reads only this voxel's D, so it can
                 * follow immediately. */
                for (d = 0; d < ndim; d++) Bb[d][c] += g[d] - Dd[d][c];
            }
        }
    }
}

REAL FN(cr_sb)(const REAL *Input, REAL *Output, REAL *infovector, REAL mu, int iter, REAL epsil, int methodTV,
               int dimX, int dimY, int dimZ)
{
    int ndim = dimZ == 1 ? 2 : 3;
    idx_t n = (idx_t)dimX * dimY * dimZ, alloc = n ? n : 1;
    REAL mu_inv = (REAL)1 / mu;
    REAL lam = (REAL)2 * mu_inv;
    REAL *prev = calloc(alloc, sizeof(REAL));
    REAL *Dd[3], *Bb[3];
    stop_state st = {0, 0};
    int it, d;

    for (d = 0; d < 3; d++) { Dd[d] = calloc(alloc, sizeof(REAL)); Bb[d] = calloc(alloc, sizeof(REAL)); }
    copy_field(Input, Output, n);
    for (it = 0; it < iter; it++) {
        copy_field(Output, prev, n);
        sb_sweep(Output, Input, prev, Dd, Bb, dimX, dimY, dimZ, ndim, lam, mu_inv);
        copy_field(Output, prev, n);
        sb_sweep(Output, Input, prev, Dd, Bb, dimX, dimY, dimZ, ndim, lam, mu_inv);
        sb_shrink_bregman(Output, Dd, Bb, dimX, dimY, dimZ, ndim, methodTV == 1, lam);
        if ((epsil != 0) && (it % 5 == 0) && stop_check(&st, Output, prev, n, epsil, it)) break;
    }
    for (d = 0; d < 3; d++) { free(Dd[d]); free(Bb[d]); }
    free(prev);
    infovector[0] = (REAL)it;
    infovector[1] = st.re;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* TGV^2-L2: primal-dual (Bredies).  TGV_core.c:43-568                       */
/* ------------------------------------------------------------------------ */

typedef struct {
    idx_t nx, ny, nz, sxy;
    int ndim;
} grid3;

/* P += sigma*(grad U - V), forward difference zero at the last index, then
 * the alpha1-ball projection (TGV_core.c:200-232, :351-387). */
static void tgv_dual_p(const grid3 *g, const REAL *U, REAL *const V[3], REAL *const P[3], REAL sigma, REAL alpha1)
{
    idx_t k;
#pragma omp parallel for
    for (k = 0; k < g->nz; k++) {
        idx_t j, i;
        for (j = 0; j < g->ny; j++)
            for (i = 0; i < g->nx; i++) {
                idx_t c = g->sxy * k + g->nx * j + i;
                REAL nrm;
                P[0][c] += sigma * ((i == g->nx - 1) ? -V[0][c] : (U[c + 1] - U[c]) - V[0][c]);
                P[1][c] += sigma * ((j == g->ny - 1) ? -V[1][c] : (U[c + g->nx] - U[c]) - V[1][c]);
                if (g->ndim == 3) {
                    P[2][c] += sigma * ((k == g->nz - 1) ? -V[2][c] : (U[c + g->sxy] - U[c]) - V[2][c]);
                    nrm = SQRT_R(P[0][c] * P[0][c] + P[1][c] * P[1][c] + P[2][c] * P[2][c]) / alpha1;
                } else {
                    nrm = SQRT_R(P[0][c] * P[0][c] + P[1][c] * P[1][c]) / alpha1;
                }
                if (nrm > 1) {
                    P[0][c] /= nrm;
                    P[1][c] /= nrm;
                    if (g->ndim == 3) P[2][c] /= nrm;
                }
            }
    }
}

/* Q += sigma * sym-grad V, forward differences zero at the last index, then
 * the alpha0-ball projection with the off-diagonals counted twice
 * (TGV_core.c:234-275, :389-446).  Channel order: 2D Q11,Q22,Q12;
 * 3D Q11,Q22,Q33,Q12,Q13,Q23. */
static void tgv_dual_q(const grid3 *g, REAL *const V[3], REAL *const Q[6], REAL sigma, REAL alpha0)
{
    idx_t k;
#pragma omp parallel for
    for (k = 0; k < g->nz; k++) {
        idx_t j, i;
        for (j = 0; j < g->ny; j++)
            for (i = 0; i < g->nx; i++) {
                idx_t c = g->sxy * k + g->nx * j + i;
                /* dV[a][b] = forward difference of V_a along axis b */
                REAL dV[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                int a;
                REAL nrm;
                for (a = 0; a < g->ndim; a++) {
                    if (i != g->nx - 1) dV[a][0] = V[a][c + 1] - V[a][c];
                    if (j != g->ny - 1) dV[a][1] = V[a][c + g->nx] - V[a][c];
                    if (g->ndim == 3 && k != g->nz - 1) dV[a][2] = V[a][c + g->sxy] - V[a][c];
                }
                if (g->ndim == 2) {
                    Q[0][c] += sigma * dV[0][0];
                    Q[1][c] += sigma * dV[1][1];
                    Q[2][c] += sigma * ((REAL)0.5 * (dV[1][0] + dV[0][1]));
                    nrm = SQRT_R(Q[0][c] * Q[0][c] + Q[1][c] * Q[1][c] + (REAL)2 * Q[2][c] * Q[2][c]);
                    nrm = nrm / alpha0;
                    if (nrm > 1) { Q[0][c] /= nrm; Q[1][c] /= nrm; Q[2][c] /= nrm; }
                } else {
                    int q;
                    Q[0][c] += sigma * dV[0][0];
                    Q[1][c] += sigma * dV[1][1];
                    Q[2][c] += sigma * dV[2][2];
                    Q[3][c] += sigma * ((REAL)0.5 * (dV[1][0] + dV[0][1]));
                    Q[4][c] += sigma * ((REAL)0.5 * (dV[2][0] + dV[0][2]));
                    Q[5][c] += sigma * ((REAL)0.5 * (dV[2][1] + dV[1][2]));
                    nrm = SQRT_R(Q[0][c] * Q[0][c] + Q[1][c] * Q[1][c] + Q[2][c] * Q[2][c]
                                 + (REAL)2 * Q[3][c] * Q[3][c] + (REAL)2 * Q[4][c] * Q[4][c]
                                 + (REAL)2 * Q[5][c] * Q[5][c]);
                    nrm = nrm / alpha0;
                    if (nrm > 1)
                        for (q = 0; q < 6; q++) Q[q][c] /= nrm;
                }
            }
    }
}

/* Adjoint (backward) difference used by the TGV primal updates: the field
 * itself at index 0, minus its predecessor at the last index, the plain
 * backward difference in between (TGV_core.c:286-292, :458-466). */
static REAL tgv_bdiff(const REAL *a, idx_t c, idx_t pos, idx_t len, idx_t stride)
{
    if (pos == 0) return a[c];
    if (pos == len - 1) return -a[c - stride];
    return a[c] - a[c - stride];
}

/* U <- extrapolated prox step, V <- extrapolated update
 * (TGV_core.c:277-345, :448-568). */
static void tgv_primal(const grid3 *g, REAL *U, REAL *Uold, const REAL *f, REAL *const V[3], REAL *const Vold[3],
                       REAL *const P[3], REAL *const Q[6], REAL lambda, REAL tau)
{
    idx_t k, t, n = g->sxy * g->nz;
    int a;
    copy_field(U, Uold, n);
#pragma omp parallel for
    for (k = 0; k < g->nz; k++) {
        idx_t j, i;
        for (j = 0; j < g->ny; j++)
            for (i = 0; i < g->nx; i++) {
                idx_t c = g->sxy * k + g->nx * j + i;
                REAL dv = tgv_bdiff(P[0], c, i, g->nx, 1) + tgv_bdiff(P[1], c, j, g->ny, g->nx);
                if (g->ndim == 3) dv = dv + tgv_bdiff(P[2], c, k, g->nz, g->sxy);
                U[c] = (lambda * (U[c] + tau * dv) + tau * f[c]) / (lambda + tau);
            }
    }
#pragma omp parallel for
    for (t = 0; t < n; t++) U[t] = (REAL)2 * U[t] - Uold[t];
    for (a = 0; a < g->ndim; a++) copy_field(V[a], Vold[a], n);
#pragma omp parallel for
    for (k = 0; k < g->nz; k++) {
        idx_t j, i;
        for (j = 0; j < g->ny; j++)
            for (i = 0; i < g->nx; i++) {
                idx_t c = g->sxy * k + g->nx * j + i;
                if (g->ndim == 2) {
                    REAL d1 = tgv_bdiff(Q[0], c, i, g->nx, 1) + tgv_bdiff(Q[2], c, j, g->ny, g->nx);
                    REAL d2 = tgv_bdiff(Q[2], c, i, g->nx, 1) + tgv_bdiff(Q[1], c, j, g->ny, g->nx);
                    V[0][c] += tau * (P[0][c] + d1);
                    V[1][c] += tau * (P[1][c] + d2);
                } else {
                    REAL d1 = tgv_bdiff(Q[0], c, i, g->nx, 1) + tgv_bdiff(Q[3], c, j, g->ny, g->nx)
                              + tgv_bdiff(Q[4], c, k, g->nz, g->sxy);
                    REAL d2 = tgv_bdiff(Q[3], c, i, g->nx, 1) + tgv_bdiff(Q[1], c, j, g->ny, g->nx)
                              + tgv_bdiff(Q[5], c, k, g->nz, g->sxy);
                    REAL d3 = tgv_bdiff(Q[4], c, i, g->nx, 1) + tgv_bdiff(Q[5], c, j, g->ny, g->nx)
                              + tgv_bdiff(Q[2], c, k, g->nz, g->sxy);
                    V[0][c] += tau * (P[0][c] + d1);
                    V[1][c] += tau * (P[1][c] + d2);
                    V[2][c] += tau * (P[2][c] + d3);
                }
            }
    }
#pragma omp parallel for private(a)
    for (t = 0; t < n; t++)
        for (a = 0; a < g->ndim; a++) V[a][t] = (REAL)2 * V[a][t] - Vold[a][t];
}

REAL FN(cr_tgv)(const REAL *U0, REAL *U, REAL *infovector, REAL lambda, REAL alpha1, REAL alpha0, int iter,
                REAL L2, REAL epsil, int dimX, int dimY, int dimZ)
{
    grid3 g;
    idx_t n = (idx_t)dimX * dimY * dimZ, alloc = n ? n : 1;
    REAL tau = (REAL)pow(L2, -0.5), sigma = (REAL)pow(L2, -0.5);
    REAL *Uold = calloc(alloc, sizeof(REAL));
    REAL *P[3], *Q[6], *V[3], *Vold[3];
    stop_state st = {0, 0};
    int it, a;

    g.nx = dimX; g.ny = dimY; g.nz = dimZ; g.sxy = (idx_t)dimX * dimY;
    g.ndim = dimZ == 1 ? 2 : 3;
    for (a = 0; a < 3; a++) {
        P[a] = calloc(alloc, sizeof(REAL));
        V[a] = calloc(alloc, sizeof(REAL));
        Vold[a] = calloc(alloc, sizeof(REAL));
    }
    for (a = 0; a < 6; a++) Q[a] = calloc(alloc, sizeof(REAL));
    copy_field(U0, U, n);
    for (it = 0; it < iter; it++) {
        tgv_dual_p(&g, U, V, P, sigma, alpha1);
        tgv_dual_q(&g, V, Q, sigma, alpha0);
        tgv_primal(&g, U, Uold, U0, V, Vold, P, Q, lambda, tau);
        if ((epsil != 0) && (it % 5 == 0) && stop_check(&st, U, Uold, n, epsil, it)) break;
    }
    for (a = 0; a < 3; a++) { free(P[a]); free(V[a]); free(Vold[a]); }
    for (a = 0; a < 6; a++) free(Q[a]);
    free(Uold);
    infovector[0] = (REAL)it;
    infovector[1] = st.re;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* LLT-ROF: explicit descent on ROF-TV plus the Lysaker-Lundervold-Tai       */
/* second-order term.  LLT_ROF_core.c:54-436                                 */
/* ------------------------------------------------------------------------ */
#define LLT_EPS 1.0e-12   /* double literals: LLT_ROF_core.c:21-22 */

/* D_d = u_dd / (|u_dd| + 1e-12) along axis d, the denominator rounded from
 * double (der2D_LLT / der3D_LLT, LLT_ROF_core.c:131-191). */
static void llt_second(const REAL *u, REAL *D1, REAL *D2, REAL *D3, idx_t nx, idx_t ny, idx_t nz, int ndim)
{
    idx_t k;
    idx_t sxy = nx * ny;
#pragma omp parallel for
    for (k = 0; k < nz; k++) {
        idx_t j, i;
        idx_t kp = mir_p(k, nz), km = mir_m(k, nz);
        for (j = 0; j < ny; j++) {
            idx_t jp = mir_p(j, ny), jm = mir_m(j, ny);
            for (i = 0; i < nx; i++) {
                idx_t c = sxy * k + nx * j + i;
                REAL uc = u[c];
                REAL dxx = u[sxy * k + nx * j + mir_p(i, nx)] - (REAL)2 * uc + u[sxy * k + nx * j + mir_m(i, nx)];
                REAL dyy = u[sxy * k + nx * jp + i] - (REAL)2 * uc + u[sxy * k + nx * jm + i];
                D1[c] = dxx / (REAL)(fabs((double)dxx) + LLT_EPS);
                D2[c] = dyy / (REAL)(fabs((double)dyy) + LLT_EPS);
                if (ndim == 3) {
                    REAL dzz = u[sxy * kp + nx * j + i] - (REAL)2 * uc + u[sxy * km + nx * j + i];
                    D3[c] = dzz / (REAL)(fabs((double)dzz) + LLT_EPS);
                }
            }
        }
    }
}

/* u += tau*(lambdaROF*div D_rof - lambdaLLT*lap D_llt - (u - f))
 * (Update2D_LLT_ROF / Update3D_LLT_ROF, LLT_ROF_core.c:369-436). */
static void llt_update(const REAL *const L[3], const REAL *const R[3], REAL *u, const REAL *f, REAL lam_rof,
                       REAL lam_llt, REAL tau, idx_t nx, idx_t ny, idx_t nz, int ndim)
{
    idx_t k;
    idx_t sxy = nx * ny;
#pragma omp parallel for
    for (k = 0; k < nz; k++) {
        idx_t j, i;
        idx_t kp = mir_p(k, nz), km = mir_m(k, nz);
        for (j = 0; j < ny; j++) {
            idx_t jp = mir_p(j, ny), jm = mir_m(j, ny);
            for (i = 0; i < nx; i++) {
                idx_t c = sxy * k + nx * j + i;
                idx_t ip = sxy * k + nx * j + mir_p(i, nx), im = sxy * k + nx * j + mir_m(i, nx);
                idx_t cjp = sxy * k + nx * jp + i, cjm = sxy * k + nx * jm + i;
                REAL lap = (L[0][ip] - (REAL)2 * L[0][c] + L[0][im]) + (L[1][cjp] - (REAL)2 * L[1][c] + L[1][cjm]);
                REAL dv = (R[0][c] - R[0][cjm]) + (R[1][c] - R[1][im]);
                if (ndim == 3) {
                    idx_t ckp = sxy * kp + nx * j + i, ckm = sxy * km + nx * j + i;
                    lap = lap + (L[2][ckp] - (REAL)2 * L[2][c] + L[2][ckm]);
                    dv = dv + (R[2][c] - R[2][ckm]);
                }
                u[c] += tau * (lam_rof * dv - lam_llt * lap - (u[c] - f[c]));
            }
        }
    }
}

REAL FN(cr_llt)(const REAL *Input, REAL *Output, REAL *infovector, REAL lambdaROF, REAL lambdaLLT, int iterationsNumb,
                REAL tau, REAL epsil, int dimX, int dimY, int dimZ)
{
    int ndim = dimZ == 1 ? 2 : 3;
    idx_t n = (idx_t)dimX * dimY * dimZ, alloc = n ? n : 1;
    REAL *L[3], *R[3], *prev = NULL;
    stop_state st = {0, 0};
    int it, d;

    for (d = 0; d < 3; d++) {
        L[d] = calloc(alloc, sizeof(REAL));
        R[d] = calloc(alloc, sizeof(REAL));
    }
    copy_field(Input, Output, n);
    if (epsil != 0) prev = calloc(alloc, sizeof(REAL));
    for (it = 0; it < iterationsNumb; it++) {
        int check = (epsil != 0) && (it % 5 == 0);
        if (check) copy_field(Output, prev, n);
        /* the ROF fluxes of ROF-TV with EPS 1e-12 (LLT_ROF_core.c:198-363) */
        rof_dir(Output, R[0], dimX, dimY, dimZ, 0, ndim, LLT_EPS);
        rof_dir(Output, R[1], dimX, dimY, dimZ, 1, ndim, LLT_EPS);
        if (ndim == 3) rof_dir(Output, R[2], dimX, dimY, dimZ, 2, ndim, LLT_EPS);
        llt_second(Output, L[0], L[1], L[2], dimX, dimY, dimZ, ndim);
        llt_update((const REAL *const *)L, (const REAL *const *)R, Output, Input, lambdaROF, lambdaLLT, tau, dimX,
                   dimY, dimZ, ndim);
        if (check && stop_check(&st, Output, prev, n, epsil, it)) break;
    }
    for (d = 0; d < 3; d++) { free(L[d]); free(R[d]); }
    free(prev);
    infovector[0] = (REAL)it;
    infovector[1] = st.re;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* NDF: linear and nonlinear (Huber, Perona-Malik, Tukey, threshold-linear,  */
/* threshold-Huber) diffusion, explicit scheme.  Diffusion_core.c:53-413     */
/* ------------------------------------------------------------------------ */

/* The penalty's flux for a neighbour difference d (Diffusion_core.c:166-231). */
static REAL ndf_flux(REAL d, REAL sigma, int penalty)
{
    switch (penalty) {
    case 1:   /* Huber */
        return fabs(d) > sigma ? (REAL)sgn_of(d) : d / sigma;
    case 2: { /* Perona-Malik; powf(x, 2) is x*x */
        REAL q = d / sigma;
        return d / ((REAL)1 + q * q);
    }
    case 3:   /* Tukey biweight */
        if (fabs(d) <= sigma) {
            REAL q = d / sigma;
            REAL t = (REAL)1 - q * q;
            return d * (t * t);
        }
        return 0;
    case 4:   /* threshold-constrained linear */
        return fabs(d) > sigma ? (REAL)0 : d;
    default:  /* 5: threshold-constrained Huber */
        if (fabs(d) <= (REAL)2 * sigma) return fabs(d) > sigma ? (REAL)sgn_of(d) : d / sigma;
        return 0;
    }
}

/* One explicit step.  in_place = 0: every neighbour is read from the
 * previous iterate (the synchronous explicit scheme).  in_place = 1: the
 * reference's loop order run serially, each point updated in place so later
 * points see earlier updates (Diffusion_core.c:108-413 on one thread). */
static void ndf_step(const REAL *f, const REAL *uin, REAL *uout, REAL lambda, REAL sigma, REAL tau, int linear,
                     int penalty, idx_t nx, idx_t ny, idx_t nz, int ndim, int in_place)
{
    idx_t k;
    idx_t sxy = nx * ny;
    const REAL *src = in_place ? uout : uin;
#pragma omp parallel for if (!in_place)
    for (k = 0; k < nz; k++) {
        idx_t j, i;
        idx_t kp = mir_p(k, nz), km = mir_m(k, nz);
        for (j = 0; j < ny; j++) {
            idx_t jp = mir_p(j, ny), jm = mir_m(j, ny);
            for (i = 0; i < nx; i++) {
                idx_t c = sxy * k + nx * j + i;
                REAL uc = src[c];
                REAL d[6];
                int nd = ndim == 3 ? 6 : 4, q;
                REAL s;
                d[0] = src[sxy * k + nx * j + mir_p(i, nx)] - uc;   /* e */
                d[1] = src[sxy * k + nx * j + mir_m(i, nx)] - uc;   /* w */
                d[2] = src[sxy * k + nx * jp + i] - uc;             /* n */
                d[3] = src[sxy * k + nx * jm + i] - uc;             /* s */
                if (ndim == 3) {
                    d[4] = src[sxy * kp + nx * j + i] - uc;         /* u */
                    d[5] = src[sxy * km + nx * j + i] - uc;         /* d */
                }
                if (!linear)
                    for (q = 0; q < nd; q++) d[q] = ndf_flux(d[q], sigma, penalty);
                s = d[0] + d[1] + d[2] + d[3];
                if (ndim == 3) s = s + d[4] + d[5];
                uout[c] = uc + tau * (lambda * s - (uc - f[c]));
            }
        }
    }
}

static REAL ndf_run(const REAL *Input, REAL *Output, REAL *infovector, REAL lambdaPar, REAL sigmaPar,
                    int iterationsNumb, REAL tau, int penaltytype, REAL epsil, int dimX, int dimY, int dimZ,
                    int in_place)
{
    int ndim = dimZ == 1 ? 2 : 3;
    idx_t n = (idx_t)dimX * dimY * dimZ, alloc = n ? n : 1;
    /* sigma / sqrt(2) with a double square root and a double division
     * (Diffusion_core.c:57) */
    REAL sigma2 = (REAL)((double)sigmaPar / sqrt(2.0));
    int linear = sigmaPar == 0;
    REAL *prev = NULL, *tmp = in_place ? NULL : calloc(alloc, sizeof(REAL));
    stop_state st = {0, 0};
    int it;

    copy_field(Input, Output, n);
    if (epsil != 0) prev = calloc(alloc, sizeof(REAL));
    for (it = 0; it < iterationsNumb; it++) {
        int check = (epsil != 0) && (it % 5 == 0);
        if (check) copy_field(Output, prev, n);
        if (in_place) {
            ndf_step(Input, Output, Output, lambdaPar, sigma2, tau, linear, penaltytype, dimX, dimY, dimZ, ndim, 1);
        } else {
            copy_field(Output, tmp, n);
            ndf_step(Input, tmp, Output, lambdaPar, sigma2, tau, linear, penaltytype, dimX, dimY, dimZ, ndim, 0);
        }
        if (check && stop_check(&st, Output, prev, n, epsil, it)) break;
    }
    free(prev);
    free(tmp);
    infovector[0] = (REAL)it;
    infovector[1] = st.re;
    return 0;
}

REAL FN(cr_ndf)(const REAL *Input, REAL *Output, REAL *infovector, REAL lambdaPar, REAL sigmaPar, int iterationsNumb,
                REAL tau, int penaltytype, REAL epsil, int dimX, int dimY, int dimZ)
{
    return ndf_run(Input, Output, infovector, lambdaPar, sigmaPar, iterationsNumb, tau, penaltytype, epsil, dimX, dimY,
                   dimZ, 0);
}

REAL FN(cr_ndf_inplace)(const REAL *Input, REAL *Output, REAL *infovector, REAL lambdaPar, REAL sigmaPar,
                        int iterationsNumb, REAL tau, int penaltytype, REAL epsil, int dimX, int dimY, int dimZ)
{
    return ndf_run(Input, Output, infovector, lambdaPar, sigmaPar, iterationsNumb, tau, penaltytype, epsil, dimX, dimY,
                   dimZ, 1);
}

/* ------------------------------------------------------------------------ */
/* Diff4th: anisotropic fourth-order diffusion (Hajiaboli), explicit scheme. */
/* Diffus4th_order_core.c:44-265                                             */
/* ------------------------------------------------------------------------ */
#define D4_EPS 1.0e-7   /* a double literal: Diffus4th_order_core.c:23 */

/* The weighted Laplacian W of u (Weighted_Laplc2D / 3D, :103-148, :176-235):
 * central first and second differences with mirrored neighbours, the
 * normal / tangential second derivatives divided by |grad|^2 (or, below
 * 1e-7, by the double 1e-7), weighted by c = 1/(1 + |grad|^2/sigma^2). */
static void d4_wlap(const REAL *u, REAL *W, REAL sigma2, idx_t nx, idx_t ny, idx_t nz, int ndim)
{
    idx_t k;
    idx_t sxy = nx * ny;
#pragma omp parallel for
    for (k = 0; k < nz; k++) {
        idx_t j, i;
        idx_t k1 = mir_p(k, nz), k2 = mir_m(k, nz);
        for (j = 0; j < ny; j++) {
            idx_t j1 = mir_p(j, ny), j2 = mir_m(j, ny);
            for (i = 0; i < nx; i++) {
                idx_t i1 = mir_p(i, nx), i2 = mir_m(i, nx);
                idx_t z0 = sxy * k;
                idx_t c = z0 + nx * j + i;
                REAL uc = u[c];
                REAL gx = (REAL)0.5 * (u[z0 + nx * j + i2] - u[z0 + nx * j + i1]);
                REAL gy = (REAL)0.5 * (u[z0 + nx * j2 + i] - u[z0 + nx * j1 + i]);
                REAL gxs = gx * gx, gys = gy * gy;
                REAL gxx = u[z0 + nx * j + i2] + u[z0 + nx * j + i1] - (REAL)2 * uc;
                REAL gyy = u[z0 + nx * j2 + i] + u[z0 + nx * j1 + i] - (REAL)2 * uc;
                REAL gxy = (REAL)0.25 * (u[z0 + nx * j2 + i2] + u[z0 + nx * j1 + i1] - u[z0 + nx * j1 + i2] -
                                         u[z0 + nx * j2 + i1]);
                REAL xy2 = (REAL)2 * gx * gy * gxy;
                REAL den, nn, no, vn, vo, cw;
                if (ndim == 3) {
                    idx_t zp = sxy * k1, zm = sxy * k2;
                    REAL gz = (REAL)0.5 * (u[zm + nx * j + i] - u[zp + nx * j + i]);
                    REAL gzs = gz * gz;
                    REAL gzz = u[zm + nx * j + i] + u[zp + nx * j + i] - (REAL)2 * uc;
                    REAL gxz = (REAL)0.25 * (u[zm + nx * j + i2] - u[zm + nx * j + i1] - u[zp + nx * j + i2] +
                                             u[zp + nx * j + i1]);
                    REAL gyz = (REAL)0.25 * (u[zm + nx * j2 + i] - u[zm + nx * j1 + i] - u[zp + nx * j2 + i] +
                                             u[zp + nx * j1 + i]);
                    REAL xz1 = (REAL)2 * gx * gz * gxz, yz2 = (REAL)2 * gy * gz * gyz;
                    den = gxs + gys + gzs;
                    nn = gxx * gxs + gyy * gys + gzz * gzs + xy2 + xz1 + yz2;
                    no = (gys + gzs) * gxx + (gxs + gzs) * gyy + (gxs + gys) * gzz - xy2 - xz1 - yz2;
                } else {
                    den = gxs + gys;
                    nn = gxx * gxs + xy2 + gyy * gys;
                    no = gxx * gys - xy2 + gyy * gxs;
                }
                if (den <= D4_EPS) {
                    vn = (REAL)(nn / D4_EPS);
                    vo = (REAL)(no / D4_EPS);
                } else {
                    vn = nn / den;
                    vo = no / den;
                }
                cw = (REAL)1 / ((REAL)1 + den / sigma2);
                W[c] = cw * cw * vn + cw * vo;
            }
        }
    }
}

/* u += tau*(-lambda*lap W - (u - f)) (Diffusion_update_step2D / 3D, :150-172, :237-265). */
static void d4_update(const REAL *W, REAL *u, const REAL *f, REAL lambda, REAL tau, idx_t nx, idx_t ny, idx_t nz,
                      int ndim)
{
    idx_t k;
    idx_t sxy = nx * ny;
#pragma omp parallel for
    for (k = 0; k < nz; k++) {
        idx_t j, i;
        idx_t k1 = mir_p(k, nz), k2 = mir_m(k, nz);
        for (j = 0; j < ny; j++) {
            idx_t j1 = mir_p(j, ny), j2 = mir_m(j, ny);
            for (i = 0; i < nx; i++) {
                idx_t z0 = sxy * k;
                idx_t c = z0 + nx * j + i;
                REAL s = (W[z0 + nx * j + mir_m(i, nx)] + W[z0 + nx * j + mir_p(i, nx)] - (REAL)2 * W[c]) +
                         (W[z0 + nx * j2 + i] + W[z0 + nx * j1 + i] - (REAL)2 * W[c]);
                if (ndim == 3) s = s + (W[sxy * k2 + nx * j + i] + W[sxy * k1 + nx * j + i] - (REAL)2 * W[c]);
                u[c] += tau * (-lambda * s - (u[c] - f[c]));
            }
        }
    }
}

REAL FN(cr_d4)(const REAL *Input, REAL *Output, REAL *infovector, REAL lambdaPar, REAL sigmaPar, int iterationsNumb,
               REAL tau, REAL epsil, int dimX, int dimY, int dimZ)
{
    int ndim = dimZ == 1 ? 2 : 3;
    idx_t n = (idx_t)dimX * dimY * dimZ, alloc = n ? n : 1;
    REAL sigma2 = sigmaPar * sigmaPar;
    REAL *W = calloc(alloc, sizeof(REAL)), *prev = NULL;
    stop_state st = {0, 0};
    int it;

    copy_field(Input, Output, n);
    if (epsil != 0) prev = calloc(alloc, sizeof(REAL));
    for (it = 0; it < iterationsNumb; it++) {
        int check = (epsil != 0) && (it % 5 == 0);
        if (check) copy_field(Output, prev, n);
        d4_wlap(Output, W, sigma2, dimX, dimY, dimZ, ndim);
        d4_update(W, Output, Input, lambdaPar, tau, dimX, dimY, dimZ, ndim);
        if (check && stop_check(&st, Output, prev, n, epsil, it)) break;
    }
    free(W);
    free(prev);
    infovector[0] = (REAL)it;
    infovector[1] = st.re;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Objective of the ROF model in FP64, used for the "final objective" of the */
/* precision study: TV_energy2D/3D of utils.c:64-117 with the same replicate */
/* boundary (the forward difference at the last index is 0) and the same    */
/* `type` (1: sum 2*lambda*|grad+ u| + sum (u - u0)^2; 2: the gradient term  */
/* alone), accumulated in double in index order instead of the reference's  */
/* serial float sum.  A 2D image is dimZ = 1 (its z difference is then 0,   */
/* which is TV_energy2D).                                                    */
/* ------------------------------------------------------------------------ */
#ifndef CR_DOUBLE
double cr_tv_energy_f64(const float *U, const float *U0, double lambda, int type, int dimX, int dimY, int dimZ)
{
    idx_t sxy = (idx_t)dimX * dimY, k, j, i;
    double e_grad = 0, e_data = 0;
    for (k = 0; k < dimZ; k++)
        for (j = 0; j < dimY; j++)
            for (i = 0; i < dimX; i++) {
                idx_t c = sxy * k + (idx_t)dimX * j + i;
                double u = U[c];
                double gx = (i + 1 < dimX ? U[c + 1] : u) - u;
                double gy = (j + 1 < dimY ? U[c + dimX] : u) - u;
                double gz = (k + 1 < dimZ ? U[c + sxy] : u) - u;
                e_grad += 2.0 * lambda * sqrt(gx * gx + gy * gy + gz * gz);
                e_data += (u - U0[c]) * (u - U0[c]);
            }
    return type == 2 ? e_grad : e_grad + e_data;
}
#endif
