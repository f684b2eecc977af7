// pd_kernels.hpp -- device code of the PD iteration: argument block,
// per-tile body and the single-volume kernel.  The body is shared by the
// single-volume kernel (pd.hip) and the grouped multi-volume kernel
// (batch.hip); the algorithm and its reference citations are described at the
// top of pd.hip.
#pragma once
#include "runtime.hpp"

namespace cilreg {

template <typename T, typename S>
struct PdArgs {
    const T* f;
    const T* U;       // u_bar in
    const S* P[3];    // P in
    T* Uout;          // u_bar out
    S* Pout[3];       // P out
    T* Ubrk;          // U before extrapolation (check iterations)
    T sigma, tau, lt, theta;
    int nonneg, check;
    double* partials;
    const Ctrl* ctrl;
    Geom g;
};

// The dual step at one point (DualP3D, PD_TV_core.c:212-229): P of the next
// iterate from the old P there and u_bar there and at the forward neighbours
// (the caller passes the mirrored backward neighbour at the last index),
// projected (utils.c:184-211).  Shared by the PD kernels.
template <int DIM, typename T, bool ISO>
__device__ inline void pd_dual_math(T sigma, T u, T ux, T uy, T uz, T p1, T p2, T p3, T& q1, T& q2, T& q3) {
    q1 = p1 + sigma * (ux - u);
    q2 = p2 + sigma * (uy - u);
    q3 = 0;
    if (DIM == 3) q3 = p3 + sigma * (uz - u);
    if (ISO) {
        T n2 = q1 * q1 + q2 * q2;
        if (DIM == 3) n2 = n2 + q3 * q3;
        if (n2 > (T)1) {
            const T s = (T)1 / dsqrt(n2);
            q1 = q1 * s;
            q2 = q2 * s;
            if (DIM == 3) q3 = q3 * s;
        }
    } else {
        T a1 = fabs(q1), a2 = fabs(q2);
        if (a1 < (T)1) a1 = 1;
        if (a2 < (T)1) a2 = 1;
        q1 = q1 / a1;
        q2 = q2 / a2;
        if (DIM == 3) {
            T a3 = fabs(q3);
            if (a3 < (T)1) a3 = 1;
            q3 = q3 / a3;
        }
    }
}

// The primal step at one point (PD_TV_core.c:120, 232-252): U from the
// (clipped) u_bar u0, the new P there (q) and at the backward neighbours
// (qxm, qym, qzm; P(-1) = 0), and f.
template <int DIM, typename T>
__device__ inline T pd_primal_math(T tau, T lt, int x, int y, int z, T q1, T q2, T q3, T qxm, T qym, T qzm, T u0,
                                   T f) {
    const T v1 = (x == 0) ? -q1 : -(q1 - qxm);
    const T v2 = (y == 0) ? -q2 : -(q2 - qym);
    T dv = v1 + v2;
    if (DIM == 3) {
        const T v3 = (z == 0) ? -q3 : -(q3 - qzm);
        dv = dv + v3;
    }
    return Mixed<T>::pd_div(u0 - tau * dv + lt * f, lt);
}

template <int DIM, int TY, typename T, typename S, bool ISO, typename I>
__device__ inline void pd_body(const PdArgs<T, S>& a, const TileCoord& tc) {
    const Geom& g = a.g;
    constexpr int NT = TX * TY, NH = TX + TY;
    const int lx = threadIdx.x, ly = threadIdx.y, tid = ly * TX + lx;
    const int x = tc.x0 + lx, y = tc.y0 + ly;
    const bool own = x < g.nx && y < g.ny;
    // -x halo column (TY points) and -y halo row (TX points), on the low or
    // high waves depending on the block index (spreads the extra pass over
    // the SIMDs of a CU)
    const int hbase = (tc.lin & 1) ? NT - NH : 0;
    const int h = tid - hbase;
    const bool hlane = h >= 0 && h < NH;
    int hx, hy, hlx, hly;
    if (h < TY) { hx = tc.x0 - 1; hy = tc.y0 + h; hlx = 0; hly = h + 1; }
    else { hx = tc.x0 + h - TY; hy = tc.y0 - 1; hlx = h - TY + 1; hly = 0; }
    const bool hown = hlane && hx >= 0 && hy >= 0 && hx < g.nx && hy < g.ny;
    const int ox = own ? x : 0, oy = own ? y : 0;
    const int hcx = hown ? hx : 0, hcy = hown ? hy : 0;

    __shared__ T Pl[2][TY + 1][TX + 1];
    __shared__ double red[kRedSlots];

    const I sxy = (I)g.sxy, nx = (I)g.nx;
    const T sigma = a.sigma;

    // updated, projected dual at (px, py, z); the forward neighbour is
    // mirrored to the backward one at the last index (PD_TV_core.c:221-226)
    auto dual = [&](int px, int py, int z, T& q1, T& q2, T& q3) {
        const I i = (I)z * sxy + (I)py * nx + (I)px;
        const I jx = g.nx == 1 ? i : (px == g.nx - 1 ? i - 1 : i + 1);
        const I jy = g.ny == 1 ? i : (py == g.ny - 1 ? i - nx : i + nx);
        const T u = ldx<T>(a.U, i), ux = ldx<T>(a.U, jx), uy = ldx<T>(a.U, jy);
        const T p1 = ldx<T>(a.P[0], i), p2 = ldx<T>(a.P[1], i);
        T uz = 0, p3 = 0;
        if (DIM == 3) {
            const I jz = g.nz == 1 ? i : (z == g.nz - 1 ? i - sxy : i + sxy);
            uz = ldx<T>(a.U, jz);
            p3 = ldx<T>(a.P[2], i);
        }
        pd_dual_math<DIM, T, ISO>(sigma, u, ux, uy, uz, p1, p2, p3, q1, q2, q3);
    };

    T p3below = 0;   // updated P3 at the own point, plane z-1
    if (DIM == 3 && tc.z0 > 0) {
        T q1, q2;
        dual(ox, oy, tc.z0 - 1, q1, q2, p3below);
    }
    double sdd = 0.0, suu = 0.0;
    for (int z = tc.z0; z < tc.z1; ++z) {
        T q1, q2, q3;
        dual(ox, oy, z, q1, q2, q3);
        Pl[0][ly + 1][lx + 1] = q1;
        Pl[1][ly + 1][lx + 1] = q2;
        if (hlane) {
            T h1, h2, h3;
            dual(hcx, hcy, z, h1, h2, h3);
            Pl[0][hly][hlx] = h1;
            Pl[1][hly][hlx] = h2;
        }
        __syncthreads();
        if (own) {
            const I i = (I)z * sxy + (I)y * nx + (I)x;
            T u0 = ldx<T>(a.U, i);
            if (a.nonneg && u0 < (T)0) u0 = 0;
            const T un = pd_primal_math<DIM, T>(a.tau, a.lt, x, y, z, q1, q2, q3, Pl[0][ly + 1][lx], Pl[1][ly][lx + 1],
                                                p3below, u0, ldx<T>(a.f, i));
            if (a.check) {
                const T df = un - u0;
                sdd += (double)df * (double)df;
                suu += (double)un * (double)un;
                stnt<T>(a.Ubrk, i, un);
            }
            stnt<T>(a.Uout, i, un + a.theta * (un - u0));
            stnt<T>(a.Pout[0], i, q1);
            stnt<T>(a.Pout[1], i, q2);
            if (DIM == 3) stnt<T>(a.Pout[2], i, q3);
        }
        __syncthreads();
        p3below = q3;
    }
    if (a.check) {
        block_sum2(sdd, suu, red);
        if (tid == 0) {
            a.partials[2 * tc.lin] = sdd;
            a.partials[2 * tc.lin + 1] = suu;
        }
    }
}

// One iteration on one volume: the workgroup's tile comes from its block index.
template <int DIM, int TY, typename T, typename S, bool ISO, typename I>
__global__ __launch_bounds__(TX * TY) void pd_iter_kernel(PdArgs<T, S> a) {
    if (solve_stopped(a.ctrl)) return;
    pd_body<DIM, TY, T, S, ISO, I>(a, tile_of_block(a.g, TY));
}

// ---------------------------------------------------------------------------
// Two iterations per pass (temporal blocking of depth 2, 3D)
// ---------------------------------------------------------------------------
// The PD state (u_bar, P) does not involve older iterates: a pass that
// advances k -> k+2 reads u_bar, f, P once and writes u_bar, P once (36 B per
// two voxel-iterations instead of 72), computing every value with
// pd_dual_math / pd_primal_math on the inputs the one-iteration kernel would
// see (P of k+1 rounded through the storage type), so the results are
// bit-identical.  Same ring and pipeline as tgv2_body: B = [x0-2, x0+64] x
// [y0-2, y0+TY]; per plane zi: S1 P of k+1 at zi, S2 u_bar of k+1 at zi, S3 P
// of k+2 at zi-1, S4 u_bar of k+2 at zi-1 on T.  The dual's mirrored forward
// difference at the last index reads the backward neighbour from the same
// LDS planes (the last plane's mirror is the plane below, which is still in
// the other parity buffer because the pass writes no plane beyond the
// volume).  On a check the un-extrapolated U goes to Ubrk0 (iteration k) or
// Ubrk1 (k+1), the output should the stopping rule fire there; a null Ubrk0 /
// Ubrk1 skips that store (the single-volume solver recomputes the stop
// iteration instead).
template <typename T, typename S>
struct Pd2Args {
    const T* f;
    const T* U;
    const S* P[3];
    T* Uout;
    S* Pout[3];
    T* Ubrk0;
    T* Ubrk1;
    T sigma, tau, lt, theta;
    int nonneg, check0, check1;
    double* partials0;
    double* partials1;
    const Ctrl* ctrl;
    Geom g;
};

template <int TY, typename T, typename S, bool ISO, typename I>
__device__ inline void pd2_body(const Pd2Args<T, S>& a, const TileCoord& tc) {
    constexpr int BW = TX + 3, BH = TY + 3, NR = 3 * BW + 3 * TY, NT = TX * TY;
    const Geom& g = a.g;
    const int lx = threadIdx.x, ly = threadIdx.y, tid = ly * TX + lx;

    int bx[2], by[2];
    bx[0] = lx + 2;
    by[0] = ly + 2;
    const int hbase = (tc.lin & 1) ? NT - NR : 0;
    const int h = tid - hbase;
    const bool ring = h >= 0 && h < NR;
    {
        int rx = 0, ry = 0;
        if (h < 3 * BW) {
            const int row = h / BW;
            rx = h - row * BW;
            ry = row < 2 ? row : TY + 2;
        } else {
            const int k = h - 3 * BW, col = k / TY;
            rx = col < 2 ? col : TX + 2;
            ry = k - col * TY + 2;
        }
        bx[1] = ring ? rx : 0;
        by[1] = ring ? ry : 0;
    }
    const int nslot = ring ? 2 : 1;
    int X[2], Y[2], cx[2], cy[2];
    #pragma unroll
    for (int s = 0; s < 2; ++s) {
        X[s] = tc.x0 + bx[s] - 2;
        Y[s] = tc.y0 + by[s] - 2;
        const bool in = X[s] >= 0 && Y[s] >= 0 && X[s] < g.nx && Y[s] < g.ny;
        cx[s] = in ? X[s] : 0;
        cy[s] = in ? Y[s] : 0;
    }
    const bool own = X[0] < g.nx && Y[0] < g.ny;

    __shared__ T Q[2][BH + 1][BW + 1];     // P1, P2 read at x-1 / y-1
    __shared__ T UB[2][BH + 1][BW + 1];    // u_bar of k+1, planes by parity
    __shared__ double red[kRedSlots];

    const I sxy = (I)g.sxy, nx = (I)g.nx;
    const int nz = g.nz, z0 = tc.z0, z1 = tc.z1;
    const T sigma = a.sigma, tau = a.tau, lt = a.lt, theta = a.theta;

    T q1v[2][3] = {};     // P of k+1 at the previous plane
    T zq2 = 0;            // P3 of k+2 at the previous plane (slot 0)
    T f0 = 0;             // f at the previous plane (slot 0)
    double sdd0 = 0.0, suu0 = 0.0, sdd1 = 0.0, suu1 = 0.0;
    const int zs = z0 >= 2 ? z0 - 2 : 0;
    for (int zi = zs; zi <= z1; ++zi) {
        const bool s12 = zi < nz;
        T q[2][3] = {}, ub[2] = {0, 0};
        // ---- S1: P of k+1 at zi (global loads, mirrored forward neighbours)
        if (s12) {
            #pragma unroll
            for (int s = 0; s < 2; ++s) {
                if (s < nslot) {
                    const I i = (I)zi * sxy + (I)cy[s] * nx + (I)cx[s];
                    const I jx = g.nx == 1 ? i : (cx[s] == g.nx - 1 ? i - 1 : i + 1);
                    const I jy = g.ny == 1 ? i : (cy[s] == g.ny - 1 ? i - nx : i + nx);
                    const I jz = nz == 1 ? i : (zi == nz - 1 ? i - sxy : i + sxy);
                    ub[s] = ldx<T>(a.U, i);
                    const T ux = ldx<T>(a.U, jx), uy = ldx<T>(a.U, jy), uz = ldx<T>(a.U, jz);
                    const T p1 = ldx<T>(a.P[0], i), p2 = ldx<T>(a.P[1], i), p3 = ldx<T>(a.P[2], i);
                    pd_dual_math<3, T, ISO>(sigma, ub[s], ux, uy, uz, p1, p2, p3, q[s][0], q[s][1], q[s][2]);
                    Q[0][by[s]][bx[s]] = q[s][0];
                    Q[1][by[s]][bx[s]] = q[s][1];
                }
            }
        }
        __syncthreads();
        // ---- S2: u_bar of k+1 at zi
        T fcur = 0;
        if (s12 && zi + 1 >= z0) {
            #pragma unroll
            for (int s = 0; s < 2; ++s) {
                if (s < nslot) {
                    const int xb = bx[s] > 0 ? bx[s] - 1 : 0, yb = by[s] > 0 ? by[s] - 1 : 0;
                    const I i = (I)zi * sxy + (I)cy[s] * nx + (I)cx[s];
                    const T fv = ldx<T>(a.f, i);
                    if (s == 0) fcur = fv;
                    T u0 = ub[s];
                    if (a.nonneg && u0 < (T)0) u0 = 0;
                    const T un = pd_primal_math<3, T>(tau, lt, X[s], Y[s], zi, q[s][0], q[s][1], q[s][2], Q[0][by[s]][xb],
                                                      Q[1][yb][bx[s]], q1v[s][2], u0, fv);
                    if (s == 0 && own && zi >= z0 && zi < z1 && a.check0) {
                        const T df = un - u0;
                        sdd0 += (double)df * (double)df;
                        suu0 += (double)un * (double)un;
                        if (a.Ubrk0) stnt<T>(a.Ubrk0, (I)zi * sxy + (I)Y[0] * nx + (I)X[0], un);
                    }
                    UB[zi & 1][by[s]][bx[s]] = un + theta * (un - u0);
                }
            }
        }
        __syncthreads();
        // ---- S3: P of k+2 at zc
        const int zc = zi - 1;
        T q2[2][3] = {};
        const bool s3 = zc >= (z0 >= 1 ? z0 - 1 : 0) && zc <= z1 - 1;
        if (s3) {
            #pragma unroll
            for (int s = 0; s < 2; ++s) {
                if (s < nslot) {
                    const int b = zc & 1;
                    const int xm = bx[s] > 0 ? bx[s] - 1 : 0, ym = by[s] > 0 ? by[s] - 1 : 0;
                    const T u = UB[b][by[s]][bx[s]];
                    const T ux = g.nx == 1 ? u : (X[s] == g.nx - 1 ? UB[b][by[s]][xm] : UB[b][by[s]][bx[s] + 1]);
                    const T uy = g.ny == 1 ? u : (Y[s] == g.ny - 1 ? UB[b][ym][bx[s]] : UB[b][by[s] + 1][bx[s]]);
                    const T uz = nz == 1 ? u : UB[b ^ 1][by[s]][bx[s]];   // zc+1, or zc-1 on the last plane
                    const T p1 = through_storage<T, S>(q1v[s][0]), p2 = through_storage<T, S>(q1v[s][1]),
                            p3 = through_storage<T, S>(q1v[s][2]);
                    pd_dual_math<3, T, ISO>(sigma, u, ux, uy, uz, p1, p2, p3, q2[s][0], q2[s][1], q2[s][2]);
                    Q[0][by[s]][bx[s]] = q2[s][0];
                    Q[1][by[s]][bx[s]] = q2[s][1];
                }
            }
        }
        __syncthreads();
        // ---- S4: u_bar of k+2 at zc on T; store the state
        if (s3 && zc >= z0 && own) {
            const int b = zc & 1;
            T u0 = UB[b][by[0]][bx[0]];
            if (a.nonneg && u0 < (T)0) u0 = 0;
            const T un = pd_primal_math<3, T>(tau, lt, X[0], Y[0], zc, q2[0][0], q2[0][1], q2[0][2],
                                              Q[0][by[0]][bx[0] - 1], Q[1][by[0] - 1][bx[0]], zq2, u0, f0);
            const I io = (I)zc * sxy + (I)Y[0] * nx + (I)X[0];
            if (a.check1) {
                const T df = un - u0;
                sdd1 += (double)df * (double)df;
                suu1 += (double)un * (double)un;
                if (a.Ubrk1) stnt<T>(a.Ubrk1, io, un);
            }
            stnt<T>(a.Uout, io, un + theta * (un - u0));
            stnt<T>(a.Pout[0], io, q2[0][0]);
            stnt<T>(a.Pout[1], io, q2[0][1]);
            stnt<T>(a.Pout[2], io, q2[0][2]);
        }
        __syncthreads();
        // ---- shift the pipeline
        if (s3) zq2 = q2[0][2];
        if (s12) {
            #pragma unroll
            for (int s = 0; s < 2; ++s) {
                #pragma unroll
                for (int c = 0; c < 3; ++c) q1v[s][c] = q[s][c];
            }
            f0 = fcur;
        }
    }
    if (a.check0) {
        block_sum2(sdd0, suu0, red);
        if (tid == 0) {
            a.partials0[2 * tc.lin] = sdd0;
            a.partials0[2 * tc.lin + 1] = suu0;
        }
        __syncthreads();
    }
    if (a.check1) {
        block_sum2(sdd1, suu1, red);
        if (tid == 0) {
            a.partials1[2 * tc.lin] = sdd1;
            a.partials1[2 * tc.lin + 1] = suu1;
        }
    }
}

// Registers of the pass: with FP32 arithmetic and 32-bit offsets it is capped
// at 64 VGPRs (eight waves per SIMD: four 512-thread workgroups per CU instead
// of three at the 68 it would take; 256^3: 0.111 -> 0.102 ms per iteration,
// nothing spilled).  FP64 and 64-bit offsets are not capped.
template <typename T, typename I>
constexpr int pd2_min_waves() { return (std::is_same<T, float>::value && sizeof(I) == 4) ? 8 : 1; }

template <int TY, typename T, typename S, bool ISO, typename I>
__global__ __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(pd2_min_waves<T, I>())))
void pd2_iter_kernel(Pd2Args<T, S> a) {
    if (solve_stopped(a.ctrl)) return;
    pd2_body<TY, T, S, ISO, I>(a, tile_of_block(a.g, TY));
}

}  // namespace cilreg
