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
    __shared__ double red[8];

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
        q1 = p1 + sigma * (ux - u);
        q2 = p2 + sigma * (uy - u);
        q3 = 0;
        if (DIM == 3) {
            const I jz = g.nz == 1 ? i : (z == g.nz - 1 ? i - sxy : i + sxy);
            const T uz = ldx<T>(a.U, jz), p3 = ldx<T>(a.P[2], i);
            q3 = p3 + sigma * (uz - u);
        }
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
            const T v1 = (x == 0) ? -q1 : -(q1 - Pl[0][ly + 1][lx]);
            const T v2 = (y == 0) ? -q2 : -(q2 - Pl[1][ly][lx + 1]);
            T dv = v1 + v2;
            if (DIM == 3) {
                const T v3 = (z == 0) ? -q3 : -(q3 - p3below);
                dv = dv + v3;
            }
            const T un = Mixed<T>::pd_div(u0 - a.tau * dv + a.lt * ldx<T>(a.f, i), a.lt);
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

}  // namespace cilreg
