// rof_kernels.hpp -- device code of the ROF iteration: argument block,
// per-tile body and the single-volume kernel.  The body is shared by the
// single-volume kernel (rof.hip) and the grouped multi-volume kernel
// (batch.hip); the algorithm and its reference citations are described at the
// top of rof.hip.
#pragma once
#include "runtime.hpp"

namespace cilreg {

template <typename T>
struct RofArgs {
    const T* f;
    const T* U;
    T* Uout;
    T lambda, tau;
    int check;
    double* partials;
    const Ctrl* ctrl;
    Geom g;
};

__device__ inline int mirror_p(int i, int n) { return i + 1 >= n ? (i - 1 >= 0 ? i - 1 : 0) : i + 1; }
__device__ inline int mirror_m(int i, int n) { return i - 1 < 0 ? (i + 1 < n ? i + 1 : 0) : i - 1; }

// The fluxes D1 (along y), D2 (along x) and, in 3D, D3 (along z) at a point
// from u there (uc) and at its mirrored neighbours (ROF_TV_core.c:102-266).
// eps is the floor added in double under the square root: 1e-8 for ROF-TV,
// 1e-12 for the ROF part of LLT-ROF (LLT_ROF_core.c:22,198-363).
template <int DIM, typename T>
__device__ inline void rof_flux_math(T uc, T uip, T uim, T ujp, T ujm, T ukp, T ukm, T& d1, T& d2, T& d3,
                                     double eps = 1.0e-8) {
    const T fi = uip - uc, bi = uc - uim, fj = ujp - uc, bj = uc - ujm;
    const T mi = Mixed<T>::minmod_sq(fi, bi), mj = Mixed<T>::minmod_sq(fj, bj);
    if (DIM == 3) {
        const T fk = ukp - uc, bk = uc - ukm;
        const T mk = Mixed<T>::minmod_sq(fk, bk);
        // ROF_TV_core.c:131-137 (double sqrt), :194-200, :257-263
        d1 = fj / Mixed<T>::rof_sqrt_d(fj * fj + mi + mk, eps);
        d2 = fi / Mixed<T>::rof_sqrt_f(fi * fi + mj + mk, eps);
        d3 = fk / Mixed<T>::rof_sqrt_f(fk * fk + mj + mi, eps);
    } else {
        // ROF_TV_core.c:157-161, :220-224
        d1 = fj / Mixed<T>::rof_sqrt_f(fj * fj + mi, eps);
        d2 = fi / Mixed<T>::rof_sqrt_f(fi * fi + mj, eps);
        d3 = 0;
    }
}

// The fluxes D1 (along y), D2 (along x) and, in 3D, D3 (along z) at
// (px, py, z) from one load of the 7-point mirrored neighbourhood.
template <int DIM, typename I, typename T>
__device__ inline void rof_fluxes(const RofArgs<T>& a, int px, int py, int z, T& d1, T& d2, T& d3) {
    const Geom& g = a.g;
    const I sxy = (I)g.sxy, nx = (I)g.nx;
    const I zo = (I)z * sxy, c = zo + (I)py * nx + (I)px;
    const T uc = ldx<T>(a.U, c);
    const T uip = ldx<T>(a.U, zo + (I)py * nx + (I)mirror_p(px, g.nx));
    const T uim = ldx<T>(a.U, zo + (I)py * nx + (I)mirror_m(px, g.nx));
    const T ujp = ldx<T>(a.U, zo + (I)mirror_p(py, g.ny) * nx + (I)px);
    const T ujm = ldx<T>(a.U, zo + (I)mirror_m(py, g.ny) * nx + (I)px);
    T ukp = 0, ukm = 0;
    if (DIM == 3) {
        ukp = ldx<T>(a.U, (I)mirror_p(z, g.nz) * sxy + (I)py * nx + (I)px);
        ukm = ldx<T>(a.U, (I)mirror_m(z, g.nz) * sxy + (I)py * nx + (I)px);
    }
    rof_flux_math<DIM, T>(uc, uip, uim, ujp, ujm, ukp, ukm, d1, d2, d3);
}

template <int DIM, int TY, typename T, typename I>
__device__ inline void rof_body(const RofArgs<T>& a, const TileCoord& tc) {
    const Geom& g = a.g;
    constexpr int NT = TX * TY, NH = TX + TY;
    const int lx = threadIdx.x, ly = threadIdx.y, tid = ly * TX + lx;
    const int x = tc.x0 + lx, y = tc.y0 + ly;
    const bool own = x < g.nx && y < g.ny;
    // -x halo column (D2, TY points) and -y halo row (D1, TX points) on the
    // low or high waves depending on the block index
    const int hbase = (tc.lin & 1) ? NT - NH : 0;
    const int h = tid - hbase;
    const bool hlane = h >= 0 && h < NH;
    const bool hcol = h < TY;
    const int hx = hcol ? tc.x0 - 1 : tc.x0 + h - TY;
    const int hy = hcol ? tc.y0 + h : tc.y0 - 1;
    const bool hown = hlane && hx >= 0 && hy >= 0 && hx < g.nx && hy < g.ny;
    const int ox = own ? x : 0, oy = own ? y : 0;
    const int hcx = hown ? hx : 0, hcy = hown ? hy : 0;

    __shared__ T D1l[TY + 1][TX];       // [ly + 1][lx]; row 0 = y0 - 1
    __shared__ T D2l[TY][TX + 1];       // [ly][lx + 1]; column 0 = x0 - 1
    __shared__ double red[8];
    const I sxy = (I)g.sxy, nx = (I)g.nx;

    // D3 at the mirrored z-predecessor of the chunk's first plane
    T d3m = 0;
    if (DIM == 3 && g.nz > 1) {
        T t1, t2;
        rof_fluxes<DIM, I>(a, ox, oy, mirror_m(tc.z0, g.nz), t1, t2, d3m);
    }
    double sdd = 0.0, suu = 0.0;
    for (int z = tc.z0; z < tc.z1; ++z) {
        T d1, d2, d3;
        rof_fluxes<DIM, I>(a, ox, oy, z, d1, d2, d3);
        D1l[ly + 1][lx] = d1;
        D2l[ly][lx + 1] = d2;
        if (hlane) {
            T h1, h2, h3;
            rof_fluxes<DIM, I>(a, hcx, hcy, z, h1, h2, h3);
            if (hown) {
                if (hcol) D2l[h][0] = h2;
                else D1l[0][h - TY] = h1;
            }
        }
        __syncthreads();
        if (own) {
            const I i = (I)z * sxy + (I)y * nx + (I)x;
            const int jm = mirror_m(y, g.ny), im = mirror_m(x, g.nx);
            const T dv1 = d1 - D1l[jm - tc.y0 + 1][lx];
            const T dv2 = d2 - D2l[ly][im - tc.x0 + 1];
            T dv = dv1 + dv2;
            if (DIM == 3) {
                // at z = 0 the mirrored predecessor is plane 1 (d3m from the prologue)
                dv = dv + (d3 - d3m);
            }
            const T u = ldx<T>(a.U, i);
            const T un = u + a.tau * (a.lambda * dv - (u - ldx<T>(a.f, i)));
            if (a.check) {
                const T df = un - u;
                sdd += (double)df * (double)df;
                suu += (double)un * (double)un;
            }
            stnt<T>(a.Uout, i, un);
        }
        __syncthreads();
        d3m = d3;
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
template <int DIM, int TY, typename T, typename I>
__global__ __launch_bounds__(TX * TY) void rof_iter_kernel(RofArgs<T> a) {
    if (solve_stopped(a.ctrl)) return;
    rof_body<DIM, TY, T, I>(a, tile_of_block(a.g, TY));
}

}  // namespace cilreg
