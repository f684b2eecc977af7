// fgp_kernels.hpp -- device code of the FGP iteration: argument block,
// per-tile body and the single-volume kernel.  The body is shared by the
// single-volume kernel (fgp.hip) and the grouped multi-volume kernel
// (batch.hip); the algorithm and its reference citations are described at the
// top of fgp.hip.
#pragma once
#include "runtime.hpp"

namespace cilreg {

template <typename T, typename S>
struct FgpArgs {
    const T* f;
    const S* P[3];      // P_k
    const S* Pp[3];     // P_{k-1}
    S* Pn[3];           // P_{k+1}
    T* Dout;            // D_k (written when write_d)
    const T* Dprev;     // D_{k-1} (read when check)
    T lambda, step, mult;
    int write_d, check, nonneg;
    double* partials;
    const Ctrl* ctrl;
    Geom g;
};

template <int DIM, int TY, typename T, typename S, bool ISO, typename I>
__device__ inline void fgp_body(const FgpArgs<T, S>& a, const TileCoord& tc) {
    const Geom& g = a.g;
    constexpr int NT = TX * TY, NH = TX + TY;
    const int lx = threadIdx.x, ly = threadIdx.y, tid = ly * TX + lx;
    const int x = tc.x0 + lx, y = tc.y0 + ly;
    const bool own = x < g.nx && y < g.ny;
    // +x halo column (TY points) and +y halo row (TX points); the halo lanes
    // alternate between the low and high waves with the block index so
    // co-resident workgroups spread the extra pass over the SIMDs.
    const int hbase = (tc.lin & 1) ? NT - NH : 0;
    const int h = tid - hbase;
    const bool hlane = h >= 0 && h < NH;
    int hx, hy, hlx, hly;
    if (h < TY) { hx = tc.x0 + TX; hy = tc.y0 + h; hlx = TX; hly = h; }
    else { hx = tc.x0 + h - TY; hy = tc.y0 + TY; hlx = h - TY; hly = TY; }
    const bool hown = hlane && hx < g.nx && hy < g.ny;
    // points outside the volume evaluate a valid stand-in and store nothing
    const int ox = own ? x : 0, oy = own ? y : 0;
    const int hcx = hown ? hx : 0, hcy = hown ? hy : 0;

    __shared__ T Ds[2][TY + 1][TX + 1];
    __shared__ double red[8];

    const I sxy = (I)g.sxy, nx = (I)g.nx;
    const T lam = a.lambda, mult = a.mult;
    auto Rv = [&](int c, I i) -> T {
        const T p = ldx<T>(a.P[c], i), q = ldx<T>(a.Pp[c], i);
        return p + mult * (p - q);
    };
    // D at (px, py, z) from R there and at the -x / -y / -z neighbours;
    // neighbour loads are unconditional (clamped) and masked afterwards.
    auto dval = [&](int px, int py, int z, T r3below, T& r1, T& r2, T& r3) -> T {
        const I i = (I)z * sxy + (I)py * nx + (I)px;
        const I im = px > 0 ? i - 1 : i, jm = py > 0 ? i - nx : i;
        r1 = Rv(0, i);
        r2 = Rv(1, i);
        const T b1r = Rv(0, im), b2r = Rv(1, jm);
        const T fv = ldx<T>(a.f, i);
        const T b1 = px > 0 ? b1r : (T)0;
        const T b2 = py > 0 ? b2r : (T)0;
        T d;
        if (DIM == 3) {
            r3 = Rv(2, i);
            const T b3 = z > 0 ? r3below : (T)0;
            d = fv - lam * (r1 + r2 + r3 - b1 - b2 - b3);
        } else {
            r3 = 0;
            d = fv - lam * (r1 + r2 - b1 - b2);
        }
        if (a.nonneg && d < (T)0) d = (T)0;
        return d;
    };

    T rc1 = 0, rc2 = 0, rc3 = 0;   // R at the own point, plane z
    T hr3 = 0;                     // R3 at the halo point, plane z
    T r3b = 0, hr3b = 0;           // R3 one plane below the chunk start
    if (DIM == 3 && tc.z0 > 0) {
        const I base = (I)(tc.z0 - 1) * sxy;
        r3b = Rv(2, base + (I)oy * nx + (I)ox);
        if (hlane) hr3b = Rv(2, base + (I)hcy * nx + (I)hcx);
    }
    Ds[0][ly][lx] = dval(ox, oy, tc.z0, r3b, rc1, rc2, rc3);
    if (hlane) { T t1, t2; const T d = dval(hcx, hcy, tc.z0, hr3b, t1, t2, hr3); if (hown) Ds[0][hly][hlx] = d; }

    double sdd = 0.0, suu = 0.0;
    int cur = 0;
    for (int z = tc.z0; z < tc.z1; ++z) {
        T rn1 = 0, rn2 = 0, rn3 = 0, hr3n = 0;
        if (DIM == 3 && z + 1 < g.nz) {
            Ds[cur ^ 1][ly][lx] = dval(ox, oy, z + 1, rc3, rn1, rn2, rn3);
            if (hlane) {
                T t1, t2;
                const T d = dval(hcx, hcy, z + 1, hr3, t1, t2, hr3n);
                if (hown) Ds[cur ^ 1][hly][hlx] = d;
            }
        }
        __syncthreads();
        if (own) {
            const I i = (I)z * sxy + (I)y * nx + (I)x;
            const T d = Ds[cur][ly][lx];
            const T g1 = (x == g.nx - 1) ? (T)0 : d - Ds[cur][ly][lx + 1];
            const T g2 = (y == g.ny - 1) ? (T)0 : d - Ds[cur][ly + 1][lx];
            T p1 = rc1 + a.step * g1;
            T p2 = rc2 + a.step * g2;
            T p3 = 0;
            if (DIM == 3) {
                const T g3 = (z == g.nz - 1) ? (T)0 : d - Ds[cur ^ 1][ly][lx];
                p3 = rc3 + a.step * g3;
            }
            if (ISO) {
                T n2 = p1 * p1 + p2 * p2;
                if (DIM == 3) n2 = n2 + p3 * p3;
                if (n2 > (T)1) {
                    const T s = (T)1 / dsqrt(n2);
                    p1 = p1 * s;
                    p2 = p2 * s;
                    if (DIM == 3) p3 = p3 * s;
                }
            } else {
                T a1 = fabs(p1), a2 = fabs(p2);
                if (a1 < (T)1) a1 = 1;
                if (a2 < (T)1) a2 = 1;
                p1 = p1 / a1;
                p2 = p2 / a2;
                if (DIM == 3) {
                    T a3 = fabs(p3);
                    if (a3 < (T)1) a3 = 1;
                    p3 = p3 / a3;
                }
            }
            stnt<T>(a.Pn[0], i, p1);
            stnt<T>(a.Pn[1], i, p2);
            if (DIM == 3) stnt<T>(a.Pn[2], i, p3);
            if (a.write_d) stnt<T>(a.Dout, i, d);
            if (a.check) {
                const T df = d - ldx<T>(a.Dprev, i);
                sdd += (double)df * (double)df;
                suu += (double)d * (double)d;
            }
        }
        __syncthreads();
        rc1 = rn1; rc2 = rn2; rc3 = rn3; hr3 = hr3n;
        cur ^= 1;
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
__global__ __launch_bounds__(TX * TY) void fgp_iter_kernel(FgpArgs<T, S> a) {
    if (solve_stopped(a.ctrl)) return;
    fgp_body<DIM, TY, T, S, ISO, I>(a, tile_of_block(a.g, TY));
}

}  // namespace cilreg
