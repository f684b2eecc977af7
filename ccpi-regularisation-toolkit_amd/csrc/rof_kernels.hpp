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

// The flux D1 (along y) from the forward difference fj and the squared
// minmods across x (mi) and, in 3D, z (mk): ROF_TV_core.c:131-137 (3D, double
// sqrt), :157-161 (2D).  eps is the floor added in double under the square
// root: 1e-8 for ROF-TV, 1e-12 for the ROF part of LLT-ROF
// (LLT_ROF_core.c:22,198-363).
template <int DIM, typename T>
__device__ inline T rof_d1(T fj, T mi, T mk, double eps) {
    if (DIM == 3) return fj / Mixed<T>::rof_sqrt_d(fj * fj + mi + mk, eps);
    return fj / Mixed<T>::rof_sqrt_f(fj * fj + mi, eps);
}

// The flux D2 (along x): ROF_TV_core.c:194-200 (3D), :220-224 (2D).
template <int DIM, typename T>
__device__ inline T rof_d2(T fi, T mj, T mk, double eps) {
    if (DIM == 3) return fi / Mixed<T>::rof_sqrt_f(fi * fi + mj + mk, eps);
    return fi / Mixed<T>::rof_sqrt_f(fi * fi + mj, eps);
}

// The fluxes D1, D2 and, in 3D, D3 (along z, ROF_TV_core.c:257-263) at a
// point from u there (uc) and at its mirrored neighbours.
template <int DIM, typename T>
__device__ inline void rof_flux_math(T uc, T uip, T uim, T ujp, T ujm, T ukp, T ukm, T& d1, T& d2, T& d3,
                                     double eps = 1.0e-8) {
    const T fi = uip - uc, bi = uc - uim, fj = ujp - uc, bj = uc - ujm;
    if (DIM == 3) {
        const T fk = ukp - uc, bk = uc - ukm;
        T m[3];
        Mixed<T>::minmod_sq_n({fi, fj, fk}, {bi, bj, bk}, m);
        const T mi = m[0], mj = m[1], mk = m[2];
        d1 = rof_d1<DIM, T>(fj, mi, mk, eps);
        d2 = rof_d2<DIM, T>(fi, mj, mk, eps);
        d3 = fk / Mixed<T>::rof_sqrt_f(fk * fk + mj + mi, eps);
    } else {
        T m[2];
        Mixed<T>::minmod_sq_n({fi, fj}, {bi, bj}, m);
        d1 = rof_d1<DIM, T>(fj, m[0], T(0), eps);
        d2 = rof_d2<DIM, T>(fi, m[1], T(0), eps);
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
    constexpr int NT = TX * TY;
    // U plane ring: rows y0-1 and y0+TY on [x0-1, x0+64] (corners included),
    // columns x0-1 and x0+64 on the tile's rows
    constexpr int NR = 2 * (TX + 2) + 2 * TY;
    static_assert(NR <= NT, "ring lanes must fit the workgroup");
    const int lx = threadIdx.x, ly = threadIdx.y, tid = ly * TX + lx;
    const int x = tc.x0 + lx, y = tc.y0 + ly;
    const bool own = x < g.nx && y < g.ny;
    const int ox = own ? x : 0, oy = own ? y : 0;
    // ring lane: one point of the U ring, on the low or high waves by block index
    const int hbase = (tc.lin & 1) ? NT - NR : 0;
    const int h = tid - hbase;
    const bool rlane = h >= 0 && h < NR;
    int sx, sy;   // ring slot in the (TY+2) x (TX+2) plane with origin (x0-1, y0-1)
    if (h < TX + 2) { sx = h; sy = 0; }
    else if (h < 2 * (TX + 2)) { sx = h - (TX + 2); sy = TY + 1; }
    else if (h < 2 * (TX + 2) + TY) { sx = 0; sy = 1 + h - 2 * (TX + 2); }
    else { sx = TX + 1; sy = 1 + h - 2 * (TX + 2) - TY; }
    const int rx = tc.x0 + sx - 1, ry = tc.y0 + sy - 1;
    const bool rin = rlane && rx >= 0 && ry >= 0 && rx < g.nx && ry < g.ny;
    const int rcx = rin ? rx : 0, rcy = rin ? ry : 0;   // clamped stand-in
    // ring lanes that also evaluate a flux: D1 on row y0-1 over x0..x0+63 and
    // D2 on column x0-1 over the tile's rows
    const bool d1lane = rlane && sy == 0 && sx >= 1 && sx <= TX;
    const bool d2lane = rlane && sx == 0 && sy >= 1 && sy <= TY;

    __shared__ T Us[TY + 2][TX + 2];    // U of plane z: tile and ring
    __shared__ T D1l[TY + 1][TX];       // [ly + 1][lx]; row 0 = y0 - 1
    __shared__ T D2l[TY][TX + 1];       // [ly][lx + 1]; column 0 = x0 - 1
    __shared__ double red[kRedSlots];
    const I sxy = (I)g.sxy, nx = (I)g.nx;

    // mirrored neighbour slots (loop-invariant) of the own point; threads
    // outside the volume read their own slot (their results are not stored)
    const int oxp = own ? mirror_p(x, g.nx) - tc.x0 + 1 : lx + 1, oxm = own ? mirror_m(x, g.nx) - tc.x0 + 1 : lx + 1;
    const int oyp = own ? mirror_p(y, g.ny) - tc.y0 + 1 : ly + 1, oym = own ? mirror_m(y, g.ny) - tc.y0 + 1 : ly + 1;
    // flux lanes: the neighbours a D1 on row y0-1 / a D2 on column x0-1 reads
    const int rxp = mirror_p(rcx, g.nx) - tc.x0 + 1, rxm = mirror_m(rcx, g.nx) - tc.x0 + 1;
    const int ryp = mirror_p(rcy, g.ny) - tc.y0 + 1, rym = mirror_m(rcy, g.ny) - tc.y0 + 1;

    const I ocol = (I)oy * nx + (I)ox, rcol = (I)rcy * nx + (I)rcx;
    // z pipeline of the own column and of the ring point
    T uc = ldx<T>(a.U, (I)tc.z0 * sxy + ocol), um = 0;
    T hc = 0, hm = 0;
    if (rlane) hc = ldx<T>(a.U, (I)tc.z0 * sxy + rcol);
    // D3 at the mirrored z-predecessor of the chunk's first plane
    T d3m = 0;
    if (DIM == 3) {
        const int zm = mirror_m(tc.z0, g.nz);
        um = ldx<T>(a.U, (I)zm * sxy + ocol);
        if (rlane) hm = ldx<T>(a.U, (I)zm * sxy + rcol);
        if (g.nz > 1) {
            T t1, t2;
            rof_fluxes<DIM, I>(a, ox, oy, zm, t1, t2, d3m);
        }
    }
    double sdd = 0.0, suu = 0.0;
    for (int z = tc.z0; z < tc.z1; ++z) {
        T up = 0, hp = 0;
        if (DIM == 3) {
            const int zp = mirror_p(z, g.nz);
            up = (zp == z + 1) ? ldx<T>(a.U, (I)zp * sxy + ocol) : um;
            if (rlane) hp = (zp == z + 1) ? ldx<T>(a.U, (I)zp * sxy + rcol) : hm;
        }
        Us[ly + 1][lx + 1] = uc;
        if (rlane) Us[sy][sx] = hc;
        __syncthreads();
        // fluxes of the own point
        T d1, d2, d3;
        rof_flux_math<DIM, T>(uc, Us[ly + 1][oxp], Us[ly + 1][oxm], Us[oyp][lx + 1], Us[oym][lx + 1], up, um, d1, d2,
                              d3);
        D1l[ly + 1][lx] = d1;
        D2l[ly][lx + 1] = d2;
        if (d1lane || d2lane) {
            // the neighbour outside the ring (y0-2 for D1, x0-2 for D2) feeds
            // only the other fluxes; it is replaced by the point itself
            const T uxp = Us[sy][rxp >= 0 && rxp < TX + 2 ? rxp : sx];
            const T uxm = Us[sy][rxm >= 0 && rxm < TX + 2 ? rxm : sx];
            const T uyp = Us[ryp >= 0 && ryp < TY + 2 ? ryp : sy][sx];
            const T uym = Us[rym >= 0 && rym < TY + 2 ? rym : sy][sx];
            // only the flux the lane stores: D2 on the column, D1 on the row
            // (the across-difference pair: y for D2, x for D1)
            const T fa = d2lane ? uyp - hc : uxp - hc, ba = d2lane ? hc - uym : hc - uxm;
            T m[2];
            Mixed<T>::minmod_sq_n({fa, DIM == 3 ? hp - hc : T(0)}, {ba, DIM == 3 ? hc - hm : T(0)}, m);
            const T mk = DIM == 3 ? m[1] : T(0);
            if (d2lane) {
                const T h2 = rof_d2<DIM, T>(uxp - hc, m[0], mk, 1.0e-8);
                if (rin) D2l[sy - 1][0] = h2;
            } else {
                const T h1 = rof_d1<DIM, T>(uyp - hc, m[0], mk, 1.0e-8);
                if (rin) D1l[0][sx - 1] = h1;
            }
        }
        __syncthreads();
        if (own) {
            const I i = (I)z * sxy + ocol;
            const T dv1 = d1 - D1l[oym][lx];
            const T dv2 = d2 - D2l[ly][oxm];
            T dv = dv1 + dv2;
            if (DIM == 3) {
                // at z = 0 the mirrored predecessor is plane 1 (d3m from the prologue)
                dv = dv + (d3 - d3m);
            }
            const T un = uc + a.tau * (a.lambda * dv - (uc - ldx<T>(a.f, i)));
            if (a.check) {
                const T df = un - uc;
                sdd += (double)df * (double)df;
                suu += (double)un * (double)un;
            }
            stnt<T>(a.Uout, i, un);
        }
        d3m = d3;
        um = uc;
        uc = up;
        hm = hc;
        hc = hp;
    }
    if (a.check) {
        block_sum2(sdd, suu, red);
        if (tid == 0) {
            a.partials[2 * tc.lin] = sdd;
            a.partials[2 * tc.lin + 1] = suu;
        }
    }
}

// ---------------------------------------------------------------------------
// Two iterations per launch on a 2D image (k -> k+2)
// ---------------------------------------------------------------------------
// A small image is launch-bound at one iteration per launch (512^2: ~1 us of
// arithmetic in ~4.4 us per iteration).  One iteration's update at p reads u
// on p's 3x3 neighbourhood (D1 at p and its mirrored -y neighbour, D2 at p
// and its mirrored -x neighbour, each from u at its +-1 neighbours), so the
// pass evaluates, for the 64 x TY tile at (x0, y0):
//   A  u_k             on [x0-2, x0+65] x [y0-2, y0+TY+1]   (loaded)
//   F  D1, D2 of k     on [x0-2, x0+64] x [y0-2, y0+TY]
//   B  u_{k+1}         on [x0-1, x0+64] x [y0-1, y0+TY]
//   G  D1, D2 of k+1   on [x0-1, x0+63] x [y0-1, y0+TY-1]
//   T  u_{k+2}         on the tile (stored)
// through LDS, every value by the expressions of rof_iter_kernel (same
// fluxes, mirrors and operation order), so the results are bit-identical to
// two one-iteration launches.  Region points outside the image take their
// own slot for every neighbour (their values are never read by a point
// inside it: mirrored neighbours of an image point lie in the image).
template <typename T>
struct Rof2Args {
    const T* f;
    const T* U;
    T* Uout;
    T lambda, tau;
    int check0, check1;
    double* partials0;
    double* partials1;
    const Ctrl* ctrl;
    Geom g;
};

template <int TY>
struct Rof2Shape {
    static constexpr int AW = TX + 4, AH = TY + 4;   // A, origin (x0-2, y0-2)
    static constexpr int FW = TX + 3, FH = TY + 3;   // F, origin (x0-2, y0-2)
    static constexpr int BW = TX + 2, BH = TY + 2;   // B, origin (x0-1, y0-1)
    static constexpr int GW = TX + 1, GH = TY + 1;   // G, origin (x0-1, y0-1)
};

// The slot (in a region of width w, height h, origin (ox, oy)) of the
// mirrored +/-1 neighbours of image point (px, py); a point outside the image
// or a neighbour outside the region gives the point's own slot.
__device__ inline void rof2_nbr_slots(int px, int py, bool in, const Geom& g, int ox, int oy, int w, int h, int& sxp,
                                      int& sxm, int& syp, int& sym) {
    const int sx = px - ox, sy = py - oy;
    sxp = sxm = sx;
    syp = sym = sy;
    if (in) {
        sxp = mirror_p(px, g.nx) - ox;
        sxm = mirror_m(px, g.nx) - ox;
        syp = mirror_p(py, g.ny) - oy;
        sym = mirror_m(py, g.ny) - oy;
        if (sxp < 0 || sxp >= w) sxp = sx;
        if (sxm < 0 || sxm >= w) sxm = sx;
        if (syp < 0 || syp >= h) syp = sy;
        if (sym < 0 || sym >= h) sym = sy;
    }
}

template <int TY, typename T, typename I>
__device__ inline void rof2_2d_body(const Rof2Args<T>& a, const TileCoord& tc) {
    using Sh = Rof2Shape<TY>;
    constexpr int NT = TX * TY;
    const Geom& g = a.g;
    const int tid = threadIdx.y * TX + threadIdx.x;
    const int x0 = tc.x0, y0 = tc.y0;
    const I nx = (I)g.nx;
    __shared__ T UA[Sh::AH][Sh::AW];
    __shared__ T F1[Sh::FH][Sh::FW], F2[Sh::FH][Sh::FW];
    __shared__ T UB[Sh::BH][Sh::BW];
    __shared__ T G1[Sh::GH][Sh::GW], G2[Sh::GH][Sh::GW];
    __shared__ double red[kRedSlots];

    // A: u_k (points outside the image load a clamped stand-in)
    for (int r = tid; r < Sh::AW * Sh::AH; r += NT) {
        const int ry = r / Sh::AW, rx = r - ry * Sh::AW;
        const int px = x0 - 2 + rx, py = y0 - 2 + ry;
        const int cx = px < 0 ? 0 : (px >= g.nx ? g.nx - 1 : px), cy = py < 0 ? 0 : (py >= g.ny ? g.ny - 1 : py);
        UA[ry][rx] = ldx<T>(a.U, (I)cy * nx + (I)cx);
    }
    __syncthreads();
    // F: fluxes of iteration k
    for (int r = tid; r < Sh::FW * Sh::FH; r += NT) {
        const int ry = r / Sh::FW, rx = r - ry * Sh::FW;
        const int px = x0 - 2 + rx, py = y0 - 2 + ry;   // A slot = (rx, ry)
        const bool in = px >= 0 && py >= 0 && px < g.nx && py < g.ny;
        int sxp, sxm, syp, sym;
        rof2_nbr_slots(px, py, in, g, x0 - 2, y0 - 2, Sh::AW, Sh::AH, sxp, sxm, syp, sym);
        T d1, d2, d3;
        rof_flux_math<2, T>(UA[ry][rx], UA[ry][sxp], UA[ry][sxm], UA[syp][rx], UA[sym][rx], T(0), T(0), d1, d2, d3);
        F1[ry][rx] = d1;
        F2[ry][rx] = d2;
    }
    __syncthreads();
    // B: u_{k+1}; the residual of iteration k on the tile
    double sdd0 = 0.0, suu0 = 0.0, sdd1 = 0.0, suu1 = 0.0;
    for (int r = tid; r < Sh::BW * Sh::BH; r += NT) {
        const int ry = r / Sh::BW, rx = r - ry * Sh::BW;
        const int px = x0 - 1 + rx, py = y0 - 1 + ry;   // F / A slot = (rx + 1, ry + 1)
        const bool in = px >= 0 && py >= 0 && px < g.nx && py < g.ny;
        // the mirrored -x / -y neighbours in F's slots
        int fxm = rx + 1, fym = ry + 1;
        if (in) {
            fxm = mirror_m(px, g.nx) - (x0 - 2);
            fym = mirror_m(py, g.ny) - (y0 - 2);
        }
        const T uc = UA[ry + 1][rx + 1];
        const T dv = (F1[ry + 1][rx + 1] - F1[fym][rx + 1]) + (F2[ry + 1][rx + 1] - F2[ry + 1][fxm]);
        const I i = (I)(in ? py : 0) * nx + (I)(in ? px : 0);
        const T un = uc + a.tau * (a.lambda * dv - (uc - ldx<T>(a.f, i)));
        UB[ry][rx] = un;
        if (a.check0 && in && rx >= 1 && rx <= TX && ry >= 1 && ry <= TY) {
            const T df = un - uc;
            sdd0 += (double)df * (double)df;
            suu0 += (double)un * (double)un;
        }
    }
    __syncthreads();
    // G: fluxes of iteration k+1
    for (int r = tid; r < Sh::GW * Sh::GH; r += NT) {
        const int ry = r / Sh::GW, rx = r - ry * Sh::GW;
        const int px = x0 - 1 + rx, py = y0 - 1 + ry;   // B slot = (rx, ry)
        const bool in = px >= 0 && py >= 0 && px < g.nx && py < g.ny;
        int sxp, sxm, syp, sym;
        rof2_nbr_slots(px, py, in, g, x0 - 1, y0 - 1, Sh::BW, Sh::BH, sxp, sxm, syp, sym);
        T d1, d2, d3;
        rof_flux_math<2, T>(UB[ry][rx], UB[ry][sxp], UB[ry][sxm], UB[syp][rx], UB[sym][rx], T(0), T(0), d1, d2, d3);
        G1[ry][rx] = d1;
        G2[ry][rx] = d2;
    }
    __syncthreads();
    // T: u_{k+2}
    {
        const int lx = threadIdx.x, ly = threadIdx.y;
        const int px = x0 + lx, py = y0 + ly;
        if (px < g.nx && py < g.ny) {
            const int gxm = mirror_m(px, g.nx) - (x0 - 1), gym = mirror_m(py, g.ny) - (y0 - 1);
            const T uc = UB[ly + 1][lx + 1];
            const T dv = (G1[ly + 1][lx + 1] - G1[gym][lx + 1]) + (G2[ly + 1][lx + 1] - G2[ly + 1][gxm]);
            const I i = (I)py * nx + (I)px;
            const T un = uc + a.tau * (a.lambda * dv - (uc - ldx<T>(a.f, i)));
            stnt<T>(a.Uout, i, un);
            if (a.check1) {
                const T df = un - uc;
                sdd1 += (double)df * (double)df;
                suu1 += (double)un * (double)un;
            }
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

template <int TY, typename T, typename I>
__global__ __launch_bounds__(TX * TY) void rof2_2d_kernel(Rof2Args<T> a) {
    if (solve_stopped(a.ctrl)) return;
    rof2_2d_body<TY, T, I>(a, tile_of_block(a.g, TY));
}

// One iteration on one volume: the workgroup's tile comes from its block index.
template <int DIM, int TY, typename T, typename I>
__global__ __launch_bounds__(TX * TY) void rof_iter_kernel(RofArgs<T> a) {
    if (solve_stopped(a.ctrl)) return;
    rof_body<DIM, TY, T, I>(a, tile_of_block(a.g, TY));
}

}  // namespace cilreg
