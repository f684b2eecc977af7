// sb.hip -- SB-TV (split Bregman), one fused kernel per iteration (two
// kernels for extents whose last tile is one column or row wide).
//
// Reference algorithm: src/Core/regularisers_CPU/SB_TV_core.c:38-371.  With
// mu' = 1/mu and lam = 2 mu', per iteration:
//   U1 = J(U0), U2 = J(U1)   two Jacobi sweeps (named gauss_seidel in the
//                            reference, :149-173, :254-282) with mirror
//                            boundaries, reading D and B at x and x - e_d
//   v_d = (U2(x + e_d) - U2(x)) + B_d      forward difference, mirrored at the
//                                          last index
//   D = shrink(v, 1/lam)     iso: vector shrinkage; aniso: per component
//   B = B + (U2(x + e_d) - U2(x)) - D      (:233-248, :353-371)
// The residual compares U2 with U1 (:123-133).
//
// Two-kernel path: kernel 1 runs the first sweep.  Kernel 2 runs the second sweep on the tile
// plus its +x / +y halo into an LDS plane (double-buffered across z), then the
// shrinkage and Bregman update of the tile.  HBM traffic per voxel-iteration:
// sweep 1 reads U0, f, D(3), B(3) and writes U1 (36 B); kernel 2 reads U1, f,
// D, B and writes U2, D, B (60 B).
#include "solvers.hpp"
#include "state.hpp"

namespace cilreg {

template <typename T>
struct SbArgs {
    const T* f;
    const T* Uin;      // sweep input
    T* Uout;           // sweep output (kernel 1) or U2 (kernel 2)
    const T* D[3];
    const T* B[3];
    T* Dout[3];
    T* Bout[3];
    T lam, mu, nc, thr;   // lam = 2/mu_in, mu = 1/mu_in, nc = 1/(mu + (4|6) lam), thr = 1/lam
    int check;
    double* partials;
    const Ctrl* ctrl;
    Geom g;
};

__device__ inline int sb_p(int i, int n) { return i + 1 == n ? (i - 1 >= 0 ? i - 1 : 0) : i + 1; }
__device__ inline int sb_m(int i, int n) { return i - 1 < 0 ? (i + 1 < n ? i + 1 : 0) : i - 1; }

// one Jacobi sweep value at (px, py, z) from a.Uin
template <int DIM, typename I, typename T>
__device__ inline T sb_sweep_at(const SbArgs<T>& a, int px, int py, int z) {
    const Geom& g = a.g;
    const I sxy = (I)g.sxy, nx = (I)g.nx;
    const I zo = (I)z * sxy;
    const I c = zo + (I)py * nx + (I)px;
    const int ip = sb_p(px, g.nx), im = sb_m(px, g.nx), jp = sb_p(py, g.ny), jm = sb_m(py, g.ny);
    const I cim = zo + (I)py * nx + (I)im, cjm = zo + (I)jm * nx + (I)px;
    const I cip = zo + (I)py * nx + (I)ip, cjp = zo + (I)jp * nx + (I)px;
    const T* U = a.Uin;
    auto L = [](const T* p, I e) { return ldx<T>(p, e); };
    T acc;
    if (DIM == 2) {
        acc = L(a.D[0], cim) - L(a.D[0], c) + L(a.D[1], cjm) - L(a.D[1], c) - L(a.B[0], cim) + L(a.B[0], c) -
              L(a.B[1], cjm) + L(a.B[1], c);
        acc += L(U, cip) + L(U, cim) + L(U, cjp) + L(U, cjm);
    } else {
        const int kp = sb_p(z, g.nz), km = sb_m(z, g.nz);
        const I ckm = (I)km * sxy + (I)py * nx + (I)px, ckp = (I)kp * sxy + (I)py * nx + (I)px;
        const T dsum = L(a.D[0], cim) - L(a.D[0], c) + L(a.D[1], cjm) - L(a.D[1], c) + L(a.D[2], ckm) - L(a.D[2], c);
        const T bsum = -L(a.B[0], cim) + L(a.B[0], c) - L(a.B[1], cjm) + L(a.B[1], c) - L(a.B[2], ckm) + L(a.B[2], c);
        acc = dsum + bsum;
        acc += L(U, cip) + L(U, cim) + L(U, cjp) + L(U, cjm) + L(U, ckp) + L(U, ckm);
    }
    acc *= a.lam;
    acc += a.mu * L(a.f, c);
    return a.nc * acc;
}

template <int DIM, int TY, typename T, typename I>
__global__ __launch_bounds__(TX * TY) void sb_sweep_kernel(SbArgs<T> a) {
    if (solve_stopped(a.ctrl)) return;
    const Geom& g = a.g;
    const TileCoord tc = tile_of_block(g, TY);
    const int x = tc.x0 + (int)threadIdx.x, y = tc.y0 + (int)threadIdx.y;
    if (x >= g.nx || y >= g.ny) return;
    for (int z = tc.z0; z < tc.z1; ++z)
        stx<T>(a.Uout, (I)z * (I)g.sxy + (I)y * (I)g.nx + (I)x, sb_sweep_at<DIM, I>(a, x, y, z));
}

template <int DIM, int TY, typename T, bool ANISO, typename I>
__global__ __launch_bounds__(TX * TY) void sb_update_kernel(SbArgs<T> a) {
    if (solve_stopped(a.ctrl)) return;
    const Geom& g = a.g;
    const TileCoord tc = tile_of_block(g, TY);
    constexpr int NT = TX * TY, NH = TX + TY;
    const int lx = threadIdx.x, ly = threadIdx.y, tid = ly * TX + lx;
    const int x = tc.x0 + lx, y = tc.y0 + ly;
    const bool own = x < g.nx && y < g.ny;
    // +x halo column and +y halo row, on the low or high waves by block index
    const int hbase = (tc.lin & 1) ? NT - NH : 0;
    const int h = tid - hbase;
    const bool hlane = h >= 0 && h < NH;
    int hx, hy, hlx, hly;
    if (h < TY) { hx = tc.x0 + TX; hy = tc.y0 + h; hlx = TX; hly = h; }
    else { hx = tc.x0 + h - TY; hy = tc.y0 + TY; hlx = h - TY; hly = TY; }
    const bool hown = hlane && hx < g.nx && hy < g.ny;
    const int ox = own ? x : 0, oy = own ? y : 0;
    const int hcx = hown ? hx : 0, hcy = hown ? hy : 0;

    __shared__ T Ul[2][TY + 1][TX + 1];
    __shared__ double red[kRedSlots];
    const I sxy = (I)g.sxy, nx = (I)g.nx;

    T below = 0;   // U2 at the own point, plane z-1
    if (DIM == 3 && tc.z0 > 0) below = sb_sweep_at<DIM, I>(a, ox, oy, tc.z0 - 1);
    Ul[0][ly][lx] = sb_sweep_at<DIM, I>(a, ox, oy, tc.z0);
    if (hlane) { const T v = sb_sweep_at<DIM, I>(a, hcx, hcy, tc.z0); if (hown) Ul[0][hly][hlx] = v; }

    double sdd = 0.0, suu = 0.0;
    int cur = 0;
    for (int z = tc.z0; z < tc.z1; ++z) {
        if (DIM == 3 && z + 1 < g.nz) {
            Ul[cur ^ 1][ly][lx] = sb_sweep_at<DIM, I>(a, ox, oy, z + 1);
            if (hlane) {
                const T v = sb_sweep_at<DIM, I>(a, hcx, hcy, z + 1);
                if (hown) Ul[cur ^ 1][hly][hlx] = v;
            }
        }
        __syncthreads();
        T u2 = 0;
        if (own) {
            const I i = (I)z * sxy + (I)y * nx + (I)x;
            u2 = Ul[cur][ly][lx];
            // mirrored forward neighbours: +1, or -1 at the last index
            T ux, uy;
            if (x + 1 < g.nx) ux = Ul[cur][ly][lx + 1];
            else if (g.nx == 1) ux = u2;
            else ux = (lx > 0) ? Ul[cur][ly][lx - 1] : sb_sweep_at<DIM, I>(a, x - 1, y, z);
            if (y + 1 < g.ny) uy = Ul[cur][ly + 1][lx];
            else if (g.ny == 1) uy = u2;
            else uy = (ly > 0) ? Ul[cur][ly - 1][lx] : sb_sweep_at<DIM, I>(a, x, y - 1, z);
            T gr[3], v[3], dn[3], bv[3];
            gr[0] = ux - u2;
            gr[1] = uy - u2;
            gr[2] = 0;
            if (DIM == 3) {
                const T uz = (z + 1 < g.nz) ? Ul[cur ^ 1][ly][lx] : below;
                gr[2] = uz - u2;
            }
            #pragma unroll
            for (int d = 0; d < DIM; ++d) {
                bv[d] = ldx<T>(a.B[d], i);
                v[d] = gr[d] + bv[d];
            }
            if (ANISO) {
                #pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    T s = fabs(v[d]) - a.thr;
                    if (s < (T)0) s = 0;
                    dn[d] = (v[d] != (T)0) ? (v[d] / fabs(v[d])) * s : (T)0;
                }
            } else {
                T n2 = v[0] * v[0] + v[1] * v[1];
                if (DIM == 3) n2 = n2 + v[2] * v[2];
                const T nrm = dsqrt(n2);
                T s = nrm - a.thr;
                if (s < (T)0) s = 0;
                #pragma unroll
                for (int d = 0; d < DIM; ++d) dn[d] = (nrm != (T)0) ? s * (v[d] / nrm) : (T)0;
            }
            #pragma unroll
            for (int d = 0; d < DIM; ++d) {
                stnt<T>(a.Dout[d], i, dn[d]);
                stnt<T>(a.Bout[d], i, bv[d] + (gr[d] - dn[d]));
            }
            stnt<T>(a.Uout, i, u2);
            if (a.check) {
                const T u1 = ldx<T>(a.Uin, i);
                const T df = u2 - u1;
                sdd += (double)df * (double)df;
                suu += (double)u2 * (double)u2;
            }
        }
        __syncthreads();
        below = u2;
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

// ---------------------------------------------------------------------------
// Fused pass: both sweeps, the shrinkage and the Bregman update in one kernel
// ---------------------------------------------------------------------------
// The D/B part of a sweep value, sum_d (D_d(m-) - D_d - B_d(m-) + B_d) in the
// reference's order, depends only on the old D and B, so both sweeps at a
// point share it: U_new = nc ((db + sum of six U neighbours) lam + mu f).
// The workgroup's region is the 67 x (TY+3) ring region of the FGP
// two-iteration kernels (slot 0 the own tile point, slot 1 a ring point).
//   S1  U1(zi) = J(U0) on the region, into a 4-plane LDS ring; db and f kept.
//       The plane's U0, D_1, B_1, D_2, B_2 go through LDS for the x / y
//       neighbours, U0 above / below and D_3, B_3 below ride in registers
//   S2  U2(zi-1) = J(U1) on the tile and its +x column / +y row (from LDS)
//   S3  shrink + Bregman on the tile at zi-2: forward differences of U2 from
//       LDS (mirrored at the last index), write U2, D, B
// HBM traffic per voxel-iteration: read U0, f, D(3), B(3); write U2, D(3),
// B(3) = 60 B (the two-kernel path moves 96 B).  The mirrored neighbour of
// the last index lies in the region unless the last tile is one column (row)
// wide; those extents use the two-kernel path.

// the D/B part of the sweep at element c with mirrored -x / -y / -z elements
template <int DIM, typename I, typename T>
__device__ inline T sb_db(const SbArgs<T>& a, I c, I cim, I cjm, I ckm) {
    auto L = [](const T* p, I e) { return ldx<T>(p, e); };
    if (DIM == 2)
        return L(a.D[0], cim) - L(a.D[0], c) + L(a.D[1], cjm) - L(a.D[1], c) - L(a.B[0], cim) + L(a.B[0], c) -
               L(a.B[1], cjm) + L(a.B[1], c);
    const T dsum = L(a.D[0], cim) - L(a.D[0], c) + L(a.D[1], cjm) - L(a.D[1], c) + L(a.D[2], ckm) - L(a.D[2], c);
    const T bsum = -L(a.B[0], cim) + L(a.B[0], c) - L(a.B[1], cjm) + L(a.B[1], c) - L(a.B[2], ckm) + L(a.B[2], c);
    return dsum + bsum;
}

template <typename T>
__device__ inline T sb_combine(const SbArgs<T>& a, T db, T usum, T fv) {
    T acc = db;
    acc += usum;
    acc *= a.lam;
    acc += a.mu * fv;
    return a.nc * acc;
}

// Ring point h (0 <= h < 3 (TX+3) + 3 TY) of the region, in region coordinates.
template <int TY>
__device__ inline void sb_ring_point(int h, int& rx, int& ry) {
    constexpr int BW = TX + 3;
    if (h < 3 * BW) {
        const int row = h / BW;
        rx = h - row * BW;
        ry = row == 0 ? 0 : TY + row;          // rows y0-1, y0+TY, y0+TY+1
    } else {
        const int k = h - 3 * BW, col = k / TY;
        rx = col == 0 ? 0 : TX + col;          // columns x0-1, x0+TX, x0+TX+1
        ry = k - col * TY + 1;
    }
}

// RW = 0: TX x TY threads, each the own tile point (slot 0) and NR of them
// also a ring point (slot 1, on the low or high waves by block index).
// RW > 0: TX x TY + 64 RW threads, the first TX x TY the tile, the RW waves
// after them the ring, one point per thread.
template <int TY, int RW>
constexpr int sb_fused_threads() { return TX * TY + 64 * RW; }

template <int DIM, int TY, typename T, bool ANISO, typename I, int RW = 0>
__global__ __launch_bounds__((sb_fused_threads<TY, RW>())) void sb_fused_kernel(SbArgs<T> a) {
    if (solve_stopped(a.ctrl)) return;
    const Geom& g = a.g;
    const TileCoord tc = tile_of_block(g, TY);
    constexpr int BW = TX + 3, BH = TY + 3, NR = 3 * BW + 3 * TY, NT = TX * TY;
    constexpr int NS = RW ? 1 : 2;   // slots per thread
    static_assert(RW == 0 || 64 * RW >= NR, "ring waves must hold the ring");
    const int lx = threadIdx.x, ly = threadIdx.y, tid = ly * TX + lx;

    int bx[2] = {0, 0}, by[2] = {0, 0};
    int nslot;
    bool tile0;   // slot 0 is a tile point
    if (RW) {
        tile0 = tid < NT;
        const int h = tid - NT;
        const bool ring = !tile0 && h < NR;
        int rx = 0, ry = 0;
        if (ring) sb_ring_point<TY>(h, rx, ry);
        bx[0] = tile0 ? lx + 1 : rx;
        by[0] = tile0 ? ly + 1 : ry;
        nslot = (tile0 || ring) ? 1 : 0;
    } else {
        tile0 = true;
        bx[0] = lx + 1;
        by[0] = ly + 1;
        const int hbase = (tc.lin & 1) ? NT - NR : 0;
        const int h = tid - hbase;
        const bool ring = h >= 0 && h < NR;
        int rx = 0, ry = 0;
        if (ring) sb_ring_point<TY>(h, rx, ry);
        bx[1] = rx;
        by[1] = ry;
        nslot = ring ? 2 : 1;
    }
    int X[2], Y[2], cx[2], cy[2];
    bool in[2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        X[s] = tc.x0 + bx[s] - 1;
        Y[s] = tc.y0 + by[s] - 1;
        in[s] = X[s] >= 0 && Y[s] >= 0 && X[s] < g.nx && Y[s] < g.ny;
        cx[s] = in[s] ? X[s] : 0;   // clamped stand-in
        cy[s] = in[s] ? Y[s] : 0;
    }
    const bool own = tile0 && in[0];
    // the ring slot also evaluates U2 on the +x column and the +y row
    const int rs = RW ? 0 : 1;
    const bool u2halo = nslot > rs && !(RW && tile0) && in[rs] &&
                        ((by[rs] == TY + 1 && bx[rs] >= 1 && bx[rs] <= TX) ||
                         (bx[rs] == TX + 1 && by[rs] >= 1 && by[rs] <= TY));
    // slots that evaluate U2 (points inside the volume: their mirrored
    // neighbours are region slots)
    bool u2ok[2];
    u2ok[0] = RW ? (tile0 ? own : u2halo) : own;
    u2ok[1] = RW ? false : u2halo;
    // mirrored x / y neighbour slots (SB's rule: +1, or -1 at the last index)
    int sxp[2], sxm[2], syp[2], sym[2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        sxp[s] = sb_p(cx[s], g.nx) - tc.x0 + 1;
        sxm[s] = sb_m(cx[s], g.nx) - tc.x0 + 1;
        syp[s] = sb_p(cy[s], g.ny) - tc.y0 + 1;
        sym[s] = sb_m(cy[s], g.ny) - tc.y0 + 1;
    }

    __shared__ T U1p[4][BH][BW];   // U1 planes by z mod 4
    __shared__ T U2p[2][BH][BW];   // U2 planes by z mod 2
    __shared__ T U0s[BH][BW];      // U0, D_1, B_1, D_2, B_2 of plane zi (x / y neighbours of S1)
    __shared__ T D0s[BH][BW], B0s[BH][BW], D1s[BH][BW], B1s[BH][BW];
    __shared__ double red[kRedSlots];
    const I sxy = (I)g.sxy, nx = (I)g.nx;
    const int nz = g.nz, z0 = tc.z0, z1 = tc.z1;

    // per-slot z pipeline of U0 and of D_3, B_3 below (mirrored at the ends)
    T u0m[2] = {}, u0c[2] = {}, d2m[2] = {}, b2m[2] = {};
    const int zstart = z0 >= 2 ? z0 - 2 : 0;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        if (s < nslot) {
            const I col = (I)cy[s] * nx + (I)cx[s];
            u0c[s] = ldx<T>(a.Uin, (I)zstart * sxy + col);
            if (DIM == 3) {
                const I cm = (I)sb_m(zstart, nz) * sxy + col;
                u0m[s] = ldx<T>(a.Uin, cm);
                d2m[s] = ldx<T>(a.D[2], cm);
                b2m[s] = ldx<T>(a.B[2], cm);
            }
        }
    }
    T dbz[2] = {}, fz[2] = {};     // db and f of plane zi - 1
    T bq1[3] = {}, bq2[3] = {};    // B of the own point at planes zi - 1, zi - 2 (for S3)
    double sdd = 0.0, suu = 0.0;
    for (int zi = zstart; zi <= z1 + 1; ++zi) {
        // ---- S1a: own values of plane zi into LDS; U0 of the plane above
        T dbn[2] = {}, fn[2] = {}, u0p[2] = {}, dv[2][3] = {}, bv[2][3] = {};
        const bool s1 = zi < nz;
        if (s1) {
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                if (s < nslot) {
                    const I col = (I)cy[s] * nx + (I)cx[s];
                    const I c = (I)zi * sxy + col;
#pragma unroll
                    for (int d = 0; d < DIM; ++d) {
                        dv[s][d] = ldx<T>(a.D[d], c);
                        bv[s][d] = ldx<T>(a.B[d], c);
                    }
                    fn[s] = ldx<T>(a.f, c);
                    if (DIM == 3) u0p[s] = (zi + 1 < nz) ? ldx<T>(a.Uin, (I)(zi + 1) * sxy + col) : u0m[s];
                    U0s[by[s]][bx[s]] = u0c[s];
                    D0s[by[s]][bx[s]] = dv[s][0];
                    B0s[by[s]][bx[s]] = bv[s][0];
                    D1s[by[s]][bx[s]] = dv[s][1];
                    B1s[by[s]][bx[s]] = bv[s][1];
                }
            }
        }
        __syncthreads();
        // ---- S1b: U1(zi) = J(U0) on the region; x / y neighbours from LDS,
        // or from memory where the mirrored neighbour lies outside the region
        if (s1) {
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                if (s < nslot) {
                    const I zo = (I)zi * sxy, row = (I)cy[s] * nx;
                    const bool xm_in = sxm[s] >= 0 && sxm[s] < BW, xp_in = sxp[s] >= 0 && sxp[s] < BW;
                    const bool ym_in = sym[s] >= 0 && sym[s] < BH, yp_in = syp[s] >= 0 && syp[s] < BH;
                    const I cim = zo + row + (I)sb_m(cx[s], g.nx), cip = zo + row + (I)sb_p(cx[s], g.nx);
                    const I cjm = zo + (I)sb_m(cy[s], g.ny) * nx + (I)cx[s];
                    const I cjp = zo + (I)sb_p(cy[s], g.ny) * nx + (I)cx[s];
                    const T d0m = xm_in ? D0s[by[s]][sxm[s]] : ldx<T>(a.D[0], cim);
                    const T b0m = xm_in ? B0s[by[s]][sxm[s]] : ldx<T>(a.B[0], cim);
                    const T d1m = ym_in ? D1s[sym[s]][bx[s]] : ldx<T>(a.D[1], cjm);
                    const T b1m = ym_in ? B1s[sym[s]][bx[s]] : ldx<T>(a.B[1], cjm);
                    T db;
                    if (DIM == 2) {
                        db = d0m - dv[s][0] + d1m - dv[s][1] - b0m + bv[s][0] - b1m + bv[s][1];
                    } else {
                        const T dsum = d0m - dv[s][0] + d1m - dv[s][1] + d2m[s] - dv[s][2];
                        const T bsum = -b0m + bv[s][0] - b1m + bv[s][1] - b2m[s] + bv[s][2];
                        db = dsum + bsum;
                    }
                    dbn[s] = db;
                    const T uxp = xp_in ? U0s[by[s]][sxp[s]] : ldx<T>(a.Uin, cip);
                    const T uxm = xm_in ? U0s[by[s]][sxm[s]] : ldx<T>(a.Uin, cim);
                    const T uyp = yp_in ? U0s[syp[s]][bx[s]] : ldx<T>(a.Uin, cjp);
                    const T uym = ym_in ? U0s[sym[s]][bx[s]] : ldx<T>(a.Uin, cjm);
                    T us = uxp + uxm + uyp + uym;
                    if (DIM == 3) us = us + u0p[s] + u0m[s];
                    U1p[zi & 3][by[s]][bx[s]] = sb_combine(a, db, us, fn[s]);
                }
            }
        }
        __syncthreads();
        // ---- S2: U2(zc) on the tile and its +x / +y halo
        const int zc = zi - 1;
        if (zc >= (z0 >= 1 ? z0 - 1 : 0) && zc <= z1 && zc < nz) {
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                if (u2ok[s]) {
                    const int b = zc & 3;
                    T us = U1p[b][by[s]][sxp[s]] + U1p[b][by[s]][sxm[s]] + U1p[b][syp[s]][bx[s]] +
                           U1p[b][sym[s]][bx[s]];
                    if (DIM == 3)
                        us = us + U1p[sb_p(zc, nz) & 3][by[s]][bx[s]] + U1p[sb_m(zc, nz) & 3][by[s]][bx[s]];
                    U2p[zc & 1][by[s]][bx[s]] = sb_combine(a, dbz[s], us, fz[s]);
                }
            }
        }
        __syncthreads();
        // ---- S3: shrinkage and Bregman update on the tile at zq
        const int zq = zi - 2;
        if (zq >= z0 && zq < z1 && own) {
            const I i = (I)zq * sxy + (I)Y[0] * nx + (I)X[0];
            const T u2 = U2p[zq & 1][by[0]][bx[0]];
            T gr[3], v[3], dn[3], bv[3];
            gr[0] = U2p[zq & 1][by[0]][sxp[0]] - u2;
            gr[1] = U2p[zq & 1][syp[0]][bx[0]] - u2;
            gr[2] = 0;
            if (DIM == 3) gr[2] = U2p[sb_p(zq, nz) & 1][by[0]][bx[0]] - u2;
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                bv[d] = bq2[d];   // loaded at S1a two planes ago
                v[d] = gr[d] + bv[d];
            }
            if (ANISO) {
#pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    T sh = fabs(v[d]) - a.thr;
                    if (sh < (T)0) sh = 0;
                    dn[d] = (v[d] != (T)0) ? (v[d] / fabs(v[d])) * sh : (T)0;
                }
            } else {
                T n2 = v[0] * v[0] + v[1] * v[1];
                if (DIM == 3) n2 = n2 + v[2] * v[2];
                const T nrm = dsqrt(n2);
                T sh = nrm - a.thr;
                if (sh < (T)0) sh = 0;
#pragma unroll
                for (int d = 0; d < DIM; ++d) dn[d] = (nrm != (T)0) ? sh * (v[d] / nrm) : (T)0;
            }
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                stnt<T>(a.Dout[d], i, dn[d]);
                stnt<T>(a.Bout[d], i, bv[d] + (gr[d] - dn[d]));
            }
            stnt<T>(a.Uout, i, u2);
            if (a.check) {
                const T df = u2 - U1p[zq & 3][by[0]][bx[0]];
                sdd += (double)df * (double)df;
                suu += (double)u2 * (double)u2;
            }
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            bq2[d] = bq1[d];
            bq1[d] = s1 ? bv[0][d] : (T)0;
        }
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            if (s1) {
                dbz[s] = dbn[s];
                fz[s] = fn[s];
                if (DIM == 3) {
                    u0m[s] = u0c[s];
                    u0c[s] = u0p[s];
                    d2m[s] = dv[s][2];
                    b2m[s] = bv[s][2];
                }
            }
        }
    }
    if (a.check) {
        block_sum2(sdd, suu, red);
        if (tid == 0) {
            a.partials[2 * tc.lin] = sdd;
            a.partials[2 * tc.lin + 1] = suu;
        }
    }
}

// The 3D pass on 64 x 8 tiles runs its ring on four waves of its own.
template <int DIM, int TY>
constexpr int sb_ring_waves() { return (DIM == 3 && TY == 8) ? 4 : 0; }

template <int DIM, int TY, typename T>
static const void* sb_fused_ptr(bool aniso, bool small) {
    constexpr int RW = sb_ring_waves<DIM, TY>();
    if (small)
        return aniso ? (const void*)sb_fused_kernel<DIM, TY, T, true, uint32_t, RW>
                     : (const void*)sb_fused_kernel<DIM, TY, T, false, uint32_t, RW>;
    return aniso ? (const void*)sb_fused_kernel<DIM, TY, T, true, uint64_t, RW>
                 : (const void*)sb_fused_kernel<DIM, TY, T, false, uint64_t, RW>;
}

template <typename T, typename S>
int sb_solve(const T* d_f, T* d_out, T mu_in, int iter, T epsil, int method, int nx, int ny, int nz, const Opts& opts,
             hipStream_t s, int* iters_done, float* re, const WarmState* ws) {
    const int dim = nz <= 1 ? 2 : 3;
    if (dim == 2) nz = 1;
    const size_t n = (size_t)nx * ny * nz;
    *iters_done = 0;
    *re = 0.f;
    if (iter <= 0) {   // Output = Input (SB_TV_core.c:58,101)
        CILREG_CK(hipMemcpyAsync(d_out, d_f, n * sizeof(T), hipMemcpyDeviceToDevice, s));
        CILREG_CK(hipStreamSynchronize(s));
        return 0;
    }
    const bool small = !opts.wide_index && fits_u32(n, sizeof(T));
    const bool aniso = method == 1;
    // one fused pass per iteration, unless the last tile is one column / row
    // wide (its mirrored neighbour lies outside the region) or depth 1 asks
    // for one pass per sweep
    const int fty = (dim == 3 && opts.tile_y != 4) ? 8 : 4;
    const bool fused = opts.depth != 1 && (nx == 1 || nx % TX != 1) && (ny == 1 || ny % fty != 1);
    Geom g;
    if (fused)
        g = dim == 2 ? make_geom(nx, ny, nz, 4, 1)
                     : (fty == 8 ? plan_geom(sb_fused_ptr<3, 8, T>(aniso, small), sb_fused_threads<8, sb_ring_waves<3, 8>()>(),
                                             nx, ny, nz, 8, opts.zchunk, 2.0)
                                 : plan_geom(sb_fused_ptr<3, 4, T>(aniso, small), TX * 4, nx, ny, nz, 4, opts.zchunk, 2.0));
    else
        // the two-kernel path is built for 64 x 4 tiles
        g = dim == 2 ? make_geom(nx, ny, nz, 4, 1)
                     : plan_geom(small ? (const void*)sb_update_kernel<3, 4, T, false, uint32_t>
                                       : (const void*)sb_update_kernel<3, 4, T, false, uint64_t>,
                                 TX * 4, nx, ny, nz, 4, opts.zchunk);
    Arena ar;
    size_t oU0, oU1, oD[2][3], oB[2][3], oCtrl, oPart;
    oU0 = ar.reserve(n * sizeof(T));
    oU1 = ar.reserve(n * sizeof(T));
    for (int b = 0; b < 2; ++b)
        for (int c = 0; c < dim; ++c) {
            oD[b][c] = ar.reserve(n * sizeof(T));
            oB[b][c] = ar.reserve(n * sizeof(T));
        }
    oCtrl = ar.reserve(sizeof(Ctrl));
    oPart = ar.reserve((size_t)g.nblocks * 2 * sizeof(double));
    if (ar.allocate() != 0) return -1;
    // an iteration reads U (buffer U0 at k = 0), D and B (set 0) and nothing
    // older, so (U, D, B) is the whole state: U = f, D = B = 0 cold
    // (SB_TV_core.c:58-60), else the caller's channels U, D1..D3, B1..B3
    if (ws && ws->warm) {
        state_in<T>((T*)ar.at(oU0), ws->d_state, n, s);
        for (int c = 0; c < dim; ++c) {
            state_in<T>((T*)ar.at(oD[0][c]), ws->d_state + (size_t)(1 + c) * n, n, s);
            state_in<T>((T*)ar.at(oB[0][c]), ws->d_state + (size_t)(4 + c) * n, n, s);
        }
    } else {
        launch_copy(ar.at(oU0), d_f, n * sizeof(T), s);
        for (int c = 0; c < dim; ++c) {
            launch_zero(ar.at(oD[0][c]), n * sizeof(T), s);
            launch_zero(ar.at(oB[0][c]), n * sizeof(T), s);
        }
    }
    Ctrl* ctrl = (Ctrl*)ar.at(oCtrl);
    double* partials = (double*)ar.at(oPart);
    // constants in the working precision (SB_TV_core.c:45-46, :153, :179, :258)
    const T mu = (T)1 / mu_in;
    const T lam = (T)2 * mu;
    const T nc = (T)1 / (mu + (dim == 3 ? (T)6 : (T)4) * lam);
    const T thr = (T)1 / lam;
    auto step = [&](int k, bool check) {
        SbArgs<T> a;
        const int cur = k & 1, nxt = cur ^ 1;
        a.f = d_f;
        for (int c = 0; c < 3; ++c) {
            const int cc = c < dim ? c : 0;
            a.D[c] = (const T*)ar.at(oD[cur][cc]);
            a.B[c] = (const T*)ar.at(oB[cur][cc]);
            a.Dout[c] = (T*)ar.at(oD[nxt][cc]);
            a.Bout[c] = (T*)ar.at(oB[nxt][cc]);
        }
        a.lam = lam;
        a.mu = mu;
        a.nc = nc;
        a.thr = thr;
        a.partials = partials;
        a.ctrl = ctrl;
        a.g = g;
        if (fused) {   // U_k -> U_{k+1}, ping-pong between the two U buffers
            a.Uin = (const T*)ar.at((k & 1) ? oU1 : oU0);
            a.Uout = (T*)ar.at((k & 1) ? oU0 : oU1);
            a.check = check;
            const int rw = (dim == 3 && fty == 8) ? sb_ring_waves<3, 8>() : 0;
            dim3 grid(g.nblocks), block(TX, fty + rw);
            SbArgs<T> arg = a;
            void* args[] = {&arg};
            const void* kp = dim == 3 ? (fty == 8 ? sb_fused_ptr<3, 8, T>(aniso, small) : sb_fused_ptr<3, 4, T>(aniso, small))
                                      : sb_fused_ptr<2, 4, T>(aniso, small);
            (void)hipLaunchKernel(kp, grid, block, args, 0, s);
            return;
        }
        dim3 grid(g.nblocks), block(TX, 4);
        // sweep 1: U0 -> U1
        a.Uin = (const T*)ar.at(oU0);
        a.Uout = (T*)ar.at(oU1);
        a.check = 0;
        if (small) {
            if (dim == 3) hipLaunchKernelGGL((sb_sweep_kernel<3, 4, T, uint32_t>), grid, block, 0, s, a);
            else hipLaunchKernelGGL((sb_sweep_kernel<2, 4, T, uint32_t>), grid, block, 0, s, a);
        } else {
            if (dim == 3) hipLaunchKernelGGL((sb_sweep_kernel<3, 4, T, uint64_t>), grid, block, 0, s, a);
            else hipLaunchKernelGGL((sb_sweep_kernel<2, 4, T, uint64_t>), grid, block, 0, s, a);
        }
        // sweep 2 + shrink + Bregman: U1 -> U2 (into U0's buffer)
        a.Uin = (const T*)ar.at(oU1);
        a.Uout = (T*)ar.at(oU0);
        a.check = check;
        if (small) {
            if (dim == 3) {
                if (method == 1) hipLaunchKernelGGL((sb_update_kernel<3, 4, T, true, uint32_t>), grid, block, 0, s, a);
                else hipLaunchKernelGGL((sb_update_kernel<3, 4, T, false, uint32_t>), grid, block, 0, s, a);
            } else {
                if (method == 1) hipLaunchKernelGGL((sb_update_kernel<2, 4, T, true, uint32_t>), grid, block, 0, s, a);
                else hipLaunchKernelGGL((sb_update_kernel<2, 4, T, false, uint32_t>), grid, block, 0, s, a);
            }
        } else {
            if (dim == 3) {
                if (method == 1) hipLaunchKernelGGL((sb_update_kernel<3, 4, T, true, uint64_t>), grid, block, 0, s, a);
                else hipLaunchKernelGGL((sb_update_kernel<3, 4, T, false, uint64_t>), grid, block, 0, s, a);
            } else {
                if (method == 1) hipLaunchKernelGGL((sb_update_kernel<2, 4, T, true, uint64_t>), grid, block, 0, s, a);
                else hipLaunchKernelGGL((sb_update_kernel<2, 4, T, false, uint64_t>), grid, block, 0, s, a);
            }
        }
    };
    int done = 0;
    if (drive_iterations(iter, (float)epsil, opts, ctrl, partials, g.nblocks, s, step, &done, re) != 0) return -1;
    *iters_done = done;
    // the fused path leaves U of the last executed iteration in buffer (last + 1) mod 2
    const int last = (done < iter) ? done : iter - 1;
    const size_t oRes = (fused && ((last + 1) & 1)) ? oU1 : oU0;
    CILREG_CK(hipMemcpyAsync(d_out, ar.at(oRes), n * sizeof(T), hipMemcpyDeviceToDevice, s));
    if (ws) {   // the state after the executed iterations: D and B in set (last + 1) mod 2
        const int set = (last + 1) & 1;
        state_out<T>(ws->d_state, (const T*)ar.at(oRes), n, s);
        for (int c = 0; c < dim; ++c) {
            state_out<T>(ws->d_state + (size_t)(1 + c) * n, (const T*)ar.at(oD[set][c]), n, s);
            state_out<T>(ws->d_state + (size_t)(4 + c) * n, (const T*)ar.at(oB[set][c]), n, s);
        }
    }
    CILREG_CK(hipStreamSynchronize(s));
    return 0;
}

template int sb_solve<float, float>(const float*, float*, float, int, float, int, int, int, int, const Opts&,
                                    hipStream_t, int*, float*, const WarmState*);

}  // namespace cilreg
