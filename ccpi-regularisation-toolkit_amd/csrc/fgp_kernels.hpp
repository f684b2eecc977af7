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
    __shared__ double red[kRedSlots];

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

// ---------------------------------------------------------------------------
// Two iterations per pass (temporal blocking of depth 2, 3D)
// ---------------------------------------------------------------------------
// One pass advances k -> k+2: it reads f, P_k, P_{k-1} once and writes
// P_{k+1} (the next pass's P_{k-1}) and P_{k+2}, so a voxel-iteration moves
// 26 B instead of 40 (32 instead of 48 with a check every iteration).  Every
// value is computed by the same expressions, in the same order, as the
// one-iteration kernel: the results are bit-identical.
//
// The workgroup owns the 64 x TY output tile T and evaluates the first
// iteration on the ring around it, B = [x0-1, x0+64] x [y0-1, y0+TY+1]
// (67 x (TY+3) points): slot 0 of every thread is its own point of T, slot 1
// of 3*67 + 3*TY ring lanes a point of B \ T.  Validity shrinks by one point
// per dependent stage (D_k on B, P_{k+1} and R_{k+1} up to x0+64 / y0+TY,
// D_{k+1} from x0 / y0 on, P_{k+2} on T), so every value a stage reads from a
// neighbour was computed by the previous stage from loaded data; values at
// ring points outside the volume are computed on a clamped stand-in and only
// ever read where the boundary rule selects zero instead.
//
// z pipeline, for input plane zi of the march over [z0 - 1, z1 + 1]:
//   S1  D_k(zi)        from P_k, P_{k-1}, f at zi and R_k,3(zi - 1)
//   S2  P_{k+1}(zi-1)  from R_k(zi - 1), D_k(zi - 1) and its +x / +y / +z
//                      neighbours; R_{k+1}(zi-1) = P + m_{k+1} (P - P_k)
//   S3  D_{k+1}(zi-1)  from R_{k+1}(zi - 1), its -x / -y neighbours and
//                      R_{k+1},3(zi - 2)
//   S4  P_{k+2}(zi-2)  on T
// P_{k+1} is rounded through the storage type before it enters R_{k+1}, as
// the one-iteration kernel reads it back from memory (through_storage).

template <typename T, typename S>
struct Fgp2Args {
    const T* f;
    const S* Pm[3];     // P_{k-1}
    const S* P[3];      // P_k
    S* Pn[3];           // P_{k+1}
    S* Pnn[3];          // P_{k+2}
    T* D0;              // D_k   (written when write_d0)
    T* D1;              // D_{k+1} (written when write_d1; the buffer of D_{k-1})
    const T* Dprev;     // D_{k-1} (read when check0)
    T lambda, step, mult0, mult1;
    int write_d0, write_d1, check0, check1, nonneg;
    double* partials0;  // residual sums of iteration k
    double* partials1;  // of iteration k + 1
    const Ctrl* ctrl;
    Geom g;
};

// Projection of P onto the unit ball (iso) or box (aniso), utils.c:184-211,
// exactly as in fgp_body.
template <int DIM, typename T, bool ISO>
__device__ inline void fgp_project(T& p1, T& p2, T& p3) {
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
}

template <int TY>
struct Fgp2Shape {
    static constexpr int BW = TX + 3, BH = TY + 3;   // ring region B
    static constexpr int NR = 3 * BW + 3 * TY;       // ring points B \ T
    static constexpr int NT = TX * TY;
};

// DIM = 2 runs the same pipeline on one plane (no z terms; P has two components).
template <int TY, typename T, typename S, bool ISO, typename I, int DIM = 3>
__device__ inline void fgp2_body(const Fgp2Args<T, S>& a, const TileCoord& tc) {
    using Sh = Fgp2Shape<TY>;
    constexpr int BW = Sh::BW, BH = Sh::BH, NR = Sh::NR, NT = Sh::NT;
    const Geom& g = a.g;
    const int lx = threadIdx.x, ly = threadIdx.y, tid = ly * TX + lx;

    // slot 0: own point of T; slot 1: a ring point (rotated over the waves)
    int bx[2], by[2];
    bx[0] = lx + 1;
    by[0] = ly + 1;
    const int hbase = (tc.lin & 1) ? NT - NR : 0;
    const int h = tid - hbase;
    const bool ring = h >= 0 && h < NR;
    {
        int rx = 0, ry = 0;
        if (h < 3 * BW) {
            const int row = h / BW;
            rx = h - row * BW;
            ry = row == 0 ? 0 : TY + row;          // rows y0-1, y0+TY, y0+TY+1
        } else {
            const int k = h - 3 * BW, col = k / TY;
            rx = col == 0 ? 0 : TX + col;          // columns x0-1, x0+TX, x0+TX+1
            ry = k - col * TY + 1;
        }
        bx[1] = ring ? rx : 0;
        by[1] = ring ? ry : 0;
    }
    const int nslot = ring ? 2 : 1;
    int X[2], Y[2], cx[2], cy[2];
    #pragma unroll
    for (int s = 0; s < 2; ++s) {
        X[s] = tc.x0 + bx[s] - 1;
        Y[s] = tc.y0 + by[s] - 1;
        const bool in = X[s] >= 0 && Y[s] >= 0 && X[s] < g.nx && Y[s] < g.ny;
        cx[s] = in ? X[s] : 0;    // clamped stand-in for loads
        cy[s] = in ? Y[s] : 0;
    }
    const bool own = X[0] < g.nx && Y[0] < g.ny;
    // per slot: the clamped column index and the boundary tests the stages
    // make, as bits (X, Y themselves are not kept)
    constexpr int F_XGT0 = 1, F_YGT0 = 2, F_XLAST = 4, F_YLAST = 8, F_CXGT0 = 16, F_CYGT0 = 32;
    I col[2];
    int fl[2];
    #pragma unroll
    for (int s = 0; s < 2; ++s) {
        col[s] = (I)cy[s] * (I)g.nx + (I)cx[s];
        fl[s] = (X[s] > 0 ? F_XGT0 : 0) | (Y[s] > 0 ? F_YGT0 : 0) | (X[s] == g.nx - 1 ? F_XLAST : 0) |
                (Y[s] == g.ny - 1 ? F_YLAST : 0) | (cx[s] > 0 ? F_CXGT0 : 0) | (cy[s] > 0 ? F_CYGT0 : 0);
    }

    __shared__ T Dk[2][BH + 1][BW + 1];    // D_k, planes by parity
    __shared__ T Dk1[2][BH + 1][BW + 1];   // D_{k+1}
    __shared__ T Rk1[BH + 1][BW + 1];      // R_k,1 and ,2 of plane zi
    __shared__ T Rk2[BH + 1][BW + 1];
    __shared__ T R1s[BH + 1][BW + 1];      // R_{k+1},1 and ,2 of the current plane
    __shared__ T R2s[BH + 1][BW + 1];
    __shared__ double red[kRedSlots];

    const I sxy = (I)g.sxy, nx = (I)g.nx;
    const T lam = a.lambda, m0 = a.mult0, m1 = a.mult1, step = a.step;
    const int nz = g.nz, z0 = tc.z0, z1 = tc.z1;

    // per-slot pipeline registers
    T rk[2][3] = {};    // R_k(zi - 1)
    T pk[2][3] = {};    // P_k(zi - 1) as loaded
    T rq[2][3] = {};    // R_{k+1}(zi - 2)
    T fz[2] = {};       // f(zi - 1)

    if (DIM == 3 && z0 >= 2) {   // R_k,3 two planes below the chunk start
        #pragma unroll
        for (int s = 0; s < 2; ++s) {
            if (s < nslot) {
                const I i = (I)(z0 - 2) * sxy + col[s];
                const T p = ldx<T>(a.P[2], i), q = ldx<T>(a.Pm[2], i);
                rk[s][2] = p + m0 * (p - q);
            }
        }
    }
    double sdd0 = 0.0, suu0 = 0.0, sdd1 = 0.0, suu1 = 0.0;
    // With 16-bit P the four per-thread residual sums live in LDS slots of
    // their own between check planes (eight registers fewer: 256^3 BF16
    // 0.103 -> 0.097 ms per iteration; FP32 measured 1% slower), added in
    // the same order.
    constexpr bool ACC_LDS = sizeof(S) == 2;
    __shared__ double acc[4][ACC_LDS ? NT : 1];
    const int at = ACC_LDS ? tid : 0;
    if (ACC_LDS) {
        #pragma unroll
        for (int c = 0; c < 4; ++c) acc[c][at] = 0.0;
    }
    auto add_acc = [&](int c0, double& r0, double& r1, double v0, double v1) {
        if (ACC_LDS) {
            acc[c0][at] += v0;
            acc[c0 + 1][at] += v1;
        } else {
            r0 += v0;
            r1 += v1;
        }
    };
    const int zstart = z0 >= 1 ? z0 - 1 : 0;
    const int zend = z1 + 1;
    for (int zi = zstart; zi <= zend; ++zi) {
        // ---- S1a: R_k(zi) on the region, into LDS
        T rn[2][3] = {}, pn[2][3] = {}, fn[2] = {};
        const bool s1 = zi < nz;
        if (s1) {
            #pragma unroll
            for (int s = 0; s < 2; ++s) {
                if (s < nslot) {
                    const I i = (I)zi * sxy + col[s];
                    #pragma unroll
                    for (int c = 0; c < DIM; ++c) {
                        const T p = ldx<T>(a.P[c], i), q = ldx<T>(a.Pm[c], i);
                        pn[s][c] = p;
                        rn[s][c] = p + m0 * (p - q);
                    }
                    fn[s] = ldx<T>(a.f, i);
                    Rk1[by[s]][bx[s]] = rn[s][0];
                    Rk2[by[s]][bx[s]] = rn[s][1];
                }
            }
        }
        __syncthreads();
        // ---- S1b: D_k(zi); the -x / -y neighbours of the region's first
        // column / row lie outside it and are evaluated from memory
        if (s1) {
            #pragma unroll
            for (int s = 0; s < 2; ++s) {
                if (s < nslot) {
                    const I i = (I)zi * sxy + col[s];
                    T b1r, b2r;
                    if (bx[s] > 0) {
                        b1r = Rk1[by[s]][bx[s] - 1];
                    } else {
                        const I im = (fl[s] & F_CXGT0) ? i - 1 : i;
                        const T b1p = ldx<T>(a.P[0], im), b1q = ldx<T>(a.Pm[0], im);
                        b1r = b1p + m0 * (b1p - b1q);
                    }
                    if (by[s] > 0) {
                        b2r = Rk2[by[s] - 1][bx[s]];
                    } else {
                        const I jm = (fl[s] & F_CYGT0) ? i - nx : i;
                        const T b2p = ldx<T>(a.P[1], jm), b2q = ldx<T>(a.Pm[1], jm);
                        b2r = b2p + m0 * (b2p - b2q);
                    }
                    const T b1 = (fl[s] & F_XGT0) ? b1r : (T)0;
                    const T b2 = (fl[s] & F_YGT0) ? b2r : (T)0;
                    T d;
                    if (DIM == 3) {
                        const T b3 = zi > 0 ? rk[s][2] : (T)0;
                        d = fn[s] - lam * (rn[s][0] + rn[s][1] + rn[s][2] - b1 - b2 - b3);
                    } else {
                        d = fn[s] - lam * (rn[s][0] + rn[s][1] - b1 - b2);
                    }
                    if (a.nonneg && d < (T)0) d = (T)0;
                    Dk[zi & 1][by[s]][bx[s]] = d;
                    if (s == 0 && own && zi >= z0 && zi < z1) {
                        const I io = (I)zi * sxy + col[0];
                        if (a.write_d0) stnt<T>(a.D0, io, d);
                        if (a.check0) {
                            const T df = d - ldx<T>(a.Dprev, io);
                            add_acc(0, sdd0, suu0, (double)df * (double)df, (double)d * (double)d);
                        }
                    }
                }
            }
        }
        __syncthreads();
        // ---- S2: P_{k+1}(zc), R_{k+1}(zc)
        const int zc = zi - 1;
        T rqn[2][3] = {};
        const bool s2 = zc >= (z0 >= 1 ? z0 - 1 : 0) && zc <= z1 && zc < nz && zc >= 0;
        if (s2) {
            #pragma unroll
            for (int s = 0; s < 2; ++s) {
                if (s < nslot) {
                    const int b = zc & 1;
                    const T d = Dk[b][by[s]][bx[s]];
                    const T g1 = (fl[s] & F_XLAST) ? (T)0 : d - Dk[b][by[s]][bx[s] + 1];
                    const T g2 = (fl[s] & F_YLAST) ? (T)0 : d - Dk[b][by[s] + 1][bx[s]];
                    T p1 = rk[s][0] + step * g1;
                    T p2 = rk[s][1] + step * g2;
                    T p3 = 0;
                    if (DIM == 3) {
                        const T g3 = (zc == nz - 1) ? (T)0 : d - Dk[b ^ 1][by[s]][bx[s]];
                        p3 = rk[s][2] + step * g3;
                    }
                    fgp_project<DIM, T, ISO>(p1, p2, p3);
                    if (s == 0 && own && zc >= z0 && zc < z1) {
                        const I io = (I)zc * sxy + col[0];
                        stnt<T>(a.Pn[0], io, p1);
                        stnt<T>(a.Pn[1], io, p2);
                        if (DIM == 3) stnt<T>(a.Pn[2], io, p3);
                    }
                    p1 = through_storage<T, S>(p1);
                    p2 = through_storage<T, S>(p2);
                    rqn[s][0] = p1 + m1 * (p1 - pk[s][0]);
                    rqn[s][1] = p2 + m1 * (p2 - pk[s][1]);
                    if (DIM == 3) {
                        p3 = through_storage<T, S>(p3);
                        rqn[s][2] = p3 + m1 * (p3 - pk[s][2]);
                    }
                    R1s[by[s]][bx[s]] = rqn[s][0];
                    R2s[by[s]][bx[s]] = rqn[s][1];
                }
            }
        }
        __syncthreads();
        // ---- S3: D_{k+1}(zc)
        const bool s3 = zc >= z0 && zc <= z1 && zc < nz;
        if (s3) {
            #pragma unroll
            for (int s = 0; s < 2; ++s) {
                if (s < nslot) {
                    const T b1r = R1s[by[s]][bx[s] > 0 ? bx[s] - 1 : 0];
                    const T b2r = R2s[by[s] > 0 ? by[s] - 1 : 0][bx[s]];
                    const T b1 = (fl[s] & F_XGT0) ? b1r : (T)0;
                    const T b2 = (fl[s] & F_YGT0) ? b2r : (T)0;
                    T d;
                    if (DIM == 3) {
                        const T b3 = zc > 0 ? rq[s][2] : (T)0;
                        d = fz[s] - lam * (rqn[s][0] + rqn[s][1] + rqn[s][2] - b1 - b2 - b3);
                    } else {
                        d = fz[s] - lam * (rqn[s][0] + rqn[s][1] - b1 - b2);
                    }
                    if (a.nonneg && d < (T)0) d = (T)0;
                    Dk1[zc & 1][by[s]][bx[s]] = d;
                    if (s == 0 && own && zc < z1) {
                        const I io = (I)zc * sxy + col[0];
                        if (a.write_d1) stnt<T>(a.D1, io, d);
                        if (a.check1) {
                            const T df = d - Dk[zc & 1][by[0]][bx[0]];
                            add_acc(2, sdd1, suu1, (double)df * (double)df, (double)d * (double)d);
                        }
                    }
                }
            }
        }
        __syncthreads();
        // ---- S4: P_{k+2}(zq) on T
        const int zq = zi - 2;
        if (zq >= z0 && zq < z1 && own) {
            const int b = zq & 1;
            const T d = Dk1[b][by[0]][bx[0]];
            const T g1 = (fl[0] & F_XLAST) ? (T)0 : d - Dk1[b][by[0]][bx[0] + 1];
            const T g2 = (fl[0] & F_YLAST) ? (T)0 : d - Dk1[b][by[0] + 1][bx[0]];
            T p1 = rq[0][0] + step * g1;
            T p2 = rq[0][1] + step * g2;
            T p3 = 0;
            if (DIM == 3) {
                const T g3 = (zq == nz - 1) ? (T)0 : d - Dk1[b ^ 1][by[0]][bx[0]];
                p3 = rq[0][2] + step * g3;
            }
            fgp_project<DIM, T, ISO>(p1, p2, p3);
            const I io = (I)zq * sxy + col[0];
            stnt<T>(a.Pnn[0], io, p1);
            stnt<T>(a.Pnn[1], io, p2);
            if (DIM == 3) stnt<T>(a.Pnn[2], io, p3);
        }
        // ---- shift the pipeline
        #pragma unroll
        for (int s = 0; s < 2; ++s) {
            #pragma unroll
            for (int c = 0; c < 3; ++c) {
                if (s1) { rk[s][c] = rn[s][c]; pk[s][c] = pn[s][c]; }
                if (s2) rq[s][c] = rqn[s][c];
            }
            if (s1) fz[s] = fn[s];
        }
    }
    if (ACC_LDS) {
        sdd0 = acc[0][at];
        suu0 = acc[1][at];
        sdd1 = acc[2][at];
        suu1 = acc[3][at];
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
// at 80 VGPRs (six waves per SIMD: three 512-thread workgroups per CU); with
// 16-bit P it would otherwise take 83 and lose one workgroup per CU (256^3
// BF16: 0.114 -> 0.101 ms per iteration with the cap, 3 VGPRs spilled).  FP64
// and 64-bit offsets are not capped: their bodies need far more.
template <typename T, typename I>
constexpr int fgp2_min_waves() { return (std::is_same<T, float>::value && sizeof(I) == 4) ? 6 : 1; }

template <int TY, typename T, typename S, bool ISO, typename I, int DIM = 3>
__global__ __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(fgp2_min_waves<T, I>())))
void fgp2_iter_kernel(Fgp2Args<T, S> a) {
    if (solve_stopped(a.ctrl)) return;
    fgp2_body<TY, T, S, ISO, I, DIM>(a, tile_of_block(a.g, TY));
}

}  // namespace cilreg
