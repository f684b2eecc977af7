// dtv.hip -- FGP-dTV (directional total variation, Ehrhardt & Betcke): FGP
// whose dual field and objective gradient are projected onto the orthogonal
// complement of a reference image's normalised gradient, one fused kernel per
// iteration.
//
// Reference algorithm: src/Core/regularisers_CPU/FGP_dTV_core.c:46-380.
// Before the iterations, from the reference image u_ref:
//   B = grad(u_ref) / sqrt(|grad(u_ref)|^2 + eta^2)        (GradNorm, :188-208, :279-306)
// with forward differences that read 0 past the last index.  Per iteration k,
// with m_k = (t_{k-1} - 1) / t_k:
//   R   = P_k + m_k (P_k - P_{k-1})                        (Rupd, :263-274, :368-380)
//   R'  = R - <R, B> B                                      (ProjectVect, :210-223, :308-323)
//   D_k = f - lambda * sum_d (R'_d(x) - R'_d(x - e_d)), R'_d(-1) = 0, max(., 0) if nonneg
//   g_d = D_k(x) - D_k(x + e_d) (0 at the last index);  g' = g - <g, B> B
//   P_{k+1} = proj(R' + g' / (26 lambda))                  (Grad_dfunc, :240-262, :342-367)
// with 1/(8 lambda) in 2D, and t_{k+1} = (1 + sqrt(1 + 4 t_k^2)) / 2 taken
// through a double square root (:95, :149).  The output is D of the last
// iteration; the stopping rule is FGP's (every 5th iteration, four hits).
//
// HBM traffic per voxel-iteration (FP32): read f, B(3), P_k(3), P_{k-1}(3),
// write P_{k+1}(3) = 52 B in 3D (40 B in 2D); R and R' are recomputed, never
// stored; B is computed once per solve.
//
// Fusion: a workgroup owns a 64 x TY column tile and marches through a chunk
// of z.  For plane z+1 it evaluates R' on the tile and on the ring of points
// the divergence reads (-x / -y neighbours of the tile and of its +x / +y halo)
// into two LDS planes, then D on the tile plus its +x / +y halo into an LDS
// plane (double-buffered across z), then updates P on plane z.  R'_3 of the
// plane below and B of the own point stay in registers.
#include "solvers.hpp"
#include "fgp_kernels.hpp"

namespace cilreg {

template <typename T>
struct DtvArgs {
    const T* f;
    const T* B[3];      // normalised reference gradient
    const T* P[3];      // P_k
    const T* Pp[3];     // P_{k-1}
    T* Pn[3];           // P_{k+1}
    T* Dout;            // D_k (written when write_d)
    const T* Dprev;     // D_{k-1} (read when check)
    T lambda, step, mult;
    int write_d, check, nonneg;
    double* partials;
    const Ctrl* ctrl;
    Geom g;
};

// B from the reference image; one thread per voxel (runs once per solve).
template <int DIM, typename T>
__global__ __launch_bounds__(256) void dtv_refgrad_kernel(const T* ref, T* B1, T* B2, T* B3, T eta, int nx, int ny,
                                                          int nz) {
    const size_t sxy = (size_t)nx * ny, n = sxy * (size_t)nz;
    const T eta_sq = eta * eta;
    for (size_t c = (size_t)blockIdx.x * 256 + threadIdx.x; c < n; c += (size_t)gridDim.x * 256) {
        const size_t k = c / sxy, r = c - k * sxy;
        const int j = (int)(r / (size_t)nx), i = (int)(r - (size_t)j * nx);
        const T rc = ref[c];
        const T v1 = (i == nx - 1) ? (T)0 : ref[c + 1];
        const T v2 = (j == ny - 1) ? (T)0 : ref[c + nx];
        const T g1 = v1 - rc, g2 = v2 - rc;
        T g3 = 0;
        T m = g1 * g1 + g2 * g2;
        if (DIM == 3) {
            const T v3 = ((int)k == nz - 1) ? (T)0 : ref[c + sxy];
            g3 = v3 - rc;
            m = m + g3 * g3;
        }
        m = (T)sqrt((double)(m + eta_sq));
        B1[c] = g1 / m;
        B2[c] = g2 / m;
        if (DIM == 3) B3[c] = g3 / m;
    }
}

template <int DIM, int TY, typename T, bool ISO, typename I>
__device__ inline void dtv_body(const DtvArgs<T>& a, const TileCoord& tc) {
    const Geom& g = a.g;
    constexpr int NT = TX * TY;
    constexpr int NH = TX + TY;                // D halo: +x column and +y row
    constexpr int NR = 2 * TX + 2 * TY + 2;    // R' ring (includes the D halo points first)
    static_assert(NR <= NT, "ring lanes must fit the workgroup");
    const int lx = threadIdx.x, ly = threadIdx.y, tid = ly * TX + lx;
    const int x = tc.x0 + lx, y = tc.y0 + ly;
    const bool own = x < g.nx && y < g.ny;
    // ring lanes alternate between the low and high waves with the block index
    const int hbase = (tc.lin & 1) ? NT - NR : 0;
    const int h = tid - hbase;
    const bool rlane = h >= 0 && h < NR;
    const bool hlane = h >= 0 && h < NH;
    // ring point (rx, ry): lanes [0, TY) the +x column on the tile's rows,
    // [TY, TY+TX) the +y row -- these also evaluate D --, then the +x column
    // at y0-1, the -x column on rows y0 .. y0+TY and the -y row on x0 .. x0+63
    int rx = 0, ry = 0;
    if (h < TY) { rx = tc.x0 + TX; ry = tc.y0 + h; }
    else if (h < TY + TX) { rx = tc.x0 + h - TY; ry = tc.y0 + TY; }
    else if (h == TY + TX) { rx = tc.x0 + TX; ry = tc.y0 - 1; }
    else if (h < 2 * TY + TX + 2) { rx = tc.x0 - 1; ry = tc.y0 + (h - TY - TX - 1); }
    else { rx = tc.x0 + (h - 2 * TY - TX - 2); ry = tc.y0 - 1; }
    const bool rin = rlane && rx >= 0 && ry >= 0 && rx < g.nx && ry < g.ny;
    const int sx = rx - tc.x0 + 1, sy = ry - tc.y0 + 1;   // slot in the R' planes
    // points outside the volume evaluate a valid stand-in and store nothing;
    // their values are only read where a boundary rule selects 0 instead
    const int ox = own ? x : 0, oy = own ? y : 0;
    const int rcx = rin ? rx : 0, rcy = rin ? ry : 0;

    __shared__ T Rs1[TY + 2][TX + 2];   // R'_1 on [x0-1, x0+64] x [y0-1, y0+TY]
    __shared__ T Rs2[TY + 2][TX + 2];   // R'_2 likewise
    __shared__ T Ds[2][TY + 1][TX + 1];
    __shared__ double red[kRedSlots];

    const I sxy = (I)g.sxy, nx = (I)g.nx;
    const T lam = a.lambda, mult = a.mult;
    // R' at element i, and B there
    auto rproj = [&](I i, T& r1, T& r2, T& r3, T& b1, T& b2, T& b3) {
        T p = ldx<T>(a.P[0], i), q = ldx<T>(a.Pp[0], i);
        r1 = p + mult * (p - q);
        p = ldx<T>(a.P[1], i);
        q = ldx<T>(a.Pp[1], i);
        r2 = p + mult * (p - q);
        b1 = ldx<T>(a.B[0], i);
        b2 = ldx<T>(a.B[1], i);
        T ip = r1 * b1 + r2 * b2;
        if (DIM == 3) {
            p = ldx<T>(a.P[2], i);
            q = ldx<T>(a.Pp[2], i);
            r3 = p + mult * (p - q);
            b3 = ldx<T>(a.B[2], i);
            ip = ip + r3 * b3;
            r3 = r3 - ip * b3;
        } else {
            r3 = 0;
            b3 = 0;
        }
        r1 = r1 - ip * b1;
        r2 = r2 - ip * b2;
    };
    // D at (px, py, z) from R' there, R'_1 / R'_2 of the -x / -y slots and R'_3 below
    auto dval = [&](int px, int py, int z, int ssx, int ssy, T r1, T r2, T r3, T r3below) -> T {
        const I i = (I)z * sxy + (I)py * nx + (I)px;
        const T fv = ldx<T>(a.f, i);
        const T b1 = px > 0 ? Rs1[ssy][ssx - 1] : (T)0;
        const T b2 = py > 0 ? Rs2[ssy - 1][ssx] : (T)0;
        T d;
        if (DIM == 3) {
            const T b3 = z > 0 ? r3below : (T)0;
            d = fv - lam * (r1 + r2 + r3 - b1 - b2 - b3);
        } else {
            d = fv - lam * (r1 + r2 - b1 - b2);
        }
        if (a.nonneg && d < (T)0) d = (T)0;
        return d;
    };
    // R' of plane z on the tile and the ring into LDS; own values and the
    // D-halo lanes' values are returned for the D stage and the update
    auto rplane = [&](int z, T& r1, T& r2, T& r3, T& b1, T& b2, T& b3, T& h1, T& h2, T& h3) {
        const I base = (I)z * sxy;
        rproj(base + (I)oy * nx + (I)ox, r1, r2, r3, b1, b2, b3);
        Rs1[ly + 1][lx + 1] = r1;
        Rs2[ly + 1][lx + 1] = r2;
        if (rlane) {
            T c1, c2, c3;
            rproj(base + (I)rcy * nx + (I)rcx, h1, h2, h3, c1, c2, c3);
            Rs1[sy][sx] = h1;
            Rs2[sy][sx] = h2;
        }
    };

    T cr1, cr2, cr3, cb1, cb2, cb3;   // own R' and B, plane z
    T hr1 = 0, hr2 = 0, hr3 = 0;      // ring lane's R', plane z
    T r3b = 0, hr3b = 0;              // R'_3 one plane below the chunk start
    if (DIM == 3 && tc.z0 > 0) {
        const I base = (I)(tc.z0 - 1) * sxy;
        T t1, t2, c1, c2, c3;
        rproj(base + (I)oy * nx + (I)ox, t1, t2, r3b, c1, c2, c3);
        if (hlane) rproj(base + (I)rcy * nx + (I)rcx, t1, t2, hr3b, c1, c2, c3);
    }
    rplane(tc.z0, cr1, cr2, cr3, cb1, cb2, cb3, hr1, hr2, hr3);
    __syncthreads();
    Ds[0][ly][lx] = dval(ox, oy, tc.z0, lx + 1, ly + 1, cr1, cr2, cr3, r3b);
    if (hlane) {
        const T d = dval(rcx, rcy, tc.z0, sx, sy, hr1, hr2, hr3, hr3b);
        if (rin) Ds[0][sy - 1][sx - 1] = d;
    }
    __syncthreads();

    double sdd = 0.0, suu = 0.0;
    int cur = 0;
    for (int z = tc.z0; z < tc.z1; ++z) {
        T nr1 = 0, nr2 = 0, nr3 = 0, nb1 = 0, nb2 = 0, nb3 = 0, hn1 = 0, hn2 = 0, hn3 = 0;
        if (DIM == 3 && z + 1 < g.nz) {
            rplane(z + 1, nr1, nr2, nr3, nb1, nb2, nb3, hn1, hn2, hn3);
            __syncthreads();
            Ds[cur ^ 1][ly][lx] = dval(ox, oy, z + 1, lx + 1, ly + 1, nr1, nr2, nr3, cr3);
            if (hlane) {
                const T d = dval(rcx, rcy, z + 1, sx, sy, hn1, hn2, hn3, hr3);
                if (rin) Ds[cur ^ 1][sy - 1][sx - 1] = d;
            }
        }
        __syncthreads();
        if (own) {
            const I i = (I)z * sxy + (I)y * nx + (I)x;
            const T d = Ds[cur][ly][lx];
            T g1 = (x == g.nx - 1) ? (T)0 : d - Ds[cur][ly][lx + 1];
            T g2 = (y == g.ny - 1) ? (T)0 : d - Ds[cur][ly + 1][lx];
            T g3 = 0;
            T ip = g1 * cb1 + g2 * cb2;
            if (DIM == 3) {
                g3 = (z == g.nz - 1) ? (T)0 : d - Ds[cur ^ 1][ly][lx];
                ip = ip + g3 * cb3;
            }
            g1 = g1 - ip * cb1;
            g2 = g2 - ip * cb2;
            T p1 = cr1 + a.step * g1;
            T p2 = cr2 + a.step * g2;
            T p3 = 0;
            if (DIM == 3) {
                g3 = g3 - ip * cb3;
                p3 = cr3 + a.step * g3;
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
        cr1 = nr1; cr2 = nr2; cr3 = nr3;
        cb1 = nb1; cb2 = nb2; cb3 = nb3;
        hr3 = hn3;
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

template <int DIM, int TY, typename T, bool ISO, typename I>
__global__ __launch_bounds__(TX * TY) void dtv_iter_kernel(DtvArgs<T> a) {
    if (solve_stopped(a.ctrl)) return;
    dtv_body<DIM, TY, T, ISO, I>(a, tile_of_block(a.g, TY));
}

// ---------------------------------------------------------------------------
// Two iterations per pass (3D)
// ---------------------------------------------------------------------------
// The FGP two-iteration pipeline (fgp2_body: ring region 67 x (TY+3), slot 0
// the own tile point, slot 1 a ring point; stages S1 D_k(zi), S2 P_{k+1} and
// R'_{k+1}(zi-1), S3 D_{k+1}(zi-1), S4 P_{k+2}(zi-2)) with dTV's projections.
// R'_k needs three loads of P, P_prev and B per component, so S1 stages it in
// LDS (S1a) and D_k reads the -x / -y neighbours from there (S1b); only the
// region's first column and row evaluate their outside neighbour from memory.
// R'_k and R'_{k+1} and both gradients are projected onto the complement of
// B at their point, and B of plane zi-1 (zi-2 for S4) rides along the z
// pipeline.  One pass reads f, B, P_k, P_{k-1} once and writes P_{k+1},
// P_{k+2}: 64 B per voxel per two iterations instead of 2 x 52.  Every value
// is computed by the one-iteration kernel's expressions in its order, so the
// results are bit-identical.

template <typename T>
struct Dtv2Args {
    const T* f;
    const T* B[3];
    const T* Pm[3];     // P_{k-1}
    const T* P[3];      // P_k
    T* Pn[3];           // P_{k+1}
    T* Pnn[3];          // P_{k+2}
    T* D0;              // D_k (written when write_d0)
    T* D1;              // D_{k+1} (written when write_d1; the buffer of D_{k-1})
    const T* Dprev;     // D_{k-1} (read when check0)
    T lambda, step, mult0, mult1;
    int write_d0, write_d1, check0, check1, nonneg;
    double* partials0;
    double* partials1;
    const Ctrl* ctrl;
    Geom g;
};

// v <- v - <v, B> B, the inner product summed in the reference's order
template <typename T>
__device__ inline void dtv_orth3(T& v1, T& v2, T& v3, T b1, T b2, T b3) {
    T ip = v1 * b1 + v2 * b2;
    ip = ip + v3 * b3;
    v1 = v1 - ip * b1;
    v2 = v2 - ip * b2;
    v3 = v3 - ip * b3;
}

template <int TY, typename T, bool ISO, typename I>
__device__ inline void dtv2_body(const Dtv2Args<T>& a, const TileCoord& tc) {
    using Sh = Fgp2Shape<TY>;
    constexpr int BW = Sh::BW, BH = Sh::BH, NR = Sh::NR, NT = Sh::NT;
    const Geom& g = a.g;
    const int lx = threadIdx.x, ly = threadIdx.y, tid = ly * TX + lx;

    // slot 0: own point of the tile; slot 1: a ring point (rotated over the waves)
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

    __shared__ T Dk[2][BH + 1][BW + 1];    // D_k, planes by parity
    __shared__ T Dk1[2][BH + 1][BW + 1];   // D_{k+1}
    __shared__ T Rk1[BH + 1][BW + 1];      // R'_k,1 and ,2 of plane zi
    __shared__ T Rk2[BH + 1][BW + 1];
    __shared__ T R1s[BH + 1][BW + 1];      // R'_{k+1},1 and ,2 of the current plane
    __shared__ T R2s[BH + 1][BW + 1];
    __shared__ double red[kRedSlots];

    const I sxy = (I)g.sxy, nx = (I)g.nx;
    const T lam = a.lambda, m0 = a.mult0, m1 = a.mult1, step = a.step;
    const int nz = g.nz, z0 = tc.z0, z1 = tc.z1;

    // R'_k at element i, with P_k and B there
    auto rk_at = [&](I i, T* r, T* p, T* b) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const T pv = ldx<T>(a.P[c], i), qv = ldx<T>(a.Pm[c], i);
            p[c] = pv;
            r[c] = pv + m0 * (pv - qv);
            b[c] = ldx<T>(a.B[c], i);
        }
        dtv_orth3(r[0], r[1], r[2], b[0], b[1], b[2]);
    };

    // per-slot pipeline registers
    T rk[2][3] = {};    // R'_k(zi - 1)
    T pk[2][3] = {};    // P_k(zi - 1)
    T bk[2][3] = {};    // B(zi - 1)
    T rq[2][3] = {};    // R'_{k+1}(zi - 2)
    T bq[3] = {};       // B(zi - 2) of the own point
    T fz[2] = {};       // f(zi - 1)

    if (z0 >= 2) {   // R'_k,3 two planes below the chunk start
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            if (s < nslot) {
                T r[3], p[3], b[3];
                rk_at((I)(z0 - 2) * sxy + (I)cy[s] * nx + (I)cx[s], r, p, b);
                rk[s][2] = r[2];
            }
        }
    }
    double sdd0 = 0.0, suu0 = 0.0, sdd1 = 0.0, suu1 = 0.0;
    const int zstart = z0 >= 1 ? z0 - 1 : 0;
    const int zend = z1 + 1;
    for (int zi = zstart; zi <= zend; ++zi) {
        // ---- S1a: R'_k(zi) on the region, into LDS
        T rn[2][3] = {}, pn[2][3] = {}, bn[2][3] = {}, fn[2] = {};
        const bool s1 = zi < nz;
        if (s1) {
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                if (s < nslot) {
                    const I i = (I)zi * sxy + (I)cy[s] * nx + (I)cx[s];
                    rk_at(i, rn[s], pn[s], bn[s]);
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
                    const I i = (I)zi * sxy + (I)cy[s] * nx + (I)cx[s];
                    T b1r, b2r;
                    if (bx[s] > 0) {
                        b1r = Rk1[by[s]][bx[s] - 1];
                    } else {
                        T t[3], tp[3], tb[3];
                        rk_at(cx[s] > 0 ? i - 1 : i, t, tp, tb);
                        b1r = t[0];
                    }
                    if (by[s] > 0) {
                        b2r = Rk2[by[s] - 1][bx[s]];
                    } else {
                        T t[3], tp[3], tb[3];
                        rk_at(cy[s] > 0 ? i - nx : i, t, tp, tb);
                        b2r = t[1];
                    }
                    const T b1 = X[s] > 0 ? b1r : (T)0;
                    const T b2 = Y[s] > 0 ? b2r : (T)0;
                    const T b3 = zi > 0 ? rk[s][2] : (T)0;
                    T d = fn[s] - lam * (rn[s][0] + rn[s][1] + rn[s][2] - b1 - b2 - b3);
                    if (a.nonneg && d < (T)0) d = (T)0;
                    Dk[zi & 1][by[s]][bx[s]] = d;
                    if (s == 0 && own && zi >= z0 && zi < z1) {
                        const I io = (I)zi * sxy + (I)Y[0] * nx + (I)X[0];
                        if (a.write_d0) stnt<T>(a.D0, io, d);
                        if (a.check0) {
                            const T df = d - ldx<T>(a.Dprev, io);
                            sdd0 += (double)df * (double)df;
                            suu0 += (double)d * (double)d;
                        }
                    }
                }
            }
        }
        __syncthreads();
        // ---- S2: P_{k+1}(zc), R'_{k+1}(zc)
        const int zc = zi - 1;
        T rqn[2][3] = {};
        const bool s2 = zc >= (z0 >= 1 ? z0 - 1 : 0) && zc <= z1 && zc < nz && zc >= 0;
        if (s2) {
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                if (s < nslot) {
                    const int b = zc & 1;
                    const T d = Dk[b][by[s]][bx[s]];
                    T g1 = (X[s] == g.nx - 1) ? (T)0 : d - Dk[b][by[s]][bx[s] + 1];
                    T g2 = (Y[s] == g.ny - 1) ? (T)0 : d - Dk[b][by[s] + 1][bx[s]];
                    T g3 = (zc == nz - 1) ? (T)0 : d - Dk[b ^ 1][by[s]][bx[s]];
                    dtv_orth3(g1, g2, g3, bk[s][0], bk[s][1], bk[s][2]);
                    T p1 = rk[s][0] + step * g1;
                    T p2 = rk[s][1] + step * g2;
                    T p3 = rk[s][2] + step * g3;
                    fgp_project<3, T, ISO>(p1, p2, p3);
                    if (s == 0 && own && zc >= z0 && zc < z1) {
                        const I io = (I)zc * sxy + (I)Y[0] * nx + (I)X[0];
                        stnt<T>(a.Pn[0], io, p1);
                        stnt<T>(a.Pn[1], io, p2);
                        stnt<T>(a.Pn[2], io, p3);
                    }
                    rqn[s][0] = p1 + m1 * (p1 - pk[s][0]);
                    rqn[s][1] = p2 + m1 * (p2 - pk[s][1]);
                    rqn[s][2] = p3 + m1 * (p3 - pk[s][2]);
                    dtv_orth3(rqn[s][0], rqn[s][1], rqn[s][2], bk[s][0], bk[s][1], bk[s][2]);
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
                    const T b1 = X[s] > 0 ? b1r : (T)0;
                    const T b2 = Y[s] > 0 ? b2r : (T)0;
                    const T b3 = zc > 0 ? rq[s][2] : (T)0;
                    T d = fz[s] - lam * (rqn[s][0] + rqn[s][1] + rqn[s][2] - b1 - b2 - b3);
                    if (a.nonneg && d < (T)0) d = (T)0;
                    Dk1[zc & 1][by[s]][bx[s]] = d;
                    if (s == 0 && own && zc < z1) {
                        const I io = (I)zc * sxy + (I)Y[0] * nx + (I)X[0];
                        if (a.write_d1) stnt<T>(a.D1, io, d);
                        if (a.check1) {
                            const T df = d - Dk[zc & 1][by[0]][bx[0]];
                            sdd1 += (double)df * (double)df;
                            suu1 += (double)d * (double)d;
                        }
                    }
                }
            }
        }
        __syncthreads();
        // ---- S4: P_{k+2}(zq) on the tile
        const int zq = zi - 2;
        if (zq >= z0 && zq < z1 && own) {
            const int b = zq & 1;
            const T d = Dk1[b][by[0]][bx[0]];
            T g1 = (X[0] == g.nx - 1) ? (T)0 : d - Dk1[b][by[0]][bx[0] + 1];
            T g2 = (Y[0] == g.ny - 1) ? (T)0 : d - Dk1[b][by[0] + 1][bx[0]];
            T g3 = (zq == nz - 1) ? (T)0 : d - Dk1[b ^ 1][by[0]][bx[0]];
            dtv_orth3(g1, g2, g3, bq[0], bq[1], bq[2]);
            T p1 = rq[0][0] + step * g1;
            T p2 = rq[0][1] + step * g2;
            T p3 = rq[0][2] + step * g3;
            fgp_project<3, T, ISO>(p1, p2, p3);
            const I io = (I)zq * sxy + (I)Y[0] * nx + (I)X[0];
            stnt<T>(a.Pnn[0], io, p1);
            stnt<T>(a.Pnn[1], io, p2);
            stnt<T>(a.Pnn[2], io, p3);
        }
        // ---- shift the pipeline
#pragma unroll
        for (int c = 0; c < 3; ++c)
            if (s2) bq[c] = bk[0][c];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                if (s1) { rk[s][c] = rn[s][c]; pk[s][c] = pn[s][c]; bk[s][c] = bn[s][c]; }
                if (s2) rq[s][c] = rqn[s][c];
            }
            if (s1) fz[s] = fn[s];
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

template <int TY, typename T, bool ISO, typename I>
__global__ __launch_bounds__(TX * TY) void dtv2_iter_kernel(Dtv2Args<T> a) {
    if (solve_stopped(a.ctrl)) return;
    dtv2_body<TY, T, ISO, I>(a, tile_of_block(a.g, TY));
}

template <int TY, typename T>
static const void* dtv2_kernel_ptr(bool iso, bool small) {
    if (small)
        return iso ? (const void*)dtv2_iter_kernel<TY, T, true, uint32_t> : (const void*)dtv2_iter_kernel<TY, T, false, uint32_t>;
    return iso ? (const void*)dtv2_iter_kernel<TY, T, true, uint64_t> : (const void*)dtv2_iter_kernel<TY, T, false, uint64_t>;
}

template <int TY, typename T>
static void launch_dtv2(const Dtv2Args<T>& a, bool iso, bool small, hipStream_t s) {
    dim3 block(TX, TY), grid(a.g.nblocks);
    Dtv2Args<T> arg = a;
    void* args[] = {&arg};
    (void)hipLaunchKernel(dtv2_kernel_ptr<TY, T>(iso, small), grid, block, args, 0, s);
}

template <int DIM, int TY, typename T, typename I>
static const void* dtv_kernel_ptr(bool iso) {
    return iso ? (const void*)dtv_iter_kernel<DIM, TY, T, true, I> : (const void*)dtv_iter_kernel<DIM, TY, T, false, I>;
}

template <int DIM, int TY, typename T>
static void launch_dtv(const DtvArgs<T>& a, bool iso, bool small, hipStream_t s) {
    dim3 block(TX, TY), grid(a.g.nblocks);
    if (small) {
        if (iso) hipLaunchKernelGGL((dtv_iter_kernel<DIM, TY, T, true, uint32_t>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((dtv_iter_kernel<DIM, TY, T, false, uint32_t>), grid, block, 0, s, a);
    } else {
        if (iso) hipLaunchKernelGGL((dtv_iter_kernel<DIM, TY, T, true, uint64_t>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((dtv_iter_kernel<DIM, TY, T, false, uint64_t>), grid, block, 0, s, a);
    }
}

// Device-resident FGP-dTV solve.  d_out receives D of the final iteration
// (left untouched when iter == 0); with epsil != 0 its prior contents seed the
// first residual (FGP_dTV_core.c:78,131).
template <typename T>
int dtv_solve(const T* d_f, const T* d_ref, T* d_out, T lambda, int iter, T epsil, T eta, int method, int nonneg,
              int nx, int ny, int nz, const Opts& opts, hipStream_t s, int* iters_done, float* re) {
    const int dim = nz <= 1 ? 2 : 3;
    if (dim == 2) nz = 1;
    *iters_done = 0;
    *re = 0.f;
    if (iter <= 0) return 0;
    const size_t n = (size_t)nx * ny * nz;
    const bool small = !opts.wide_index && fits_u32(n, sizeof(T));
    const int ty = (opts.tile_y == 4 || dim == 2) ? 4 : 8;
    const bool iso = method == 0;
    const int width = (dim == 3 && opts.depth != 1 && iter > 1) ? 2 : 1;
    Geom g;
    if (dim == 2)
        g = make_geom(nx, ny, nz, 4, 1);
    else if (width == 2)
        // the pass evaluates about three planes below and above its chunk
        g = ty == 8 ? plan_geom(dtv2_kernel_ptr<8, T>(iso, small), TX * 8, nx, ny, nz, 8, opts.zchunk, 3.0)
                    : plan_geom(dtv2_kernel_ptr<4, T>(iso, small), TX * 4, nx, ny, nz, 4, opts.zchunk, 3.0);
    else if (ty == 8)
        g = plan_geom(small ? dtv_kernel_ptr<3, 8, T, uint32_t>(iso) : dtv_kernel_ptr<3, 8, T, uint64_t>(iso), TX * 8,
                      nx, ny, nz, 8, opts.zchunk);
    else
        g = plan_geom(small ? dtv_kernel_ptr<3, 4, T, uint32_t>(iso) : dtv_kernel_ptr<3, 4, T, uint64_t>(iso), TX * 4,
                      nx, ny, nz, 4, opts.zchunk);

    // P_j lives in buffer j mod nbuf (4 with two-iteration passes, which write
    // P_{k+1} and P_{k+2} while reading P_k and P_{k-1})
    const int nbuf = width == 2 ? 4 : 3;
    Arena ar;
    size_t oP[4][3], oB[3], oD[2], oCtrl, oPart;
    for (int b = 0; b < nbuf; ++b)
        for (int c = 0; c < dim; ++c) oP[b][c] = ar.reserve(n * sizeof(T));
    for (int c = 0; c < dim; ++c) oB[c] = ar.reserve(n * sizeof(T));
    for (int b = 0; b < 2; ++b) oD[b] = ar.reserve(n * sizeof(T));
    oCtrl = ar.reserve(sizeof(Ctrl));
    oPart = ar.reserve((size_t)g.nblocks * 4 * sizeof(double));
    if (ar.allocate() != 0) return -1;
    for (int c = 0; c < dim; ++c) {   // P_0 = P_{-1} = 0
        launch_zero(ar.at(oP[0][c]), n * sizeof(T), s);
        launch_zero(ar.at(oP[nbuf - 1][c]), n * sizeof(T), s);
    }
    auto Pbuf = [&](int j, int c) -> T* { return (T*)ar.at(oP[((j % nbuf) + nbuf) % nbuf][c < dim ? c : 0]); };
    T* B[3];
    for (int c = 0; c < 3; ++c) B[c] = (T*)ar.at(oB[c < dim ? c : 0]);
    {
        size_t blocks = (n + 255) / 256;
        if (blocks > 65536) blocks = 65536;
        if (dim == 3)
            hipLaunchKernelGGL((dtv_refgrad_kernel<3, T>), dim3((unsigned)blocks), dim3(256), 0, s, d_ref, B[0], B[1],
                               B[2], eta, nx, ny, nz);
        else
            hipLaunchKernelGGL((dtv_refgrad_kernel<2, T>), dim3((unsigned)blocks), dim3(256), 0, s, d_ref, B[0], B[1],
                               B[2], eta, nx, ny, nz);
    }
    const bool use_tol = epsil != (T)0;
    T* Dbuf[2] = {(T*)ar.at(oD[0]), (T*)ar.at(oD[1])};
    if (use_tol) launch_copy(Dbuf[1], d_out, n * sizeof(T), s);   // D_{-1} = caller's Output
    Ctrl* ctrl = (Ctrl*)ar.at(oCtrl);
    double* partials = (double*)ar.at(oPart);
    double* partials1 = partials + 2 * (size_t)g.nblocks;
    const int ce = opts.check_every;
    auto needs_d = [&](int k) { return k == iter - 1 || (use_tol && (k % ce == 0 || (k + 1) % ce == 0)); };
    const T stepc = (T)1 / ((dim == 3 ? (T)26 : (T)8) * lambda);

    // t_k with a double square root and a double sum rounded once
    // (FGP_dTV_core.c:95,149); mult[k] multiplies (P_k - P_{k-1}) in iteration k
    std::vector<T> mult(iter + 1);
    {
        T tk = 1, m = 0;
        for (int k = 0; k <= iter; ++k) {
            mult[k] = m;
            const T tkp1 = (T)((1.0 + sqrt((double)((T)1 + (T)4 * tk * tk))) * 0.5);
            m = (tk - (T)1) / tkp1;
            tk = tkp1;
        }
    }
    // one iteration k on the one-iteration kernel
    auto one = [&](int k, bool write_d, bool check, const Ctrl* c_ptr) {
        DtvArgs<T> a;
        a.f = d_f;
        for (int c = 0; c < 3; ++c) {
            a.B[c] = B[c];
            a.P[c] = Pbuf(k, c);
            a.Pp[c] = Pbuf(k - 1, c);
            a.Pn[c] = Pbuf(k + 1, c);
        }
        a.Dout = Dbuf[k & 1];
        a.Dprev = Dbuf[(k + 1) & 1];
        a.lambda = lambda;
        a.step = stepc;
        a.mult = mult[k];
        a.write_d = write_d;
        a.check = check;
        a.nonneg = nonneg != 0;
        a.partials = partials;
        a.ctrl = c_ptr;
        a.g = g;
        if (dim == 3) {
            if (ty == 8) launch_dtv<3, 8>(a, iso, small, s);
            else launch_dtv<3, 4>(a, iso, small, s);
        } else {
            launch_dtv<2, 4>(a, iso, small, s);
        }
    };
    auto step = [&](int k, int nk, bool check, bool check1) {
        if (nk == 2) {
            Dtv2Args<T> a;
            a.f = d_f;
            for (int c = 0; c < 3; ++c) {
                a.B[c] = B[c];
                a.Pm[c] = Pbuf(k - 1, c);
                a.P[c] = Pbuf(k, c);
                a.Pn[c] = Pbuf(k + 1, c);
                a.Pnn[c] = Pbuf(k + 2, c);
            }
            a.D0 = Dbuf[k & 1];
            a.D1 = Dbuf[(k + 1) & 1];
            a.Dprev = Dbuf[(k + 1) & 1];
            a.lambda = lambda;
            a.step = stepc;
            a.mult0 = mult[k];
            a.mult1 = mult[k + 1];
            // D_k of a pass's first iteration stays on chip (its check and
            // iteration k+1 read it from LDS); a stop there recomputes it
            a.write_d0 = 0;
            a.write_d1 = needs_d(k + 1) || check1;
            a.check0 = check;
            a.check1 = check1;
            a.nonneg = nonneg != 0;
            a.partials0 = partials;
            a.partials1 = partials1;
            a.ctrl = ctrl;
            a.g = g;
            if (ty == 8) launch_dtv2<8>(a, iso, small, s);
            else launch_dtv2<4>(a, iso, small, s);
            return;
        }
        one(k, needs_d(k) || check, check, ctrl);
    };
    int done = 0;
    if (drive_passes(iter, width, (float)epsil, opts, ctrl, partials, partials1, g.nblocks, s, step, &done, re) != 0)
        return -1;
    *iters_done = done;
    // a stop on the first iteration of a two-iteration pass: D_done from
    // P_done and P_{done-1}, left intact by the no-op passes after the stop
    if (done < iter && width == 2 && done % 2 == 0 && done + 1 < iter) one(done, true, false, nullptr);
    const int last = (done < iter) ? done : iter - 1;
    CILREG_CK(hipMemcpyAsync(d_out, Dbuf[last & 1], n * sizeof(T), hipMemcpyDeviceToDevice, s));
    CILREG_CK(hipStreamSynchronize(s));
    return 0;
}
template int dtv_solve<float>(const float*, const float*, float*, float, int, float, float, int, int, int, int, int,
                              const Opts&, hipStream_t, int*, float*);

}  // namespace cilreg
