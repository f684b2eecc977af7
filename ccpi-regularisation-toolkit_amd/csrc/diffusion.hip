// diffusion.hip -- the explicit-diffusion family: LLT-ROF, NDF (linear and
// nonlinear diffusion) and Diff4th (anisotropic fourth-order diffusion), one
// fused kernel per iteration each.
//
// Reference algorithms (src/Core/regularisers_CPU/):
//  * LLT-ROF, LLT_ROF_core.c:54-436: the ROF fluxes D_rof of ROF-TV with the
//    floor 1e-12, the second-order terms D_llt,d = u_dd / (|u_dd| + 1e-12) and
//      u += tau (lambda_rof div D_rof - lambda_llt sum_d d_dd D_llt,d - (u - f))
//    with mirrored neighbours (index -1 -> 1, N -> N-2).
//  * NDF, Diffusion_core.c:53-413: with e, w, n, s, u, d the differences to
//    the six mirrored neighbours and g the penalty's flux (Huber, Perona-Malik,
//    Tukey, threshold-linear, threshold-Huber; identity when sigma = 0),
//      u += tau (lambda sum g(.) - (u - f)),  sigma' = sigma / sqrt(2) in double.
//    The reference updates u in place inside its parallel loops, so its
//    result depends on the thread schedule; this kernel computes the
//    synchronous explicit step (every neighbour from the previous iterate).
//  * Diff4th, Diffus4th_order_core.c:44-265: the weighted Laplacian
//      W = c^2 V_norm + c V_orth,  c = 1 / (1 + |grad u|^2 / sigma^2)
//    from central differences (|grad u|^2 floored at the double 1e-7), then
//      u += tau (-lambda lap W - (u - f)).
// Every kernel keeps the reference's expressions, their order and their
// double promotions, so FP32 results equal the CPU oracle bit for bit.
//
// HBM traffic per voxel-iteration: read u, f; write u = 12 B for all three.
// The work per voxel is arithmetic (divisions and square roots on every
// flux), so these kernels are bounded by the vector ALUs rather than HBM.
//
// Fusion: a workgroup owns a 64 x TY column tile and marches through a chunk
// of z.  For each plane the quantities the update reads at x +- 1 / y +- 1
// (u for NDF, the LLT and ROF fluxes, W for Diff4th) go through LDS planes of
// (TY + 2) x (TX + 2) points covering the tile and a one-point ring; the
// values of the own column at z +- 1 stay in registers.  Mirrored neighbours
// of a boundary point always fall inside the tile or its ring.
#include "solvers.hpp"
#include "rof_kernels.hpp"

namespace cilreg {

namespace {

template <typename T>
struct DiffArgs {
    const T* f;
    const T* U;
    T* Uout;
    T lambda, lambda2;   // NDF / Diff4th: lambda; LLT-ROF: lambda_rof, lambda_llt
    T sigma, tau;        // NDF: sigma / sqrt(2); Diff4th: sigma^2
    int penalty;         // NDF: 0 linear, 1..5 the reference's penalties
    int check;
    double* partials;
    const Ctrl* ctrl;
    Geom g;
};

// The ring around a 64 x TY tile, one point per lane: lanes [0, TY) the -x
// column, [TY, 2TY) the +x column, [2TY, 2TY+TX) the -y row and
// [2TY+TX, 2TY+2TX) the +y row.  Slots are in the (TY+2) x (TX+2) plane
// whose origin is (x0 - 1, y0 - 1).
template <int TY>
struct Ring {
    static constexpr int N = 2 * TY + 2 * TX;
    bool lane, in;
    int x, y, sx, sy;   // point (clamped into the volume) and its slot
    __device__ Ring(const TileCoord& tc, const Geom& g, int tid, int nt) {
        const int hbase = (tc.lin & 1) ? nt - N : 0;
        const int h = tid - hbase;
        lane = h >= 0 && h < N;
        int rx, ry;
        if (h < TY) { rx = tc.x0 - 1; ry = tc.y0 + h; }
        else if (h < 2 * TY) { rx = tc.x0 + TX; ry = tc.y0 + h - TY; }
        else if (h < 2 * TY + TX) { rx = tc.x0 + h - 2 * TY; ry = tc.y0 - 1; }
        else { rx = tc.x0 + h - 2 * TY - TX; ry = tc.y0 + TY; }
        in = lane && rx >= 0 && ry >= 0 && rx < g.nx && ry < g.ny;
        sx = rx - tc.x0 + 1;
        sy = ry - tc.y0 + 1;
        x = in ? rx : 0;
        y = in ? ry : 0;
    }
};

// residual partials of the step u -> un (the check compares the iterate
// with the one before the step)
__device__ inline void add_resid(float un, float u, double& sdd, double& suu) {
    const float df = un - u;
    sdd += (double)df * (double)df;
    suu += (double)un * (double)un;
}

template <int TY>
__device__ inline void store_partials(const DiffArgs<float>& a, const TileCoord& tc, int tid, double sdd, double suu) {
    __shared__ double red[kRedSlots];
    block_sum2(sdd, suu, red);
    if (tid == 0) {
        a.partials[2 * tc.lin] = sdd;
        a.partials[2 * tc.lin + 1] = suu;
    }
}

// ---------------------------------------------------------------------------
// NDF
// ---------------------------------------------------------------------------
// The penalty's flux for a neighbour difference d (Diffusion_core.c:166-231);
// powf(x, 2) is x * x.  PEN is the penalty (0 linear), a template argument so
// each kernel holds one straight-line flux.
template <int PEN>
__device__ inline float ndf_flux(float d, float sigma) {
    const float ad = fabsf(d);
    const float sg = (float)((d > 0.f) - (d < 0.f));
    if constexpr (PEN == 1) return ad > sigma ? sg : d / sigma;
    if constexpr (PEN == 2) {
        const float q = d / sigma;
        return d / (1.0f + q * q);
    }
    if constexpr (PEN == 3) {
        if (ad <= sigma) {
            const float q = d / sigma;
            const float t = 1.0f - q * q;
            return d * (t * t);
        }
        return 0.0f;
    }
    if constexpr (PEN == 4) return ad > sigma ? 0.0f : d;
    if constexpr (PEN == 5) return ad <= 2.0f * sigma ? (ad > sigma ? sg : d / sigma) : 0.0f;
    return d;   // linear diffusion
}

template <int DIM, int TY, typename I, int PEN>
__device__ inline void ndf_body(const DiffArgs<float>& a, const TileCoord& tc) {
    const Geom& g = a.g;
    constexpr int NT = TX * TY;
    const int lx = threadIdx.x, ly = threadIdx.y, tid = ly * TX + lx;
    const int x = tc.x0 + lx, y = tc.y0 + ly;
    const bool own = x < g.nx && y < g.ny;
    const int ox = own ? x : 0, oy = own ? y : 0;
    const Ring<TY> r(tc, g, tid, NT);
    __shared__ float Us[TY + 2][TX + 2];
    const I sxy = (I)g.sxy, nx = (I)g.nx;
    const I col = (I)oy * nx + (I)ox, rcol = (I)r.y * nx + (I)r.x;
    // mirrored neighbour slots of the own point
    // threads outside the volume read their own slot (results not stored)
    const int sxp = own ? mirror_p(x, g.nx) - tc.x0 + 1 : lx + 1, sxm = own ? mirror_m(x, g.nx) - tc.x0 + 1 : lx + 1;
    const int syp = own ? mirror_p(y, g.ny) - tc.y0 + 1 : ly + 1, sym = own ? mirror_m(y, g.ny) - tc.y0 + 1 : ly + 1;

    float uc = ldx<float>(a.U, (I)tc.z0 * sxy + col);
    float um = 0.f;
    if (DIM == 3) um = ldx<float>(a.U, (I)mirror_m(tc.z0, g.nz) * sxy + col);
    double sdd = 0.0, suu = 0.0;
    for (int z = tc.z0; z < tc.z1; ++z) {
        float up = 0.f;
        if (DIM == 3) up = (z + 1 < g.nz) ? ldx<float>(a.U, (I)(z + 1) * sxy + col) : um;
        Us[ly + 1][lx + 1] = uc;
        if (r.lane) Us[r.sy][r.sx] = ldx<float>(a.U, (I)z * sxy + rcol);
        __syncthreads();
        if (own) {
            const I i = (I)z * sxy + col;
            float d[6];
            d[0] = Us[ly + 1][sxp] - uc;   // e
            d[1] = Us[ly + 1][sxm] - uc;   // w
            d[2] = Us[syp][lx + 1] - uc;   // n
            d[3] = Us[sym][lx + 1] - uc;   // s
            if (DIM == 3) {
                d[4] = up - uc;            // u
                d[5] = um - uc;            // d
            }
            if constexpr (PEN != 0) {
#pragma unroll
                for (int q = 0; q < (DIM == 3 ? 6 : 4); ++q) d[q] = ndf_flux<PEN>(d[q], a.sigma);
            }
            float s = d[0] + d[1] + d[2] + d[3];
            if (DIM == 3) s = s + d[4] + d[5];
            const float un = uc + a.tau * (a.lambda * s - (uc - ldx<float>(a.f, i)));
            stnt<float>(a.Uout, i, un);
            if (a.check) add_resid(un, uc, sdd, suu);
        }
        __syncthreads();
        um = uc;
        uc = up;
    }
    if (a.check) store_partials<TY>(a, tc, tid, sdd, suu);
}

// g(dw) from F = g(D), the flux a neighbour computed for its own difference
// D across the shared edge (penalty PEN >= 1, sigma > 0).  dw = -D exactly
// unless both are +0 (a - b = -(b - a) in round-to-nearest, except equal
// operands, which give +0 both ways), and every penalty is odd except where
// it returns the literal 0: there g(dw) = 0, on zero g(dw) = dw, elsewhere
// g(dw) = -F, bit for bit (NaN operands included: the literal-zero tests are
// written as the penalty's own).
template <int PEN>
__device__ inline float ndf_flux_mirror(float dw, float F, float sigma) {
    if (dw == 0.f) return dw;
    const float ad = fabsf(dw);
    if constexpr (PEN == 3) {
        if (!(ad <= sigma)) return 0.0f;
    }
    if constexpr (PEN == 4) {
        if (ad > sigma) return 0.0f;
    }
    if constexpr (PEN == 5) {
        if (!(ad <= 2.0f * sigma)) return 0.0f;
    }
    return -F;
}

// NDF with each edge's flux evaluated once: every point evaluates its +x,
// +y, +z fluxes (the mirrored ones at the far boundary), shares the in-plane
// ones through LDS (ring lanes add the tile's -x column and -y row) and
// carries the z one to the next plane; the -x, -y, -z fluxes are the
// neighbours' by ndf_flux_mirror, or at the near boundary (mirror index 1)
// the point's own +x, +y, +z flux.  Same terms, same sum order as ndf_body.
template <int DIM, int TY, typename I, int PEN>
__device__ inline void ndf_shared_body(const DiffArgs<float>& a, const TileCoord& tc) {
    const Geom& g = a.g;
    constexpr int NT = TX * TY;
    const int lx = threadIdx.x, ly = threadIdx.y, tid = ly * TX + lx;
    const int x = tc.x0 + lx, y = tc.y0 + ly;
    const bool own = x < g.nx && y < g.ny;
    const int ox = own ? x : 0, oy = own ? y : 0;
    const Ring<TY> r(tc, g, tid, NT);
    __shared__ float Us[TY + 2][TX + 2];
    __shared__ float FE[TY][TX + 1];   // [ly][lx + 1]: +x flux; column 0: x0 - 1
    __shared__ float FN[TY + 1][TX];   // [ly + 1][lx]: +y flux; row 0: y0 - 1
    const I sxy = (I)g.sxy, nx = (I)g.nx;
    const I col = (I)oy * nx + (I)ox, rcol = (I)r.y * nx + (I)r.x;
    // threads outside the volume read their own slot (results not stored)
    const int sxp = own ? mirror_p(x, g.nx) - tc.x0 + 1 : lx + 1, sxm = own ? mirror_m(x, g.nx) - tc.x0 + 1 : lx + 1;
    const int syp = own ? mirror_p(y, g.ny) - tc.y0 + 1 : ly + 1, sym = own ? mirror_m(y, g.ny) - tc.y0 + 1 : ly + 1;
    const float sigma = a.sigma;

    float uc = ldx<float>(a.U, (I)tc.z0 * sxy + col);
    float um = 0.f, fu_prev = 0.f;
    if (DIM == 3) {
        um = ldx<float>(a.U, (I)mirror_m(tc.z0, g.nz) * sxy + col);
        if (tc.z0 > 0) fu_prev = ndf_flux<PEN>(uc - um, sigma);   // +z flux of plane z0 - 1
    }
    double sdd = 0.0, suu = 0.0;
    for (int z = tc.z0; z < tc.z1; ++z) {
        float up = 0.f;
        if (DIM == 3) up = (z + 1 < g.nz) ? ldx<float>(a.U, (I)(z + 1) * sxy + col) : um;
        Us[ly + 1][lx + 1] = uc;
        if (r.lane) Us[r.sy][r.sx] = ldx<float>(a.U, (I)z * sxy + rcol);
        __syncthreads();
        const float fe = ndf_flux<PEN>(Us[ly + 1][sxp] - uc, sigma);
        const float fn = ndf_flux<PEN>(Us[syp][lx + 1] - uc, sigma);
        const float dw = Us[ly + 1][sxm] - uc, ds = Us[sym][lx + 1] - uc;
        float fu = 0.f;
        if (DIM == 3) fu = ndf_flux<PEN>(up - uc, sigma);
        FE[ly][lx + 1] = fe;
        FN[ly + 1][lx] = fn;
        if (r.lane) {
            if (r.sx == 0) FE[r.sy - 1][0] = ndf_flux<PEN>(Us[r.sy][1] - Us[r.sy][0], sigma);
            else if (r.sy == 0) FN[0][r.sx - 1] = ndf_flux<PEN>(Us[1][r.sx] - Us[0][r.sx], sigma);
        }
        __syncthreads();
        if (own) {
            const I i = (I)z * sxy + col;
            const float fw = x == 0 ? fe : ndf_flux_mirror<PEN>(dw, FE[ly][lx], sigma);
            const float fs = y == 0 ? fn : ndf_flux_mirror<PEN>(ds, FN[ly][lx], sigma);
            float s = fe + fw + fn + fs;
            if (DIM == 3) {
                const float fd = z == 0 ? fu : ndf_flux_mirror<PEN>(um - uc, fu_prev, sigma);
                s = s + fu + fd;
            }
            const float un = uc + a.tau * (a.lambda * s - (uc - ldx<float>(a.f, i)));
            stnt<float>(a.Uout, i, un);
            if (a.check) add_resid(un, uc, sdd, suu);
        }
        fu_prev = fu;
        um = uc;
        uc = up;
    }
    if (a.check) store_partials<TY>(a, tc, tid, sdd, suu);
}

// SH: the shared-flux body (penalties 1-5 with sigma > 0).
template <int DIM, int TY, typename I, int PEN, bool SH>
__global__ __launch_bounds__(TX * TY) void ndf_iter_kernel(DiffArgs<float> a) {
    if (solve_stopped(a.ctrl)) return;
    if constexpr (SH) ndf_shared_body<DIM, TY, I, PEN>(a, tile_of_block(a.g, TY));
    else ndf_body<DIM, TY, I, PEN>(a, tile_of_block(a.g, TY));
}

// ---------------------------------------------------------------------------
// LLT-ROF
// ---------------------------------------------------------------------------
// u_dd / (|u_dd| + 1e-12), the denominator rounded from double (LLT_ROF_core.c:145-152)
__device__ inline float llt_ratio(float dd) { return dd / (float)(fabs((double)dd) + 1.0e-12); }

// ROF fluxes (floor 1e-12) and the x / y second-order terms at a point
template <int DIM, typename I>
__device__ inline void llt_point(const DiffArgs<float>& a, int px, int py, int z, float& r1, float& r2, float& r3,
                                 float& l1, float& l2) {
    const Geom& g = a.g;
    const I sxy = (I)g.sxy, nx = (I)g.nx;
    const I zo = (I)z * sxy, c = zo + (I)py * nx + (I)px;
    const float uc = ldx<float>(a.U, c);
    const float uip = ldx<float>(a.U, zo + (I)py * nx + (I)mirror_p(px, g.nx));
    const float uim = ldx<float>(a.U, zo + (I)py * nx + (I)mirror_m(px, g.nx));
    const float ujp = ldx<float>(a.U, zo + (I)mirror_p(py, g.ny) * nx + (I)px);
    const float ujm = ldx<float>(a.U, zo + (I)mirror_m(py, g.ny) * nx + (I)px);
    float ukp = 0.f, ukm = 0.f;
    if (DIM == 3) {
        ukp = ldx<float>(a.U, (I)mirror_p(z, g.nz) * sxy + (I)py * nx + (I)px);
        ukm = ldx<float>(a.U, (I)mirror_m(z, g.nz) * sxy + (I)py * nx + (I)px);
    }
    rof_flux_math<DIM, float>(uc, uip, uim, ujp, ujm, ukp, ukm, r1, r2, r3, 1.0e-12);
    l1 = llt_ratio(uip - 2.0f * uc + uim);
    l2 = llt_ratio(ujp - 2.0f * uc + ujm);
}

// the z second-order term of the own column at plane z
template <typename I>
__device__ inline float llt_l3(const DiffArgs<float>& a, I col, int z) {
    const Geom& g = a.g;
    const I sxy = (I)g.sxy;
    const float uc = ldx<float>(a.U, (I)z * sxy + col);
    const float up = ldx<float>(a.U, (I)mirror_p(z, g.nz) * sxy + col);
    const float um = ldx<float>(a.U, (I)mirror_m(z, g.nz) * sxy + col);
    return llt_ratio(up - 2.0f * uc + um);
}

template <int DIM, int TY, typename I>
__device__ inline void llt_body(const DiffArgs<float>& a, const TileCoord& tc) {
    const Geom& g = a.g;
    constexpr int NT = TX * TY;
    const int lx = threadIdx.x, ly = threadIdx.y, tid = ly * TX + lx;
    const int x = tc.x0 + lx, y = tc.y0 + ly;
    const bool own = x < g.nx && y < g.ny;
    const int ox = own ? x : 0, oy = own ? y : 0;
    const Ring<TY> r(tc, g, tid, NT);
    __shared__ float L1s[TY + 2][TX + 2], L2s[TY + 2][TX + 2], R1s[TY + 2][TX + 2], R2s[TY + 2][TX + 2];
    const I sxy = (I)g.sxy, nx = (I)g.nx;
    const I col = (I)oy * nx + (I)ox;
    // threads outside the volume read their own slot (results not stored)
    const int sxp = own ? mirror_p(x, g.nx) - tc.x0 + 1 : lx + 1, sxm = own ? mirror_m(x, g.nx) - tc.x0 + 1 : lx + 1;
    const int syp = own ? mirror_p(y, g.ny) - tc.y0 + 1 : ly + 1, sym = own ? mirror_m(y, g.ny) - tc.y0 + 1 : ly + 1;

    float r3m = 0.f, l3m = 0.f, l3c = 0.f;   // D_rof,3 and D_llt,3 of the own column below / here
    if (DIM == 3) {
        float t1, t2, t4, t5;
        llt_point<DIM, I>(a, ox, oy, mirror_m(tc.z0, g.nz), t1, t2, r3m, t4, t5);
        l3m = llt_l3<I>(a, col, mirror_m(tc.z0, g.nz));
        l3c = llt_l3<I>(a, col, tc.z0);
    }
    double sdd = 0.0, suu = 0.0;
    for (int z = tc.z0; z < tc.z1; ++z) {
        float r1, r2, r3, l1, l2;
        llt_point<DIM, I>(a, ox, oy, z, r1, r2, r3, l1, l2);
        float l3p = 0.f;
        if (DIM == 3) l3p = (z + 1 < g.nz) ? llt_l3<I>(a, col, z + 1) : l3m;
        L1s[ly + 1][lx + 1] = l1;
        L2s[ly + 1][lx + 1] = l2;
        R1s[ly + 1][lx + 1] = r1;
        R2s[ly + 1][lx + 1] = r2;
        if (r.lane) {
            // only what the tile reads at the lane's slot: L1 on the columns
            // and D2 on the left one, L2 on the rows and D1 on the upper one
            const I zo = (I)z * sxy, rc = zo + (I)r.y * nx + (I)r.x;
            const float uc = ldx<float>(a.U, rc);
            const float uip = ldx<float>(a.U, zo + (I)r.y * nx + (I)mirror_p(r.x, g.nx));
            const float uim = ldx<float>(a.U, zo + (I)r.y * nx + (I)mirror_m(r.x, g.nx));
            const float ujp = ldx<float>(a.U, zo + (I)mirror_p(r.y, g.ny) * nx + (I)r.x);
            const float ujm = ldx<float>(a.U, zo + (I)mirror_m(r.y, g.ny) * nx + (I)r.x);
            float mk = 0.f;
            if (DIM == 3) {
                const float ukp = ldx<float>(a.U, (I)mirror_p(z, g.nz) * sxy + (I)r.y * nx + (I)r.x);
                const float ukm = ldx<float>(a.U, (I)mirror_m(z, g.nz) * sxy + (I)r.y * nx + (I)r.x);
                mk = Mixed<float>::minmod_sq(ukp - uc, uc - ukm);
            }
            if (r.sx == 0 || r.sx == TX + 1) {
                L1s[r.sy][r.sx] = llt_ratio(uip - 2.0f * uc + uim);
                if (r.sx == 0)
                    R2s[r.sy][r.sx] = rof_d2<DIM, float>(uip - uc, Mixed<float>::minmod_sq(ujp - uc, uc - ujm), mk,
                                                          1.0e-12);
            } else {
                L2s[r.sy][r.sx] = llt_ratio(ujp - 2.0f * uc + ujm);
                if (r.sy == 0)
                    R1s[r.sy][r.sx] = rof_d1<DIM, float>(ujp - uc, Mixed<float>::minmod_sq(uip - uc, uc - uim), mk,
                                                          1.0e-12);
            }
        }
        __syncthreads();
        if (own) {
            const I i = (I)z * sxy + col;
            const float uc = ldx<float>(a.U, i);
            float lap = (L1s[ly + 1][sxp] - 2.0f * l1 + L1s[ly + 1][sxm]) + (L2s[syp][lx + 1] - 2.0f * l2 + L2s[sym][lx + 1]);
            float dv = (r1 - R1s[sym][lx + 1]) + (r2 - R2s[ly + 1][sxm]);
            if (DIM == 3) {
                lap = lap + (l3p - 2.0f * l3c + l3m);
                dv = dv + (r3 - r3m);
            }
            const float un = uc + a.tau * (a.lambda * dv - a.lambda2 * lap - (uc - ldx<float>(a.f, i)));
            stnt<float>(a.Uout, i, un);
            if (a.check) add_resid(un, uc, sdd, suu);
        }
        __syncthreads();
        r3m = r3;
        l3m = l3c;
        l3c = l3p;
    }
    if (a.check) store_partials<TY>(a, tc, tid, sdd, suu);
}

template <int DIM, int TY, typename I>
__global__ __launch_bounds__(TX * TY) void llt_iter_kernel(DiffArgs<float> a) {
    if (solve_stopped(a.ctrl)) return;
    llt_body<DIM, TY, I>(a, tile_of_block(a.g, TY));
}

// ---------------------------------------------------------------------------
// Diff4th
// ---------------------------------------------------------------------------
// The weighted Laplacian at (px, py, z) (Weighted_Laplc2D / 3D,
// Diffus4th_order_core.c:103-148, :176-235), sigma2 = sigma^2.
template <int DIM, typename I>
__device__ inline float d4_w(const DiffArgs<float>& a, int px, int py, int z) {
    const Geom& g = a.g;
    const I sxy = (I)g.sxy, nx = (I)g.nx;
    const I x1 = (I)mirror_p(px, g.nx), x2 = (I)mirror_m(px, g.nx);
    const I y1 = (I)mirror_p(py, g.ny) * nx, y2 = (I)mirror_m(py, g.ny) * nx, yc = (I)py * nx;
    const I zc = (I)z * sxy;
    auto U = [&](I zo, I yo, I xo) { return ldx<float>(a.U, zo + yo + xo); };
    const float uc = U(zc, yc, (I)px);
    const float ux2 = U(zc, yc, x2), ux1 = U(zc, yc, x1);
    const float uy2 = U(zc, y2, (I)px), uy1 = U(zc, y1, (I)px);
    const float gx = 0.5f * (ux2 - ux1);
    const float gy = 0.5f * (uy2 - uy1);
    const float gxs = gx * gx, gys = gy * gy;
    const float gxx = ux2 + ux1 - 2.0f * uc;
    const float gyy = uy2 + uy1 - 2.0f * uc;
    const float gxy = 0.25f * (U(zc, y2, x2) + U(zc, y1, x1) - U(zc, y1, x2) - U(zc, y2, x1));
    const float xy2 = 2.0f * gx * gy * gxy;
    float den, nn, no;
    if (DIM == 3) {
        const I z1 = (I)mirror_p(z, g.nz) * sxy, z2 = (I)mirror_m(z, g.nz) * sxy;
        const float uz2 = U(z2, yc, (I)px), uz1 = U(z1, yc, (I)px);
        const float gz = 0.5f * (uz2 - uz1);
        const float gzs = gz * gz;
        const float gzz = uz2 + uz1 - 2.0f * uc;
        const float gxz = 0.25f * (U(z2, yc, x2) - U(z2, yc, x1) - U(z1, yc, x2) + U(z1, yc, x1));
        const float gyz = 0.25f * (U(z2, y2, (I)px) - U(z2, y1, (I)px) - U(z1, y2, (I)px) + U(z1, y1, (I)px));
        const float xz1 = 2.0f * gx * gz * gxz, yz2 = 2.0f * gy * gz * gyz;
        den = gxs + gys + gzs;
        nn = gxx * gxs + gyy * gys + gzz * gzs + xy2 + xz1 + yz2;
        no = (gys + gzs) * gxx + (gxs + gzs) * gyy + (gxs + gys) * gzz - xy2 - xz1 - yz2;
    } else {
        den = gxs + gys;
        nn = gxx * gxs + xy2 + gyy * gys;
        no = gxx * gys - xy2 + gyy * gxs;
    }
    float vn, vo;
    if ((double)den <= 1.0e-7) {   // the double floor, Diffus4th_order_core.c:23,133
        vn = (float)((double)nn / 1.0e-7);
        vo = (float)((double)no / 1.0e-7);
    } else {
        vn = nn / den;
        vo = no / den;
    }
    const float c = 1.0f / (1.0f + den / a.sigma);
    return c * c * vn + c * vo;
}

template <int DIM, int TY, typename I>
__device__ inline void d4_body(const DiffArgs<float>& a, const TileCoord& tc) {
    const Geom& g = a.g;
    constexpr int NT = TX * TY;
    const int lx = threadIdx.x, ly = threadIdx.y, tid = ly * TX + lx;
    const int x = tc.x0 + lx, y = tc.y0 + ly;
    const bool own = x < g.nx && y < g.ny;
    const int ox = own ? x : 0, oy = own ? y : 0;
    const Ring<TY> r(tc, g, tid, NT);
    __shared__ float Ws[TY + 2][TX + 2];
    const I sxy = (I)g.sxy, nx = (I)g.nx;
    const I col = (I)oy * nx + (I)ox;
    // threads outside the volume read their own slot (results not stored)
    const int sxp = own ? mirror_p(x, g.nx) - tc.x0 + 1 : lx + 1, sxm = own ? mirror_m(x, g.nx) - tc.x0 + 1 : lx + 1;
    const int syp = own ? mirror_p(y, g.ny) - tc.y0 + 1 : ly + 1, sym = own ? mirror_m(y, g.ny) - tc.y0 + 1 : ly + 1;

    float wm = 0.f;   // W of the own column at the mirrored z - 1
    if (DIM == 3) wm = d4_w<DIM, I>(a, ox, oy, mirror_m(tc.z0, g.nz));
    float wc = d4_w<DIM, I>(a, ox, oy, tc.z0);
    double sdd = 0.0, suu = 0.0;
    for (int z = tc.z0; z < tc.z1; ++z) {
        Ws[ly + 1][lx + 1] = wc;
        if (r.lane) Ws[r.sy][r.sx] = d4_w<DIM, I>(a, r.x, r.y, z);
        float wp = 0.f;
        if (DIM == 3) wp = (z + 1 < g.nz) ? d4_w<DIM, I>(a, ox, oy, z + 1) : wm;
        __syncthreads();
        if (own) {
            const I i = (I)z * sxy + col;
            const float uc = ldx<float>(a.U, i);
            float s = (Ws[ly + 1][sxm] + Ws[ly + 1][sxp] - 2.0f * wc) + (Ws[sym][lx + 1] + Ws[syp][lx + 1] - 2.0f * wc);
            if (DIM == 3) s = s + (wm + wp - 2.0f * wc);
            const float un = uc + a.tau * (-a.lambda * s - (uc - ldx<float>(a.f, i)));
            stnt<float>(a.Uout, i, un);
            if (a.check) add_resid(un, uc, sdd, suu);
        }
        __syncthreads();
        wm = wc;
        wc = wp;
    }
    if (a.check) store_partials<TY>(a, tc, tid, sdd, suu);
}

template <int DIM, int TY, typename I>
__global__ __launch_bounds__(TX * TY) void d4_iter_kernel(DiffArgs<float> a) {
    if (solve_stopped(a.ctrl)) return;
    d4_body<DIM, TY, I>(a, tile_of_block(a.g, TY));
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
enum Method { M_LLT, M_NDF, M_D4 };

template <int DIM, int TY, typename I>
const void* kernel_of(Method m, int penalty, bool share) {
    switch (m) {
        case M_LLT: return (const void*)llt_iter_kernel<DIM, TY, I>;
        case M_NDF:
            switch (penalty) {
                case 1: return share ? (const void*)ndf_iter_kernel<DIM, TY, I, 1, true>
                                     : (const void*)ndf_iter_kernel<DIM, TY, I, 1, false>;
                case 2: return share ? (const void*)ndf_iter_kernel<DIM, TY, I, 2, true>
                                     : (const void*)ndf_iter_kernel<DIM, TY, I, 2, false>;
                case 3: return share ? (const void*)ndf_iter_kernel<DIM, TY, I, 3, true>
                                     : (const void*)ndf_iter_kernel<DIM, TY, I, 3, false>;
                case 4: return share ? (const void*)ndf_iter_kernel<DIM, TY, I, 4, true>
                                     : (const void*)ndf_iter_kernel<DIM, TY, I, 4, false>;
                case 5: return share ? (const void*)ndf_iter_kernel<DIM, TY, I, 5, true>
                                     : (const void*)ndf_iter_kernel<DIM, TY, I, 5, false>;
                default: return (const void*)ndf_iter_kernel<DIM, TY, I, 0, false>;
            }
        default: return (const void*)d4_iter_kernel<DIM, TY, I>;
    }
}

template <int DIM, int TY>
void launch(Method m, const DiffArgs<float>& a, bool small, hipStream_t s) {
    dim3 block(TX, TY), grid(a.g.nblocks);
    const void* k = small ? kernel_of<DIM, TY, uint32_t>(m, a.penalty, a.sigma > 0.f) : kernel_of<DIM, TY, uint64_t>(m, a.penalty, a.sigma > 0.f);
    DiffArgs<float> arg = a;
    void* args[] = {&arg};
    (void)hipLaunchKernel(k, grid, block, args, 0, s);
}

// One explicit solve: u_0 = f, `iter` steps of method m, the reference's
// stopping rule; d_out receives the last iterate (f itself when iter <= 0).
int diffusion_solve(Method m, const float* d_f, float* d_out, const DiffArgs<float>& proto, int iter, float epsil,
                    int nx, int ny, int nz, const Opts& opts, hipStream_t s, int* iters_done, float* re) {
    if (nz < 1) nz = 1;
    const int dim = nz > 1 ? 3 : 2;
    const size_t n = (size_t)nx * ny * nz;
    *iters_done = 0;
    *re = 0.f;
    if (iter <= 0) {   // Output = Input (copyIm before the loop)
        CILREG_CK(hipMemcpyAsync(d_out, d_f, n * sizeof(float), hipMemcpyDeviceToDevice, s));
        CILREG_CK(hipStreamSynchronize(s));
        return 0;
    }
    const bool small = !opts.wide_index && fits_u32(n, sizeof(float));
    const int ty = (opts.tile_y == 4 || dim == 2) ? 4 : 8;
    Geom g;
    if (dim == 2)
        g = make_geom(nx, ny, nz, 4, 1);
    else if (ty == 8)
        g = plan_geom(small ? kernel_of<3, 8, uint32_t>(m, proto.penalty, proto.sigma > 0.f)
                            : kernel_of<3, 8, uint64_t>(m, proto.penalty, proto.sigma > 0.f),
                      TX * 8, nx, ny, nz, 8, opts.zchunk);
    else
        g = plan_geom(small ? kernel_of<3, 4, uint32_t>(m, proto.penalty, proto.sigma > 0.f)
                            : kernel_of<3, 4, uint64_t>(m, proto.penalty, proto.sigma > 0.f),
                      TX * 4, nx, ny, nz, 4, opts.zchunk);
    Arena ar;
    size_t oU[2], oCtrl, oPart;
    oU[0] = ar.reserve(n * sizeof(float));
    oU[1] = ar.reserve(n * sizeof(float));
    oCtrl = ar.reserve(sizeof(Ctrl));
    oPart = ar.reserve((size_t)g.nblocks * 2 * sizeof(double));
    if (ar.allocate() != 0) return -1;
    launch_copy(ar.at(oU[0]), d_f, n * sizeof(float), s);
    Ctrl* ctrl = (Ctrl*)ar.at(oCtrl);
    double* partials = (double*)ar.at(oPart);
    auto step = [&](int k, bool check) {
        DiffArgs<float> a = proto;
        a.f = d_f;
        a.U = (const float*)ar.at(oU[k & 1]);
        a.Uout = (float*)ar.at(oU[(k + 1) & 1]);
        a.check = check;
        a.partials = partials;
        a.ctrl = ctrl;
        a.g = g;
        if (dim == 3) {
            if (ty == 8) launch<3, 8>(m, a, small, s);
            else launch<3, 4>(m, a, small, s);
        } else {
            launch<2, 4>(m, a, small, s);
        }
    };
    int done = 0;
    if (drive_iterations(iter, epsil, opts, ctrl, partials, g.nblocks, s, step, &done, re) != 0) return -1;
    *iters_done = done;
    const int last = (done < iter) ? done : iter - 1;
    CILREG_CK(hipMemcpyAsync(d_out, ar.at(oU[(last + 1) & 1]), n * sizeof(float), hipMemcpyDeviceToDevice, s));
    CILREG_CK(hipStreamSynchronize(s));
    return 0;
}

}  // namespace

int llt_solve(const float* d_f, float* d_out, float lambda_rof, float lambda_llt, int iter, float tau, float epsil,
              int nx, int ny, int nz, const Opts& opts, hipStream_t s, int* iters_done, float* re) {
    DiffArgs<float> p{};
    p.lambda = lambda_rof;
    p.lambda2 = lambda_llt;
    p.tau = tau;
    return diffusion_solve(M_LLT, d_f, d_out, p, iter, epsil, nx, ny, nz, opts, s, iters_done, re);
}

int ndf_solve(const float* d_f, float* d_out, float lambda, float sigma, int iter, float tau, int penalty, float epsil,
              int nx, int ny, int nz, const Opts& opts, hipStream_t s, int* iters_done, float* re) {
    if (sigma != 0.0f && (penalty < 1 || penalty > 5)) {
        set_error("NDF: penalty type %d is not one of 1 (Huber), 2 (Perona-Malik), 3 (Tukey), 4, 5", penalty);
        return -1;
    }
    DiffArgs<float> p{};
    p.lambda = lambda;
    // sigma / sqrt(2) with a double square root and division (Diffusion_core.c:57)
    p.sigma = (float)((double)sigma / sqrt(2.0));
    p.tau = tau;
    p.penalty = sigma == 0.0f ? 0 : penalty;   // sigma = 0: linear diffusion (:74,79)
    return diffusion_solve(M_NDF, d_f, d_out, p, iter, epsil, nx, ny, nz, opts, s, iters_done, re);
}

int d4_solve(const float* d_f, float* d_out, float lambda, float sigma, int iter, float tau, float epsil, int nx,
             int ny, int nz, const Opts& opts, hipStream_t s, int* iters_done, float* re) {
    DiffArgs<float> p{};
    p.lambda = lambda;
    p.sigma = sigma * sigma;   // Diffus4th_order_core.c:51
    p.tau = tau;
    return diffusion_solve(M_D4, d_f, d_out, p, iter, epsil, nx, ny, nz, opts, s, iters_done, re);
}

}  // namespace cilreg
