// pd.hip -- PD-TV (Chambolle-Pock primal-dual), one fused kernel per iteration.
//
// Reference algorithm: src/Core/regularisers_CPU/PD_TV_core.c:40-252.  State:
// the extrapolated primal u_bar and the dual P.  Per iteration:
//   P   = proj(P + sigma * grad u_bar)   forward differences, mirrored at the
//                                        last index (DualP3D, :212-229)
//   U0  = max(u_bar, 0) if nonneg        (:120)
//   U   = (U0 - tau * div P + lt * f) / (1.0 + lt)   div with P(-1) = 0,
//                                        double denominator (DivProj3D, :232-252)
//   residual check |U - U0| / |U| every 5th iteration (:131-141)
//   u_bar = U + theta (U - U0)           (getX, :197-205)
// The output is u_bar, or U when the stopping rule fires (the check precedes
// the extrapolation in the reference loop).
//
// HBM traffic per voxel-iteration (FP32): read u_bar, f, P(3); write u_bar,
// P(3) = 36 B (+4 B on check iterations for the break candidate U).
//
// Fusion: the updated P of plane z is computed on the tile plus its -x / -y
// halo into LDS; the primal update of plane z reads it there and P3 of plane
// z-1 from a register.
#include "solvers.hpp"

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
__global__ __launch_bounds__(TX * TY) void pd_iter_kernel(PdArgs<T, S> a) {
    if (solve_stopped(a.ctrl)) return;
    const Geom& g = a.g;
    const TileCoord tc = tile_of_block(g, TY);
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

template <int DIM, int TY, typename T, typename S, typename I>
static const void* pd_kernel_ptr(bool iso) {
    return iso ? (const void*)pd_iter_kernel<DIM, TY, T, S, true, I> : (const void*)pd_iter_kernel<DIM, TY, T, S, false, I>;
}

template <int DIM, int TY, typename T, typename S>
static void launch_pd(const PdArgs<T, S>& a, bool iso, bool small, hipStream_t s) {
    dim3 block(TX, TY), grid(a.g.nblocks);
    if (small) {
        if (iso) hipLaunchKernelGGL((pd_iter_kernel<DIM, TY, T, S, true, uint32_t>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((pd_iter_kernel<DIM, TY, T, S, false, uint32_t>), grid, block, 0, s, a);
    } else {
        if (iso) hipLaunchKernelGGL((pd_iter_kernel<DIM, TY, T, S, true, uint64_t>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((pd_iter_kernel<DIM, TY, T, S, false, uint64_t>), grid, block, 0, s, a);
    }
}

template <typename T, typename S>
int pd_solve(const T* d_f, T* d_out, T lambda, int iter, T epsil, T lipschitz, int method, int nonneg, int nx, int ny,
             int nz, const Opts& opts, hipStream_t s, int* iters_done, float* re) {
    const int dim = nz <= 1 ? 2 : 3;
    if (dim == 2) nz = 1;
    const size_t n = (size_t)nx * ny * nz;
    *iters_done = 0;
    *re = 0.f;
    if (iter <= 0) {   // U = Input (PD_TV_core.c:59)
        CILREG_CK(hipMemcpyAsync(d_out, d_f, n * sizeof(T), hipMemcpyDeviceToDevice, s));
        CILREG_CK(hipStreamSynchronize(s));
        return 0;
    }
    const bool small = !opts.wide_index && fits_u32(n, sizeof(T) > sizeof(S) ? sizeof(T) : sizeof(S));
    const int ty = (opts.tile_y == 4 || dim == 2) ? 4 : 8;
    const bool iso = method == 0;
    Geom g;
    if (dim == 2)
        g = make_geom(nx, ny, nz, 4, 1);
    else if (ty == 8)
        g = plan_geom(small ? pd_kernel_ptr<3, 8, T, S, uint32_t>(iso) : pd_kernel_ptr<3, 8, T, S, uint64_t>(iso),
                      TX * 8, nx, ny, nz, 8, opts.zchunk);
    else
        g = plan_geom(small ? pd_kernel_ptr<3, 4, T, S, uint32_t>(iso) : pd_kernel_ptr<3, 4, T, S, uint64_t>(iso),
                      TX * 4, nx, ny, nz, 4, opts.zchunk);

    Arena ar;
    size_t oU[2], oP[2][3], oB, oCtrl, oPart;
    for (int b = 0; b < 2; ++b) oU[b] = ar.reserve(n * sizeof(T));
    for (int b = 0; b < 2; ++b)
        for (int c = 0; c < dim; ++c) oP[b][c] = ar.reserve(n * sizeof(S));
    oB = ar.reserve(n * sizeof(T));
    oCtrl = ar.reserve(sizeof(Ctrl));
    oPart = ar.reserve((size_t)g.nblocks * 2 * sizeof(double));
    if (ar.allocate() != 0) return -1;
    launch_copy(ar.at(oU[0]), d_f, n * sizeof(T), s);
    for (int c = 0; c < dim; ++c) launch_zero(ar.at(oP[0][c]), n * sizeof(S), s);
    Ctrl* ctrl = (Ctrl*)ar.at(oCtrl);
    double* partials = (double*)ar.at(oPart);

    // step sizes in the working precision (PD_TV_core.c:50-53)
    const T tau = lambda * (T)0.1;
    const T sigma = (T)(1.0 / (double)(lipschitz * tau));
    const T lt = tau / lambda;
    auto step = [&](int k, bool check) {
        PdArgs<T, S> a;
        const int cur = k & 1, nxt = cur ^ 1;
        a.f = d_f;
        a.U = (const T*)ar.at(oU[cur]);
        a.Uout = (T*)ar.at(oU[nxt]);
        for (int c = 0; c < 3; ++c) {
            const int cc = c < dim ? c : 0;
            a.P[c] = (const S*)ar.at(oP[cur][cc]);
            a.Pout[c] = (S*)ar.at(oP[nxt][cc]);
        }
        a.Ubrk = (T*)ar.at(oB);
        a.sigma = sigma;
        a.tau = tau;
        a.lt = lt;
        a.theta = 1;
        a.nonneg = nonneg != 0;
        a.check = check;
        a.partials = partials;
        a.ctrl = ctrl;
        a.g = g;
        if (dim == 3) {
            if (ty == 8) launch_pd<3, 8>(a, iso, small, s);
            else launch_pd<3, 4>(a, iso, small, s);
        } else {
            launch_pd<2, 4>(a, iso, small, s);
        }
    };
    int done = 0;
    if (drive_iterations(iter, (float)epsil, opts, ctrl, partials, g.nblocks, s, step, &done, re) != 0) return -1;
    *iters_done = done;
    const T* src = (done < iter) ? (const T*)ar.at(oB) : (const T*)ar.at(oU[iter & 1]);
    CILREG_CK(hipMemcpyAsync(d_out, src, n * sizeof(T), hipMemcpyDeviceToDevice, s));
    CILREG_CK(hipStreamSynchronize(s));
    return 0;
}

template int pd_solve<float, float>(const float*, float*, float, int, float, float, int, int, int, int, int,
                                    const Opts&, hipStream_t, int*, float*);
template int pd_solve<float, bf16_t>(const float*, float*, float, int, float, float, int, int, int, int, int,
                                     const Opts&, hipStream_t, int*, float*);
template int pd_solve<float, _Float16>(const float*, float*, float, int, float, float, int, int, int, int, int,
                                       const Opts&, hipStream_t, int*, float*);

}  // namespace cilreg
