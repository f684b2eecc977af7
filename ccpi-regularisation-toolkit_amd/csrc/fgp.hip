// fgp.hip -- FGP-TV (FISTA on the dual of the ROF model, Beck & Teboulle),
// one fused kernel per iteration.
//
// Reference algorithm: src/Core/regularisers_CPU/FGP_TV_core.c:40-269 with the
// projection of utils.c:149-213.  Per iteration k, with multiplier
// m_k = (t_{k-1} - 1) / t_k:
//   R   = P_k + m_k (P_k - P_{k-1})                 (Rupd, :257-269)
//   D_k = f - lambda * sum_d (R_d(x) - R_d(x - e_d)), R_d(-1) = 0   (:220-236)
//   D_k = max(D_k, 0) if nonneg                     (:129)
//   P_{k+1} = proj(R + (1/(12 lambda)) (D_k(x) - D_k(x + e_d)))   (:237-256)
// with 1/(8 lambda) in 2D.  The output is D of the last iteration.
//
// HBM traffic per voxel-iteration (FP32): read f, P_k(3), P_{k-1}(3), write
// P_{k+1}(3) = 40 B; R is recomputed, never stored; P is triple-buffered.
// D is written only when a later residual check or the caller needs it.
//
// Fusion: a workgroup owns a 64 x TY column tile and marches through a chunk
// of z.  For plane z+1 it evaluates D on the tile plus its +x / +y halo into
// an LDS plane (double-buffered across z), then updates P on plane z from D
// at (x, x+e1, x+e2) and (z+1).  R3 of the plane below stays in a register.
#include "solvers.hpp"

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
__global__ __launch_bounds__(TX * TY) void fgp_iter_kernel(FgpArgs<T, S> a) {
    if (solve_stopped(a.ctrl)) return;
    const Geom& g = a.g;
    const TileCoord tc = tile_of_block(g, TY);
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

template <int DIM, int TY, typename T, typename S, typename I>
static const void* fgp_kernel_ptr(bool iso) {
    return iso ? (const void*)fgp_iter_kernel<DIM, TY, T, S, true, I> : (const void*)fgp_iter_kernel<DIM, TY, T, S, false, I>;
}

template <int DIM, int TY, typename T, typename S>
static void launch_fgp(const FgpArgs<T, S>& a, bool iso, bool small, hipStream_t s) {
    dim3 block(TX, TY), grid(a.g.nblocks);
    if (small) {
        if (iso) hipLaunchKernelGGL((fgp_iter_kernel<DIM, TY, T, S, true, uint32_t>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((fgp_iter_kernel<DIM, TY, T, S, false, uint32_t>), grid, block, 0, s, a);
    } else {
        if (iso) hipLaunchKernelGGL((fgp_iter_kernel<DIM, TY, T, S, true, uint64_t>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((fgp_iter_kernel<DIM, TY, T, S, false, uint64_t>), grid, block, 0, s, a);
    }
}

// Device-resident FGP solve.  d_out receives D of the final iteration (left
// untouched when iter == 0); with epsil != 0 its prior contents seed the first
// residual (FGP_TV_core.c:66,123).
template <typename T, typename S>
int fgp_solve(const T* d_f, T* d_out, T lambda, int iter, T epsil, int method, int nonneg, int nx, int ny, int nz,
              const Opts& opts, hipStream_t s, int* iters_done, float* re) {
    const int dim = nz <= 1 ? 2 : 3;
    if (dim == 2) nz = 1;
    *iters_done = 0;
    *re = 0.f;
    if (iter <= 0) return 0;
    const size_t n = (size_t)nx * ny * nz;
    const bool small = !opts.wide_index && fits_u32(n, sizeof(T) > sizeof(S) ? sizeof(T) : sizeof(S));
    const int ty = (opts.tile_y == 4 || dim == 2) ? 4 : 8;
    const bool iso = method == 0;
    Geom g;
    if (dim == 2)
        g = make_geom(nx, ny, nz, 4, 1);
    else if (ty == 8)
        g = plan_geom(small ? fgp_kernel_ptr<3, 8, T, S, uint32_t>(iso) : fgp_kernel_ptr<3, 8, T, S, uint64_t>(iso),
                      TX * 8, nx, ny, nz, 8, opts.zchunk);
    else
        g = plan_geom(small ? fgp_kernel_ptr<3, 4, T, S, uint32_t>(iso) : fgp_kernel_ptr<3, 4, T, S, uint64_t>(iso),
                      TX * 4, nx, ny, nz, 4, opts.zchunk);

    Arena ar;
    size_t oP[3][3], oD[2], oCtrl, oPart;
    for (int b = 0; b < 3; ++b)
        for (int c = 0; c < dim; ++c) oP[b][c] = ar.reserve(n * sizeof(S));
    for (int b = 0; b < 2; ++b) oD[b] = ar.reserve(n * sizeof(T));
    oCtrl = ar.reserve(sizeof(Ctrl));
    oPart = ar.reserve((size_t)g.nblocks * 2 * sizeof(double));
    if (ar.allocate() != 0) return -1;
    // P_0 = P_{-1} = 0 (buffers 0 and 2); buffer 1 receives P_1.
    for (int c = 0; c < dim; ++c) {
        launch_zero(ar.at(oP[0][c]), n * sizeof(S), s);
        launch_zero(ar.at(oP[2][c]), n * sizeof(S), s);
    }
    const bool use_tol = epsil != (T)0;
    T* Dbuf[2] = {(T*)ar.at(oD[0]), (T*)ar.at(oD[1])};
    if (use_tol) launch_copy(Dbuf[1], d_out, n * sizeof(T), s);   // D_{-1} = caller's Output
    Ctrl* ctrl = (Ctrl*)ar.at(oCtrl);
    double* partials = (double*)ar.at(oPart);
    const int ce = opts.check_every;

    // t_k sequence in the working precision (FGP_TV_core.c:138,159)
    T tk = 1, mult = 0;
    auto step = [&](int k, bool check) {
        FgpArgs<T, S> a;
        a.f = d_f;
        const int cur = k % 3, prv = (k + 2) % 3, nxt = (k + 1) % 3;
        for (int c = 0; c < 3; ++c) {
            const int cc = c < dim ? c : 0;
            a.P[c] = (const S*)ar.at(oP[cur][cc]);
            a.Pp[c] = (const S*)ar.at(oP[prv][cc]);
            a.Pn[c] = (S*)ar.at(oP[nxt][cc]);
        }
        a.Dout = Dbuf[k & 1];
        a.Dprev = Dbuf[(k + 1) & 1];
        a.lambda = lambda;
        a.step = (T)1 / ((dim == 3 ? (T)12 : (T)8) * lambda);
        a.mult = mult;
        const bool next_checks = use_tol && ((k + 1) % ce == 0);
        a.write_d = (k == iter - 1) || check || next_checks;
        a.check = check;
        a.nonneg = nonneg != 0;
        a.partials = partials;
        a.ctrl = ctrl;
        a.g = g;
        if (dim == 3) {
            if (ty == 8) launch_fgp<3, 8>(a, iso, small, s);
            else launch_fgp<3, 4>(a, iso, small, s);
        } else {
            launch_fgp<2, 4>(a, iso, small, s);
        }
        // advance t for the next iteration's R (the reference updates R at the
        // end of iteration k with t_k and t_{k+1})
        const T tkp1 = ((T)1 + dsqrt_host(((T)1 + (T)4 * tk * tk))) * (T)0.5;
        mult = (tk - (T)1) / tkp1;
        tk = tkp1;
    };
    int done = 0;
    if (drive_iterations(iter, (float)epsil, opts, ctrl, partials, g.nblocks, s, step, &done, re) != 0) return -1;
    *iters_done = done;
    // the last D written: iteration `done` on a stop, else iter - 1
    const int last = (done < iter) ? done : iter - 1;
    CILREG_CK(hipMemcpyAsync(d_out, Dbuf[last & 1], n * sizeof(T), hipMemcpyDeviceToDevice, s));
    CILREG_CK(hipStreamSynchronize(s));
    return 0;
}

template int fgp_solve<float, float>(const float*, float*, float, int, float, int, int, int, int, int, const Opts&,
                                     hipStream_t, int*, float*);
template int fgp_solve<float, bf16_t>(const float*, float*, float, int, float, int, int, int, int, int, const Opts&,
                                      hipStream_t, int*, float*);
template int fgp_solve<float, _Float16>(const float*, float*, float, int, float, int, int, int, int, int,
                                        const Opts&, hipStream_t, int*, float*);
template int fgp_solve<double, double>(const double*, double*, double, int, double, int, int, int, int, int,
                                       const Opts&, hipStream_t, int*, float*);

}  // namespace cilreg
