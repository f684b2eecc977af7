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
#include "fgp_kernels.hpp"

namespace cilreg {

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
