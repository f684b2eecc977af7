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
template <int TY, typename T, typename S, int DIM = 3>
static const void* fgp2_kernel_ptr(bool iso, bool small) {
    if (small)
        return iso ? (const void*)fgp2_iter_kernel<TY, T, S, true, uint32_t, DIM>
                   : (const void*)fgp2_iter_kernel<TY, T, S, false, uint32_t, DIM>;
    return iso ? (const void*)fgp2_iter_kernel<TY, T, S, true, uint64_t, DIM>
               : (const void*)fgp2_iter_kernel<TY, T, S, false, uint64_t, DIM>;
}

template <int TY, typename T, typename S, int DIM = 3>
static void launch_fgp2(const Fgp2Args<T, S>& a, bool iso, bool small, hipStream_t s) {
    dim3 block(TX, TY), grid(a.g.nblocks);
    if (small) {
        if (iso) hipLaunchKernelGGL((fgp2_iter_kernel<TY, T, S, true, uint32_t, DIM>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((fgp2_iter_kernel<TY, T, S, false, uint32_t, DIM>), grid, block, 0, s, a);
    } else {
        if (iso) hipLaunchKernelGGL((fgp2_iter_kernel<TY, T, S, true, uint64_t, DIM>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((fgp2_iter_kernel<TY, T, S, false, uint64_t, DIM>), grid, block, 0, s, a);
    }
}

// Device-resident FGP solve.  d_out receives D of the final iteration (left
// untouched when iter == 0); with epsil != 0 its prior contents seed the first
// residual (FGP_TV_core.c:66,123).  Iterations run two per pass
// (fgp2_iter_kernel) unless opts.depth == 1 or, by default, a 2D image has
// kFgp2dPassMaxPixels or more; an odd last iteration runs on the
// one-iteration kernel.  P_j lives in buffer
// j mod nbuf (nbuf = 4 with two-iteration passes, which write P_{k+1} and
// P_{k+2} while reading P_k and P_{k-1}; 3 otherwise).
template <typename T, typename S>
int fgp_solve(const T* d_f, T* d_out, T lambda, int iter, T epsil, int method, int nonneg, int nx, int ny, int nz,
              const Opts& opts, hipStream_t s, int* iters_done, float* re, const WarmState* ws) {
    const int dim = nz <= 1 ? 2 : 3;
    if (dim == 2) nz = 1;
    *iters_done = 0;
    *re = 0.f;
    if (iter <= 0) return 0;
    const size_t n = (size_t)nx * ny * nz;
    const bool small = !opts.wide_index && fits_u32(n, sizeof(T) > sizeof(S) ? sizeof(T) : sizeof(S));
    const int ty = (opts.tile_y == 4 || dim == 2) ? 4 : 8;
    const bool iso = method == 0;
    // returning the state after a stop needs the stop iteration's own pass
    const bool pass2 = opts.depth == 2 || (opts.depth == 0 && (dim == 3 || n < kFgp2dPassMaxPixels));
    const int width = (pass2 && iter > 1 && !(ws && epsil != (T)0)) ? 2 : 1;
    // 2D two-iteration passes: 64 x 4 tiles unless 8 rows are asked for
    const int ty2d = opts.tile_y == 8 ? 8 : 4;
    Geom g;
    if (dim == 2 && width == 2)
        g = make_geom(nx, ny, nz, ty2d, 1);
    else if (dim == 2)
        g = make_geom(nx, ny, nz, 4, 1);
    else if (width == 2)
        // the pass evaluates about three planes below and above its chunk
        g = ty == 8 ? plan_geom(fgp2_kernel_ptr<8, T, S>(iso, small), TX * 8, nx, ny, nz, 8, opts.zchunk, 3.0)
                    : plan_geom(fgp2_kernel_ptr<4, T, S>(iso, small), TX * 4, nx, ny, nz, 4, opts.zchunk, 3.0);
    else if (ty == 8)
        g = plan_geom(small ? fgp_kernel_ptr<3, 8, T, S, uint32_t>(iso) : fgp_kernel_ptr<3, 8, T, S, uint64_t>(iso),
                      TX * 8, nx, ny, nz, 8, opts.zchunk);
    else
        g = plan_geom(small ? fgp_kernel_ptr<3, 4, T, S, uint32_t>(iso) : fgp_kernel_ptr<3, 4, T, S, uint64_t>(iso),
                      TX * 4, nx, ny, nz, 4, opts.zchunk);

    const int nbuf = width == 2 ? 4 : 3;
    Arena ar;
    if (opts.scratch) ar.use_external(opts.scratch, opts.scratch_bytes);   // the allocation-free entry point
    size_t oP[4][3], oD[2], oCtrl, oPart;
    for (int b = 0; b < nbuf; ++b)
        for (int c = 0; c < dim; ++c) oP[b][c] = ar.reserve(n * sizeof(S));
    for (int b = 0; b < 2; ++b) oD[b] = ar.reserve(n * sizeof(T));
    oCtrl = ar.reserve(sizeof(Ctrl));
    oPart = ar.reserve((size_t)g.nblocks * 4 * sizeof(double));
    if (ar.allocate() != 0) return -1;
    // P_0 = P_{-1} = 0
    for (int c = 0; c < dim; ++c) {
        launch_zero(ar.at(oP[0][c]), n * sizeof(S), s);
        launch_zero(ar.at(oP[nbuf - 1][c]), n * sizeof(S), s);
    }
    auto Pbuf = [&](int j, int c) -> S* { return (S*)ar.at(oP[((j % nbuf) + nbuf) % nbuf][c < dim ? c : 0]); };
    if constexpr (std::is_same<T, float>::value) {
        if (ws && ws->warm)   // P_0 and P_{-1} from the caller's state
            for (int c = 0; c < dim; ++c) {
                state_in<S>(Pbuf(0, c), ws->d_state + (size_t)c * n, n, s);
                state_in<S>(Pbuf(-1, c), ws->d_state + (size_t)(3 + c) * n, n, s);
            }
    }
    const bool use_tol = epsil != (T)0;
    T* Dbuf[2] = {(T*)ar.at(oD[0]), (T*)ar.at(oD[1])};
    if (use_tol) launch_copy(Dbuf[1], d_out, n * sizeof(T), s);   // D_{-1} = caller's Output
    Ctrl* ctrl = (Ctrl*)ar.at(oCtrl);
    double* partials0 = (double*)ar.at(oPart);
    double* partials1 = partials0 + 2 * (size_t)g.nblocks;
    const int ce = opts.check_every;
    // D_k is needed when iteration k is the last, is checked, or precedes a check
    auto needs_d = [&](int k) {
        return k == iter - 1 || (use_tol && (k % ce == 0 || (k + 1) % ce == 0));
    };
    const T stepc = (T)1 / ((dim == 3 ? (T)12 : (T)8) * lambda);

    // t_k sequence in the working precision (FGP_TV_core.c:138,159): mult[k]
    // multiplies (P_k - P_{k-1}) in iteration k
    std::vector<T> mult(iter + 1), tks(iter + 1);
    {
        T tk = 1, m = 0;
        if (ws && ws->warm) {   // resume the sequence where the caller's state left it
            tk = (T)ws->tstate[0];
            m = (T)ws->tstate[1];
        }
        for (int k = 0; k <= iter; ++k) {
            mult[k] = m;
            tks[k] = tk;
            const T tkp1 = ((T)1 + dsqrt_host(((T)1 + (T)4 * tk * tk))) * (T)0.5;
            m = (tk - (T)1) / tkp1;
            tk = tkp1;
        }
    }
    // one iteration k on the one-iteration kernel
    auto one = [&](int k, bool write_d, bool check, const Ctrl* c_ptr) {
        FgpArgs<T, S> a;
        a.f = d_f;
        for (int c = 0; c < 3; ++c) {
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
        a.partials = partials0;
        a.ctrl = c_ptr;
        a.g = g;
        if (dim == 3) {
            if (ty == 8) launch_fgp<3, 8>(a, iso, small, s);
            else launch_fgp<3, 4>(a, iso, small, s);
        } else if (width == 2 && ty2d == 8) {
            launch_fgp<2, 8>(a, iso, small, s);   // odd last iteration on the pass geometry
        } else {
            launch_fgp<2, 4>(a, iso, small, s);
        }
    };
    auto step = [&](int k, int nk, bool c0, bool c1) {
        if (nk == 2) {
            Fgp2Args<T, S> a;
            a.f = d_f;
            for (int c = 0; c < 3; ++c) {
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
            // D_k of a pass's first iteration is needed in memory only as the
            // output of a stop there (its check and iteration k+1 read it from
            // LDS); that case recomputes it after the passes
            a.write_d0 = 0;
            a.write_d1 = needs_d(k + 1) || c1;
            a.check0 = c0;
            a.check1 = c1;
            a.nonneg = nonneg != 0;
            a.partials0 = partials0;
            a.partials1 = partials1;
            a.ctrl = ctrl;
            a.g = g;
            if (dim == 2) {
                if (ty2d == 8) launch_fgp2<8, T, S, 2>(a, iso, small, s);
                else launch_fgp2<4, T, S, 2>(a, iso, small, s);
            } else if (ty == 8) {
                launch_fgp2<8>(a, iso, small, s);
            } else {
                launch_fgp2<4>(a, iso, small, s);
            }
            return;
        }
        one(k, needs_d(k) || c0, c0, ctrl);
    };
    int done = 0;
    if (drive_passes(iter, width, (float)epsil, opts, ctrl, partials0, partials1, g.nblocks, s, step, &done, re) != 0)
        return -1;
    *iters_done = done;
    // a stop on the first iteration of a two-iteration pass: D_done from
    // P_done and P_{done-1}, which the no-op passes after the stop left intact
    // (same kernel and expressions as every one-iteration pass)
    if (done < iter && width == 2 && done % 2 == 0 && done + 1 < iter) one(done, true, false, nullptr);
    // the last D written: iteration `done` on a stop, else iter - 1
    const int last = (done < iter) ? done : iter - 1;
    CILREG_CK(hipMemcpyAsync(d_out, Dbuf[last & 1], n * sizeof(T), hipMemcpyDeviceToDevice, s));
    if constexpr (std::is_same<T, float>::value) {
        if (ws) {   // P and P_prev after the executed iterations, t and the next multiplier
            const int executed = last + 1;
            for (int c = 0; c < dim; ++c) {
                state_out<S>(ws->d_state + (size_t)c * n, Pbuf(executed, c), n, s);
                state_out<S>(ws->d_state + (size_t)(3 + c) * n, Pbuf(executed - 1, c), n, s);
            }
            ws->tstate[0] = (float)tks[executed];
            ws->tstate[1] = (float)mult[executed];
        }
    }
    CILREG_CK(hipStreamSynchronize(s));
    return 0;
}
template int fgp_solve<float, float>(const float*, float*, float, int, float, int, int, int, int, int, const Opts&,
                                     hipStream_t, int*, float*, const WarmState*);
template int fgp_solve<float, bf16_t>(const float*, float*, float, int, float, int, int, int, int, int, const Opts&,
                                      hipStream_t, int*, float*, const WarmState*);
template int fgp_solve<float, _Float16>(const float*, float*, float, int, float, int, int, int, int, int,
                                        const Opts&, hipStream_t, int*, float*, const WarmState*);
template int fgp_solve<double, double>(const double*, double*, double, int, double, int, int, int, int, int,
                                       const Opts&, hipStream_t, int*, float*, const WarmState*);

}  // namespace cilreg
