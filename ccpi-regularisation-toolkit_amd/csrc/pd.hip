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
#include "pd_kernels.hpp"
#include <vector>

namespace cilreg {

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

template <int TY, typename T, typename S>
static const void* pd2_kernel_ptr(bool iso, bool small) {
    if (small)
        return iso ? (const void*)pd2_iter_kernel<TY, T, S, true, uint32_t>
                   : (const void*)pd2_iter_kernel<TY, T, S, false, uint32_t>;
    return iso ? (const void*)pd2_iter_kernel<TY, T, S, true, uint64_t>
               : (const void*)pd2_iter_kernel<TY, T, S, false, uint64_t>;
}

template <int TY, typename T, typename S>
static void launch_pd2(const Pd2Args<T, S>& a, bool iso, bool small, hipStream_t s) {
    dim3 block(TX, TY), grid(a.g.nblocks);
    if (small) {
        if (iso) hipLaunchKernelGGL((pd2_iter_kernel<TY, T, S, true, uint32_t>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((pd2_iter_kernel<TY, T, S, false, uint32_t>), grid, block, 0, s, a);
    } else {
        if (iso) hipLaunchKernelGGL((pd2_iter_kernel<TY, T, S, true, uint64_t>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((pd2_iter_kernel<TY, T, S, false, uint64_t>), grid, block, 0, s, a);
    }
}

// Device-resident PD solve.  In 3D, iterations run two per pass
// (pd2_iter_kernel) unless opts.depth == 1; an odd last iteration, and every
// 2D iteration, runs on the one-iteration kernel.  Each pass reads state set
// `cur` (u_bar, P) and writes set cur ^ 1.  A stop returns the un-extrapolated
// U of the stop iteration: one-iteration passes store it on checks (Ubrk); a
// stop inside a two-iteration pass replays that pass's iterations on the
// one-iteration kernel afterwards.
template <typename T, typename S>
int pd_solve(const T* d_f, T* d_out, T lambda, int iter, T epsil, T lipschitz, int method, int nonneg, int nx, int ny,
             int nz, const Opts& opts, hipStream_t s, int* iters_done, float* re, const WarmState* ws) {
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
    // returning the state after a stop needs the stop iteration's own pass
    const int width = (dim == 3 && opts.depth != 1 && iter > 1 && !(ws && epsil != (T)0)) ? 2 : 1;
    Geom g;
    if (dim == 2)
        g = make_geom(nx, ny, nz, 4, 1);
    else if (width == 2)
        g = ty == 8 ? plan_geom(pd2_kernel_ptr<8, T, S>(iso, small), TX * 8, nx, ny, nz, 8, opts.zchunk, 2.5)
                    : plan_geom(pd2_kernel_ptr<4, T, S>(iso, small), TX * 4, nx, ny, nz, 4, opts.zchunk, 2.5);
    else if (ty == 8)
        g = plan_geom(small ? pd_kernel_ptr<3, 8, T, S, uint32_t>(iso) : pd_kernel_ptr<3, 8, T, S, uint64_t>(iso),
                      TX * 8, nx, ny, nz, 8, opts.zchunk);
    else
        g = plan_geom(small ? pd_kernel_ptr<3, 4, T, S, uint32_t>(iso) : pd_kernel_ptr<3, 4, T, S, uint64_t>(iso),
                      TX * 4, nx, ny, nz, 4, opts.zchunk);

    Arena ar;
    size_t oU[2], oP[2][3], oB[1] = {0}, oCtrl, oPart;
    for (int b = 0; b < 2; ++b) oU[b] = ar.reserve(n * sizeof(T));
    for (int b = 0; b < 2; ++b)
        for (int c = 0; c < dim; ++c) oP[b][c] = ar.reserve(n * sizeof(S));
    oB[0] = ar.reserve(n * sizeof(T));
    oCtrl = ar.reserve(sizeof(Ctrl));
    oPart = ar.reserve((size_t)g.nblocks * 4 * sizeof(double));
    if (ar.allocate() != 0) return -1;
    launch_copy(ar.at(oU[0]), d_f, n * sizeof(T), s);
    for (int c = 0; c < dim; ++c) launch_zero(ar.at(oP[0][c]), n * sizeof(S), s);
    if constexpr (std::is_same<T, float>::value) {
        if (ws && ws->warm) {   // u_bar and P from the caller's state [u_bar, P1, P2, P3]
            state_in<T>((T*)ar.at(oU[0]), ws->d_state, n, s);
            for (int c = 0; c < dim; ++c) state_in<S>((S*)ar.at(oP[0][c]), ws->d_state + (size_t)(1 + c) * n, n, s);
        }
    }
    Ctrl* ctrl = (Ctrl*)ar.at(oCtrl);
    double* partials0 = (double*)ar.at(oPart);
    double* partials1 = partials0 + 2 * (size_t)g.nblocks;

    // step sizes in the working precision (PD_TV_core.c:50-53)
    const T tau = lambda * (T)0.1;
    const T sigma = (T)(1.0 / (double)(lipschitz * tau));
    const T lt = tau / lambda;
    int cur = 0;
    std::vector<int> pass_k, pass_n, pass_out;
    // one iteration from state set `in` to set in ^ 1; with `check` the
    // un-extrapolated U goes to Ubrk
    auto one = [&](int in, bool check, const Ctrl* c_ptr) {
        const int out = in ^ 1;
        PdArgs<T, S> a;
        a.f = d_f;
        a.U = (const T*)ar.at(oU[in]);
        a.Uout = (T*)ar.at(oU[out]);
        for (int c = 0; c < 3; ++c) {
            const int cc = c < dim ? c : 0;
            a.P[c] = (const S*)ar.at(oP[in][cc]);
            a.Pout[c] = (S*)ar.at(oP[out][cc]);
        }
        a.Ubrk = (T*)ar.at(oB[0]);
        a.sigma = sigma;
        a.tau = tau;
        a.lt = lt;
        a.theta = 1;
        a.nonneg = nonneg != 0;
        a.check = check;
        a.partials = partials0;
        a.ctrl = c_ptr;
        a.g = g;
        if (dim == 3) {
            if (ty == 8) launch_pd<3, 8>(a, iso, small, s);
            else launch_pd<3, 4>(a, iso, small, s);
        } else {
            launch_pd<2, 4>(a, iso, small, s);
        }
    };
    auto step = [&](int k, int nk, bool c0, bool c1) {
        const int nxt = cur ^ 1;
        pass_k.push_back(k);
        pass_n.push_back(nk);
        pass_out.push_back(nxt);
        if (nk == 2) {
            Pd2Args<T, S> a;
            a.f = d_f;
            a.U = (const T*)ar.at(oU[cur]);
            a.Uout = (T*)ar.at(oU[nxt]);
            for (int c = 0; c < 3; ++c) {
                a.P[c] = (const S*)ar.at(oP[cur][c]);
                a.Pout[c] = (S*)ar.at(oP[nxt][c]);
            }
            // the un-extrapolated U of a stop inside the pass is recomputed
            // afterwards, so the pass stores none
            a.Ubrk0 = nullptr;
            a.Ubrk1 = nullptr;
            a.sigma = sigma;
            a.tau = tau;
            a.lt = lt;
            a.theta = 1;
            a.nonneg = nonneg != 0;
            a.check0 = c0;
            a.check1 = c1;
            a.partials0 = partials0;
            a.partials1 = partials1;
            a.ctrl = ctrl;
            a.g = g;
            if (ty == 8) launch_pd2<8>(a, iso, small, s);
            else launch_pd2<4>(a, iso, small, s);
        } else {
            one(cur, c0, ctrl);
        }
        cur = nxt;
    };
    int done = 0;
    if (drive_passes(iter, width, (float)epsil, opts, ctrl, partials0, partials1, g.nblocks, s, step, &done, re) != 0)
        return -1;
    *iters_done = done;
    // a stop returns U of the stop iteration (before extrapolation); else u_bar after the last pass
    const void* src = nullptr;
    if (done < iter) {
        for (size_t q = 0; q < pass_k.size(); ++q)
            if (done >= pass_k[q] && done < pass_k[q] + pass_n[q]) {
                if (pass_n[q] == 2) {
                    // replay the pass's iterations up to the stop one from its
                    // input set (intact: the passes after the stop were no-ops)
                    // with the one-iteration kernel, whose results are the same
                    const int in = pass_out[q] ^ 1;
                    if (done == pass_k[q]) {
                        one(in, true, nullptr);
                    } else {
                        one(in, false, nullptr);
                        one(in ^ 1, true, nullptr);
                    }
                }
                src = ar.at(oB[0]);
                break;
            }
    } else {
        src = ar.at(oU[pass_out.back()]);
    }
    if (!src) { set_error("PD: no pass covers iteration %d", done); return -1; }
    CILREG_CK(hipMemcpyAsync(d_out, src, n * sizeof(T), hipMemcpyDeviceToDevice, s));
    if constexpr (std::is_same<T, float>::value) {
        if (ws) {   // the state set left by the pass holding the last executed iteration
            const int last = done < iter ? done : iter - 1;
            int set = -1;
            for (size_t q = 0; q < pass_k.size(); ++q)
                if (last >= pass_k[q] && last < pass_k[q] + pass_n[q]) set = pass_out[q];
            if (set < 0) { set_error("PD: no pass covers iteration %d", last); return -1; }
            state_out<T>(ws->d_state, (const T*)ar.at(oU[set]), n, s);
            for (int c = 0; c < dim; ++c)
                state_out<S>(ws->d_state + (size_t)(1 + c) * n, (const S*)ar.at(oP[set][c]), n, s);
        }
    }
    CILREG_CK(hipStreamSynchronize(s));
    return 0;
}
template int pd_solve<float, float>(const float*, float*, float, int, float, float, int, int, int, int, int,
                                    const Opts&, hipStream_t, int*, float*, const WarmState*);
template int pd_solve<float, bf16_t>(const float*, float*, float, int, float, float, int, int, int, int, int,
                                     const Opts&, hipStream_t, int*, float*, const WarmState*);
template int pd_solve<float, _Float16>(const float*, float*, float, int, float, float, int, int, int, int, int,
                                       const Opts&, hipStream_t, int*, float*, const WarmState*);
template int pd_solve<double, double>(const double*, double*, double, int, double, double, int, int, int, int, int,
                                      const Opts&, hipStream_t, int*, float*, const WarmState*);

}  // namespace cilreg
