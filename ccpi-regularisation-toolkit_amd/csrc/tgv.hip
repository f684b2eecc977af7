// tgv.hip -- TGV^2-L2 denoising by primal-dual iterations, one fused kernel
// per iteration.
//
// Reference algorithm: src/Core/regularisers_CPU/TGV_core.c:43-568.  State:
// the extrapolated primal u_bar, the extrapolated auxiliary field v_bar
// (3 channels; 2 in 2D), the duals P (3; 2) and Q (6 symmetric-tensor
// channels Q11,Q22,Q33,Q12,Q13,Q23; 3 in 2D).  tau = sigma = L2^-1/2.
//   P = proj_alpha1(P + sigma (grad+ u_bar - v_bar))    (DualP/ProjP, :351-387)
//   Q = proj_alpha0(Q + sigma sym-grad+ v_bar)           (DualQ/ProjQ, :389-446)
//   U = (lambda (u_bar + tau div P) + tau f) / (lambda + tau)    (:448-472)
//   u_bar <- 2 U - u_bar                                 (:548-554)
//   V = v_bar + tau (P + div Q);  v_bar <- 2 V - v_bar   (:474-533, :558-568)
// Forward differences are zero at the last index; the divergences are their
// adjoints (the field at index 0, minus the predecessor at the last index).
// The residual compares the new u_bar with the old one.
//
// HBM traffic per voxel-iteration (FP32): read u_bar, f, v_bar(3), P(3),
// Q(6) = 14 fields, write 13 = 108 B.  With 16-bit storage of v_bar, P and Q:
// 60 B.
//
// Fusion: for plane z the updated P and Q are computed on the tile plus its
// -x / -y halo; the channels the backward divergences need from a neighbour
// go through LDS, those needed from plane z-1 stay in registers.
#include "solvers.hpp"
#include "tgv_kernels.hpp"

namespace cilreg {

template <int DIM, int TY, typename T, typename S>
static void launch_tgv(const TgvArgs<T, S>& a, bool small, hipStream_t s) {
    dim3 grid(a.g.nblocks), block(TX, TY);
    if (small) hipLaunchKernelGGL((tgv_iter_kernel<DIM, TY, T, S, uint32_t>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((tgv_iter_kernel<DIM, TY, T, S, uint64_t>), grid, block, 0, s, a);
}

template <int TY, typename T, typename S>
static const void* tgv2_kernel_ptr(bool small) {
    if constexpr (std::is_same<T, float>::value) {
        if (small) return (const void*)tgv2_iter_kernel<TY, T, S>;
        return (const void*)tgv2_iter_kernel_uncapped<TY, T, S, uint64_t>;
    } else {
        return small ? (const void*)tgv2_iter_kernel_uncapped<TY, T, S, uint32_t>
                     : (const void*)tgv2_iter_kernel_uncapped<TY, T, S, uint64_t>;
    }
}

// threads of the pass kernel tgv2_kernel_ptr returns
template <int TY, typename T>
static int tgv2_block_threads(bool small) {
    return (std::is_same<T, float>::value && small) ? tgv2_threads<TY>() : TX * TY;
}

template <int TY, typename T, typename S>
static void launch_tgv2(const Tgv2Args<T, S>& a, bool small, hipStream_t s) {
    dim3 grid(a.g.nblocks), block(TX, TY);
    if constexpr (std::is_same<T, float>::value) {
        if (small) hipLaunchKernelGGL((tgv2_iter_kernel<TY, T, S>), grid, dim3(TX, tgv2_threads<TY>() / TX), 0, s, a);
        else hipLaunchKernelGGL((tgv2_iter_kernel_uncapped<TY, T, S, uint64_t>), grid, block, 0, s, a);
    } else {
        if (small) hipLaunchKernelGGL((tgv2_iter_kernel_uncapped<TY, T, S, uint32_t>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((tgv2_iter_kernel_uncapped<TY, T, S, uint64_t>), grid, block, 0, s, a);
    }
}

// Device-resident TGV solve.  In 3D, iterations run two per pass
// (tgv2_iter_kernel) unless opts.depth == 1; an odd last iteration, and every
// 2D iteration, runs on the one-iteration kernel.  Each pass reads state set
// `cur` and writes set cur ^ 1.  The output of a stop on a pass's first
// iteration (u_bar after it) stays on chip; that case replays the iteration on
// the one-iteration kernel from the pass's input set afterwards.
template <typename T, typename S>
int tgv_solve(const T* d_f, T* d_out, T lambda, T alpha1, T alpha0, int iter, T L2, T epsil, int nx, int ny, int nz,
              const Opts& opts, hipStream_t s, int* iters_done, float* re, const WarmState* ws) {
    const int dim = nz == 1 ? 2 : 3;
    const size_t n = (size_t)nx * ny * nz;
    *iters_done = 0;
    *re = 0.f;
    if (iter <= 0) {   // U = U0 (TGV_core.c:53)
        CILREG_CK(hipMemcpyAsync(d_out, d_f, n * sizeof(T), hipMemcpyDeviceToDevice, s));
        CILREG_CK(hipStreamSynchronize(s));
        return 0;
    }
    const int nq = dim == 3 ? 6 : 3;
    // state in FP32 or 16-bit storage lives in pair fields (two channels side by side, tgv_pairs)
    const bool pairs = tgv_pairs<3, S>();
    const size_t pes = pairs ? 2 * sizeof(S) : sizeof(S);   // bytes per point of one field
    const bool small = !opts.wide_index && fits_u32(n, sizeof(T) > pes ? sizeof(T) : pes);
    const int ty = (opts.tile_y == 4 || dim == 2) ? 4 : 8;
    // returning the state after a stop needs the stop iteration's own pass
    const int width = (dim == 3 && opts.depth != 1 && iter > 1 && !(ws && epsil != (T)0)) ? 2 : 1;
    Geom g;
    if (dim == 2)
        g = make_geom(nx, ny, nz, 4, 1);
    else if (width == 2)
        g = ty == 8 ? plan_geom(tgv2_kernel_ptr<8, T, S>(small), tgv2_block_threads<8, T>(small), nx, ny, nz, 8, opts.zchunk, 2.5)
                    : plan_geom(tgv2_kernel_ptr<4, T, S>(small), tgv2_block_threads<4, T>(small), nx, ny, nz, 4, opts.zchunk, 2.5);
    else if (ty == 8)
        g = plan_geom(small ? (const void*)tgv_iter_kernel<3, 8, T, S, uint32_t> : (const void*)tgv_iter_kernel<3, 8, T, S, uint64_t>,
                      TX * 8, nx, ny, nz, 8, opts.zchunk);
    else
        g = plan_geom(small ? (const void*)tgv_iter_kernel<3, 4, T, S, uint32_t> : (const void*)tgv_iter_kernel<3, 4, T, S, uint64_t>,
                      TX * 4, nx, ny, nz, 4, opts.zchunk);

    Arena ar;
    size_t oU[2], oV[2][3], oP[2][3], oQ[2][6], oCtrl, oPart;
    for (int b = 0; b < 2; ++b) {
        oU[b] = ar.reserve(n * sizeof(T));
        if (pairs && dim == 3) {   // six pair fields; the second channel of a pair names the same buffer
            const size_t pb = 2 * n * sizeof(S);
            oV[b][0] = oV[b][1] = ar.reserve(pb);    // V1, V2
            oV[b][2] = oP[b][0] = ar.reserve(pb);    // V3, P1
            oP[b][1] = oP[b][2] = ar.reserve(pb);    // P2, P3
            for (int c = 0; c < 6; c += 2) oQ[b][c] = oQ[b][c + 1] = ar.reserve(pb);
        } else if (pairs) {   // 2D: three pair fields and Q12 alone
            const size_t pb = 2 * n * sizeof(S);
            oV[b][0] = oV[b][1] = oV[b][2] = ar.reserve(pb);   // V1, V2
            oP[b][0] = oP[b][1] = oP[b][2] = ar.reserve(pb);   // P1, P2
            oQ[b][0] = oQ[b][1] = ar.reserve(pb);              // Q11, Q22
            oQ[b][2] = ar.reserve(n * sizeof(S));              // Q12
            for (int c = 3; c < 6; ++c) oQ[b][c] = oQ[b][2];
        } else {
            for (int c = 0; c < dim; ++c) oV[b][c] = ar.reserve(n * sizeof(S));
            for (int c = 0; c < dim; ++c) oP[b][c] = ar.reserve(n * sizeof(S));
            for (int c = 0; c < nq; ++c) oQ[b][c] = ar.reserve(n * sizeof(S));
        }
    }
    oCtrl = ar.reserve(sizeof(Ctrl));
    oPart = ar.reserve((size_t)g.nblocks * 4 * sizeof(double));
    if (ar.allocate() != 0) return -1;
    // State channel ch (V1..3, P1..3, then Q) of set b: its first element and
    // element stride.
    auto chan = [&](int b, int ch, int& stride) -> S* {
        if (pairs && dim == 3) {
            const size_t pf[6] = {oV[b][0], oV[b][2], oP[b][1], oQ[b][0], oQ[b][2], oQ[b][4]};
            stride = 2;
            return (S*)ar.at(pf[ch / 2]) + (ch & 1);
        }
        if (pairs) {   // 2D channels: V 0-1, P 3-4, Q 6-8
            stride = ch == 8 ? 1 : 2;
            if (ch < 3) return (S*)ar.at(oV[b][0]) + ch;
            if (ch < 6) return (S*)ar.at(oP[b][0]) + (ch - 3);
            return ch == 8 ? (S*)ar.at(oQ[b][2]) : (S*)ar.at(oQ[b][0]) + (ch - 6);
        }
        stride = 1;
        return (S*)ar.at(ch < 3 ? oV[b][ch] : ch < 6 ? oP[b][ch - 3] : oQ[b][ch - 6]);
    };
    const int nch = 2 * dim + nq;   // channels in state order (3D: 12, 2D: 7)
    auto chan_of = [&](int k) { return k < dim ? k : k < 2 * dim ? 3 + (k - dim) : 6 + (k - 2 * dim); };
    launch_copy(ar.at(oU[0]), d_f, n * sizeof(T), s);
    if (pairs && dim == 3) {
        for (int k = 0; k < 6; ++k) {
            int st;
            launch_zero(chan(0, 2 * k, st), 2 * n * sizeof(S), s);
        }
    } else if (pairs) {
        launch_zero(ar.at(oV[0][0]), 2 * n * sizeof(S), s);
        launch_zero(ar.at(oP[0][0]), 2 * n * sizeof(S), s);
        launch_zero(ar.at(oQ[0][0]), 2 * n * sizeof(S), s);
        launch_zero(ar.at(oQ[0][2]), n * sizeof(S), s);
    } else {
        for (int c = 0; c < dim; ++c) {
            launch_zero(ar.at(oV[0][c]), n * sizeof(S), s);
            launch_zero(ar.at(oP[0][c]), n * sizeof(S), s);
        }
        for (int c = 0; c < nq; ++c) launch_zero(ar.at(oQ[0][c]), n * sizeof(S), s);
    }
    // caller state channels: u_bar, V1..3, P1..3, Q (Q11, Q22, Q33, Q12, Q13, Q23 in 3D; Q11, Q22, Q12 in 2D)
    if constexpr (std::is_same<T, float>::value) {
        if (ws && ws->warm) {
            state_in<T>((T*)ar.at(oU[0]), ws->d_state, n, s);
            for (int k = 0; k < nch; ++k) {
                const int ch = chan_of(k);
                const size_t at = ch < 3 ? 1 + ch : ch < 6 ? 4 + (ch - 3) : 7 + (ch - 6);
                int st;
                S* p = chan(0, ch, st);
                state_in<S>(p, ws->d_state + at * n, n, s, st);
            }
        }
    }
    Ctrl* ctrl = (Ctrl*)ar.at(oCtrl);
    double* partials0 = (double*)ar.at(oPart);
    double* partials1 = partials0 + 2 * (size_t)g.nblocks;
    const T tau = (T)pow((double)L2, -0.5);   // TGV_core.c:54-55

    int cur = 0;
    std::vector<int> pass_k, pass_n, pass_out;   // where each pass left its state
    // one iteration from state set `in` to set in ^ 1
    auto one = [&](int in, bool check, const Ctrl* c_ptr) {
        const int out = in ^ 1;
        TgvArgs<T, S> a;
        a.f = d_f;
        a.U = (const T*)ar.at(oU[in]);
        a.Uout = (T*)ar.at(oU[out]);
        for (int c = 0; c < 3; ++c) {
            const int cc = c < dim ? c : 0;
            a.V[c] = (const S*)ar.at(oV[in][cc]);
            a.Vout[c] = (S*)ar.at(oV[out][cc]);
            a.P[c] = (const S*)ar.at(oP[in][cc]);
            a.Pout[c] = (S*)ar.at(oP[out][cc]);
        }
        for (int c = 0; c < 6; ++c) {
            const int cc = c < nq ? c : 0;
            a.Q[c] = (const S*)ar.at(oQ[in][cc]);
            a.Qout[c] = (S*)ar.at(oQ[out][cc]);
        }
        a.lambda = lambda;
        a.tau = tau;
        a.sigma = tau;
        a.alpha1 = alpha1;
        a.alpha0 = alpha0;
        a.check = check;
        a.partials = partials0;
        a.ctrl = c_ptr;
        a.g = g;
        if (dim == 3) {
            if (ty == 8) launch_tgv<3, 8>(a, small, s);
            else launch_tgv<3, 4>(a, small, s);
        } else {
            launch_tgv<2, 4>(a, small, s);
        }
    };
    auto step = [&](int k, int nk, bool c0, bool c1) {
        const int nxt = cur ^ 1;
        pass_k.push_back(k);
        pass_n.push_back(nk);
        pass_out.push_back(nxt);
        if (nk == 2) {
            Tgv2Args<T, S> a;
            a.f = d_f;
            a.U = (const T*)ar.at(oU[cur]);
            a.Uout = (T*)ar.at(oU[nxt]);
            for (int c = 0; c < 3; ++c) {
                a.V[c] = (const S*)ar.at(oV[cur][c]);
                a.Vout[c] = (S*)ar.at(oV[nxt][c]);
                a.P[c] = (const S*)ar.at(oP[cur][c]);
                a.Pout[c] = (S*)ar.at(oP[nxt][c]);
            }
            for (int c = 0; c < 6; ++c) {
                a.Q[c] = (const S*)ar.at(oQ[cur][c]);
                a.Qout[c] = (S*)ar.at(oQ[nxt][c]);
            }
            a.Ualt = nullptr;   // a stop on the first iteration is replayed afterwards
            a.lambda = lambda;
            a.tau = tau;
            a.sigma = tau;
            a.alpha1 = alpha1;
            a.alpha0 = alpha0;
            a.check0 = c0;
            a.check1 = c1;
            a.single = 0;
            a.partials0 = partials0;
            a.partials1 = partials1;
            a.ctrl = ctrl;
            a.g = g;
            if (ty == 8) launch_tgv2<8>(a, small, s);
            else launch_tgv2<4>(a, small, s);
        } else {
            one(cur, c0, ctrl);
        }
        cur = nxt;
    };
    int done = 0;
    if (drive_passes(iter, width, (float)epsil, opts, ctrl, partials0, partials1, g.nblocks, s, step, &done, re) != 0)
        return -1;
    *iters_done = done;
    // the output is u_bar after iteration `last`: the stop iteration, or iter - 1
    const int last = (done < iter) ? done : iter - 1;
    const void* src = nullptr;
    for (size_t q = 0; q < pass_k.size(); ++q) {
        if (last >= pass_k[q] && last < pass_k[q] + pass_n[q]) {
            if (pass_n[q] == 2 && last == pass_k[q]) {
                // a stop on the pass's first iteration: replay it from the
                // pass's input set (intact: the passes after the stop were
                // no-ops) into its output set
                one(pass_out[q] ^ 1, false, nullptr);
            }
            src = ar.at(oU[pass_out[q]]);
            break;
        }
    }
    if (!src) { set_error("TGV: no pass covers iteration %d", last); return -1; }
    CILREG_CK(hipMemcpyAsync(d_out, src, n * sizeof(T), hipMemcpyDeviceToDevice, s));
    if constexpr (std::is_same<T, float>::value) {
        if (ws) {   // the state set left by the pass holding the last executed iteration
            int set = -1;
            for (size_t q = 0; q < pass_k.size(); ++q)
                if (last >= pass_k[q] && last < pass_k[q] + pass_n[q]) set = pass_out[q];
            if (set < 0) { set_error("TGV: no pass covers iteration %d", last); return -1; }
            state_out<T>(ws->d_state, (const T*)ar.at(oU[set]), n, s);
            for (int k = 0; k < nch; ++k) {
                const int ch = chan_of(k);
                const size_t at = ch < 3 ? 1 + ch : ch < 6 ? 4 + (ch - 3) : 7 + (ch - 6);
                int st;
                const S* p = chan(set, ch, st);
                state_out<S>(ws->d_state + at * n, p, n, s, st);
            }
        }
    }
    CILREG_CK(hipStreamSynchronize(s));
    return 0;
}

template int tgv_solve<float, float>(const float*, float*, float, float, float, int, float, float, int, int, int,
                                     const Opts&, hipStream_t, int*, float*, const WarmState*);
template int tgv_solve<float, bf16_t>(const float*, float*, float, float, float, int, float, float, int, int, int,
                                      const Opts&, hipStream_t, int*, float*, const WarmState*);
template int tgv_solve<float, _Float16>(const float*, float*, float, float, float, int, float, float, int, int, int,
                                        const Opts&, hipStream_t, int*, float*, const WarmState*);
template int tgv_solve<double, double>(const double*, double*, double, double, double, int, double, double, int, int,
                                       int, const Opts&, hipStream_t, int*, float*, const WarmState*);

}  // namespace cilreg
