// rof.hip -- ROF-TV by explicit gradient descent on the smoothed TV
// functional, one fused kernel per iteration.
//
// Reference algorithm: src/Core/regularisers_CPU/ROF_TV_core.c:50-317.  Per
// iteration, with mirror boundaries (index -1 -> 1, index N -> N-2):
//   D_d = a+_d / sqrt(a+_d^2 + sum_{d' != d} minmod(a+_{d'}, a-_{d'})^2 + 1e-8)
//   u  += tau * (lambda * sum_d (D_d(x) - D_d(x - e_d)) - (u - f))
// where a+/a- are the one-sided differences of u.  D1 is the forward
// difference along y (the reference's "x"), D2 along x, D3 along z
// (:122-123).  The 3D D1 takes a double sqrt of the double sum (:136); every
// other flux rounds the double sum to float and calls sqrtf.
//
// HBM traffic per voxel-iteration: read u, f; write u = 12 B.
//
// Fusion: for plane z the fluxes D1 (tile + -y halo row) and D2 (tile + -x
// halo column) go through LDS; D3 of the own column stays in a register from
// the plane below (at z = 0 the mirror reads plane 1, computed up front).
#include "solvers.hpp"
#include "rof_kernels.hpp"

namespace cilreg {

template <int DIM, int TY, typename T>
static void launch_rof(const RofArgs<T>& a, bool small, hipStream_t s) {
    dim3 block(TX, TY), grid(a.g.nblocks);
    if (small) hipLaunchKernelGGL((rof_iter_kernel<DIM, TY, T, uint32_t>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((rof_iter_kernel<DIM, TY, T, uint64_t>), grid, block, 0, s, a);
}

template <int TY, typename T>
static void launch_rof2_2d(const Rof2Args<T>& a, bool small, hipStream_t s) {
    dim3 block(TX, TY), grid(a.g.nblocks);
    if (small) hipLaunchKernelGGL((rof2_2d_kernel<TY, T, uint32_t>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((rof2_2d_kernel<TY, T, uint64_t>), grid, block, 0, s, a);
}

// 2D images below this many pixels run two iterations per launch: there the
// one-iteration kernel is launch-bound (the pass evaluates about 7% more
// fluxes than two launches would, for one launch less).
constexpr size_t kRof2dPassMaxPixels = (size_t)1 << 20;

template <typename T, typename S>
int rof_solve(const T* d_f, T* d_out, T lambda, int iter, T tau, T epsil, int nx, int ny, int nz, const Opts& opts,
              hipStream_t s, int* iters_done, float* re) {
    if (nz < 1) nz = 1;
    const int dim = nz > 1 ? 3 : 2;
    const size_t n = (size_t)nx * ny * nz;
    *iters_done = 0;
    *re = 0.f;
    if (iter <= 0) {   // Output = Input (ROF_TV_core.c:65)
        CILREG_CK(hipMemcpyAsync(d_out, d_f, n * sizeof(T), hipMemcpyDeviceToDevice, s));
        CILREG_CK(hipStreamSynchronize(s));
        return 0;
    }
    const bool small = !opts.wide_index && fits_u32(n, sizeof(T));
    const int ty = (opts.tile_y == 4 || dim == 2) ? 4 : 8;
    // two iterations per launch for small 2D images (64 x 8 tiles, also for
    // the one-iteration launches of the same solve, so both share one
    // partials layout)
    const int width = (dim == 2 && iter > 1 && opts.depth != 1 &&
                       (opts.depth == 2 || n < kRof2dPassMaxPixels)) ? 2 : 1;
    Geom g;
    if (dim == 2 && width == 2)
        g = make_geom(nx, ny, nz, 8, 1);
    else if (dim == 2)
        g = make_geom(nx, ny, nz, 4, 1);
    else if (ty == 8)
        g = plan_geom(small ? (const void*)rof_iter_kernel<3, 8, T, uint32_t> : (const void*)rof_iter_kernel<3, 8, T, uint64_t>,
                      TX * 8, nx, ny, nz, 8, opts.zchunk);
    else
        g = plan_geom(small ? (const void*)rof_iter_kernel<3, 4, T, uint32_t> : (const void*)rof_iter_kernel<3, 4, T, uint64_t>,
                      TX * 4, nx, ny, nz, 4, opts.zchunk);
    Arena ar;
    size_t oU[2], oCtrl, oPart;
    oU[0] = ar.reserve(n * sizeof(T));
    oU[1] = ar.reserve(n * sizeof(T));
    oCtrl = ar.reserve(sizeof(Ctrl));
    oPart = ar.reserve((size_t)g.nblocks * 4 * sizeof(double));   // two iterations' sums
    if (ar.allocate() != 0) return -1;
    launch_copy(ar.at(oU[0]), d_f, n * sizeof(T), s);
    Ctrl* ctrl = (Ctrl*)ar.at(oCtrl);
    double* partials0 = (double*)ar.at(oPart);
    double* partials1 = partials0 + 2 * (size_t)g.nblocks;
    // launch q (one or two iterations) reads U[q mod 2] and writes the other
    int launches = 0;
    auto one = [&](int in, bool check, const Ctrl* c) {
        RofArgs<T> a;
        a.f = d_f;
        a.U = (const T*)ar.at(oU[in]);
        a.Uout = (T*)ar.at(oU[in ^ 1]);
        a.lambda = lambda;
        a.tau = tau;
        a.check = check;
        a.partials = partials0;
        a.ctrl = c;
        a.g = g;
        if (dim == 3) {
            if (ty == 8) launch_rof<3, 8>(a, small, s);
            else launch_rof<3, 4>(a, small, s);
        } else if (width == 2) {
            launch_rof<2, 8>(a, small, s);
        } else {
            launch_rof<2, 4>(a, small, s);
        }
    };
    auto step = [&](int, int nk, bool c0, bool c1) {
        const int in = launches & 1;
        ++launches;
        if (nk == 2) {
            Rof2Args<T> a;
            a.f = d_f;
            a.U = (const T*)ar.at(oU[in]);
            a.Uout = (T*)ar.at(oU[in ^ 1]);
            a.lambda = lambda;
            a.tau = tau;
            a.check0 = c0;
            a.check1 = c1;
            a.partials0 = partials0;
            a.partials1 = partials1;
            a.ctrl = ctrl;
            a.g = g;
            launch_rof2_2d<8>(a, small, s);
            return;
        }
        one(in, c0, ctrl);
    };
    int done = 0;
    if (drive_passes(iter, width, (float)epsil, opts, ctrl, partials0, partials1, g.nblocks, s, step, &done, re) != 0)
        return -1;
    *iters_done = done;
    const int last = (done < iter) ? done : iter - 1;
    int res = (last + 1) & 1;
    if (width == 2) {
        // iteration `last` ran in launch last / 2; a stop on the first
        // iteration of a two-iteration launch recomputes that iteration from
        // the launch's input, which the no-op launches after the stop left
        // intact (same kernel and expressions as every one-iteration launch)
        const int q = last / 2;
        res = (q + 1) & 1;
        if (done < iter && (done & 1) == 0 && done + 1 < iter) one(q & 1, false, nullptr);
    }
    CILREG_CK(hipMemcpyAsync(d_out, ar.at(oU[res]), n * sizeof(T), hipMemcpyDeviceToDevice, s));
    CILREG_CK(hipStreamSynchronize(s));
    return 0;
}

template int rof_solve<float, float>(const float*, float*, float, int, float, float, int, int, int, const Opts&,
                                     hipStream_t, int*, float*);

}  // namespace cilreg
