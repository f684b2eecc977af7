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
    Geom g;
    if (dim == 2)
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
    oPart = ar.reserve((size_t)g.nblocks * 2 * sizeof(double));
    if (ar.allocate() != 0) return -1;
    launch_copy(ar.at(oU[0]), d_f, n * sizeof(T), s);
    Ctrl* ctrl = (Ctrl*)ar.at(oCtrl);
    double* partials = (double*)ar.at(oPart);
    auto step = [&](int k, bool check) {
        RofArgs<T> a;
        a.f = d_f;
        a.U = (const T*)ar.at(oU[k & 1]);
        a.Uout = (T*)ar.at(oU[(k + 1) & 1]);
        a.lambda = lambda;
        a.tau = tau;
        a.check = check;
        a.partials = partials;
        a.ctrl = ctrl;
        a.g = g;
        if (dim == 3) {
            if (ty == 8) launch_rof<3, 8>(a, small, s);
            else launch_rof<3, 4>(a, small, s);
        } else {
            launch_rof<2, 4>(a, small, s);
        }
    };
    int done = 0;
    if (drive_iterations(iter, (float)epsil, opts, ctrl, partials, g.nblocks, s, step, &done, re) != 0) return -1;
    *iters_done = done;
    const int last = (done < iter) ? done : iter - 1;
    CILREG_CK(hipMemcpyAsync(d_out, ar.at(oU[(last + 1) & 1]), n * sizeof(T), hipMemcpyDeviceToDevice, s));
    CILREG_CK(hipStreamSynchronize(s));
    return 0;
}

template int rof_solve<float, float>(const float*, float*, float, int, float, float, int, int, int, const Opts&,
                                     hipStream_t, int*, float*);

}  // namespace cilreg
