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

namespace cilreg {

template <typename T>
struct RofArgs {
    const T* f;
    const T* U;
    T* Uout;
    T lambda, tau;
    int check;
    double* partials;
    const Ctrl* ctrl;
    Geom g;
};

__device__ inline int mirror_p(int i, int n) { return i + 1 >= n ? (i - 1 >= 0 ? i - 1 : 0) : i + 1; }
__device__ inline int mirror_m(int i, int n) { return i - 1 < 0 ? (i + 1 < n ? i + 1 : 0) : i - 1; }

// flux along `dir` (0: y, 1: x, 2: z) at (px, py, z)
template <int DIM, typename T>
__device__ inline T rof_flux(const RofArgs<T>& a, int px, int py, int z, int dir) {
    const Geom& g = a.g;
    const size_t sxy = (size_t)g.sxy, nx = (size_t)g.nx;
    const size_t row = (size_t)z * sxy + (size_t)py * nx;
    const size_t c = row + px;
    const T uc = a.U[c];
    const T fi = a.U[row + mirror_p(px, g.nx)] - uc, bi = uc - a.U[row + mirror_m(px, g.nx)];
    const T fj = a.U[(size_t)z * sxy + (size_t)mirror_p(py, g.ny) * nx + px] - uc;
    const T bj = uc - a.U[(size_t)z * sxy + (size_t)mirror_m(py, g.ny) * nx + px];
    T fk = 0, bk = 0;
    if (DIM == 3) {
        fk = a.U[(size_t)mirror_p(z, g.nz) * sxy + (size_t)py * nx + px] - uc;
        bk = uc - a.U[(size_t)mirror_m(z, g.nz) * sxy + (size_t)py * nx + px];
    }
    T num, s;
    if (dir == 0) {
        num = fj;
        s = num * num + Mixed<T>::minmod_sq(fi, bi);
        if (DIM == 3) s = s + Mixed<T>::minmod_sq(fk, bk);
        return num / (DIM == 3 ? Mixed<T>::rof_sqrt_d(s) : Mixed<T>::rof_sqrt_f(s));
    } else if (dir == 1) {
        num = fi;
        s = num * num + Mixed<T>::minmod_sq(fj, bj);
        if (DIM == 3) s = s + Mixed<T>::minmod_sq(fk, bk);
        return num / Mixed<T>::rof_sqrt_f(s);
    } else {
        num = fk;
        s = num * num + Mixed<T>::minmod_sq(fj, bj);
        s = s + Mixed<T>::minmod_sq(fi, bi);
        return num / Mixed<T>::rof_sqrt_f(s);
    }
}

template <int DIM, int TY, typename T>
__global__ __launch_bounds__(TX * TY) void rof_iter_kernel(RofArgs<T> a) {
    if (solve_stopped(a.ctrl)) return;
    const Geom& g = a.g;
    const TileCoord tc = tile_of_block(g, TY);
    const int lx = threadIdx.x, ly = threadIdx.y, tid = ly * TX + lx;
    const int x = tc.x0 + lx, y = tc.y0 + ly;
    const bool own = x < g.nx && y < g.ny;
    // -x halo column (D2) for tid < TY, -y halo row (D1) for TY <= tid < TY + TX
    int hx = -1, hy = -1;
    if (tid < TY) { hx = tc.x0 - 1; hy = tc.y0 + tid; }
    else if (tid < TY + TX) { hx = tc.x0 + tid - TY; hy = tc.y0 - 1; }
    const bool hown = hx >= 0 && hy >= 0 && hx < g.nx && hy < g.ny;

    __shared__ T D1l[TY + 1][TX];       // [ly + 1][lx]; row 0 = y0 - 1
    __shared__ T D2l[TY][TX + 1];       // [ly][lx + 1]; column 0 = x0 - 1
    __shared__ double red[8];
    const size_t sxy = (size_t)g.sxy, nx = (size_t)g.nx;

    // D3 at the mirrored z-predecessor of the chunk's first plane
    T d3m = 0;
    if (DIM == 3 && own && g.nz > 1) d3m = rof_flux<DIM>(a, x, y, mirror_m(tc.z0, g.nz), 2);

    double sdd = 0.0, suu = 0.0;
    for (int z = tc.z0; z < tc.z1; ++z) {
        T d1 = 0, d2 = 0, d3 = 0;
        if (own) {
            d1 = rof_flux<DIM>(a, x, y, z, 0);
            d2 = rof_flux<DIM>(a, x, y, z, 1);
            if (DIM == 3) d3 = rof_flux<DIM>(a, x, y, z, 2);
            D1l[ly + 1][lx] = d1;
            D2l[ly][lx + 1] = d2;
        }
        if (hown) {
            if (tid < TY) D2l[tid][0] = rof_flux<DIM>(a, hx, hy, z, 1);
            else D1l[0][tid - TY] = rof_flux<DIM>(a, hx, hy, z, 0);
        }
        __syncthreads();
        if (own) {
            const size_t i = (size_t)z * sxy + (size_t)y * nx + x;
            const int jm = mirror_m(y, g.ny), im = mirror_m(x, g.nx);
            const T dv1 = d1 - D1l[jm - tc.y0 + 1][lx];
            const T dv2 = d2 - D2l[ly][im - tc.x0 + 1];
            T dv = dv1 + dv2;
            if (DIM == 3) {
                T d3p = d3m;
                if (z == 0) d3p = (g.nz > 1) ? d3m : d3;
                dv = dv + (d3 - d3p);
            }
            const T u = a.U[i];
            const T un = u + a.tau * (a.lambda * dv - (u - a.f[i]));
            if (a.check) {
                const T df = un - u;
                sdd += (double)df * (double)df;
                suu += (double)un * (double)un;
            }
            a.Uout[i] = un;
        }
        __syncthreads();
        d3m = d3;
    }
    if (a.check) {
        block_sum2(sdd, suu, red);
        if (tid == 0) {
            a.partials[2 * tc.lin] = sdd;
            a.partials[2 * tc.lin + 1] = suu;
        }
    }
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
    int ty = opts.tile_y, zc = opts.zchunk;
    auto_tiling(nx, ny, nz, &ty, &zc);
    const Geom g = make_geom(nx, ny, nz, ty, zc);
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
        dim3 grid(g.nblocks);
        if (dim == 3) {
            if (ty == 8) hipLaunchKernelGGL((rof_iter_kernel<3, 8, T>), grid, dim3(TX, 8), 0, s, a);
            else hipLaunchKernelGGL((rof_iter_kernel<3, 4, T>), grid, dim3(TX, 4), 0, s, a);
        } else {
            hipLaunchKernelGGL((rof_iter_kernel<2, 4, T>), grid, dim3(TX, 4), 0, s, a);
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
