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

// The fluxes D1 (along y), D2 (along x) and, in 3D, D3 (along z) at
// (px, py, z) from one load of the 7-point mirrored neighbourhood.
template <int DIM, typename I, typename T>
__device__ inline void rof_fluxes(const RofArgs<T>& a, int px, int py, int z, T& d1, T& d2, T& d3) {
    const Geom& g = a.g;
    const I sxy = (I)g.sxy, nx = (I)g.nx;
    const I zo = (I)z * sxy, c = zo + (I)py * nx + (I)px;
    const T uc = ldx<T>(a.U, c);
    const T uip = ldx<T>(a.U, zo + (I)py * nx + (I)mirror_p(px, g.nx));
    const T uim = ldx<T>(a.U, zo + (I)py * nx + (I)mirror_m(px, g.nx));
    const T ujp = ldx<T>(a.U, zo + (I)mirror_p(py, g.ny) * nx + (I)px);
    const T ujm = ldx<T>(a.U, zo + (I)mirror_m(py, g.ny) * nx + (I)px);
    T ukp = 0, ukm = 0;
    if (DIM == 3) {
        ukp = ldx<T>(a.U, (I)mirror_p(z, g.nz) * sxy + (I)py * nx + (I)px);
        ukm = ldx<T>(a.U, (I)mirror_m(z, g.nz) * sxy + (I)py * nx + (I)px);
    }
    const T fi = uip - uc, bi = uc - uim, fj = ujp - uc, bj = uc - ujm;
    const T mi = Mixed<T>::minmod_sq(fi, bi), mj = Mixed<T>::minmod_sq(fj, bj);
    if (DIM == 3) {
        const T fk = ukp - uc, bk = uc - ukm;
        const T mk = Mixed<T>::minmod_sq(fk, bk);
        // ROF_TV_core.c:131-137 (double sqrt), :194-200, :257-263
        d1 = fj / Mixed<T>::rof_sqrt_d(fj * fj + mi + mk);
        d2 = fi / Mixed<T>::rof_sqrt_f(fi * fi + mj + mk);
        d3 = fk / Mixed<T>::rof_sqrt_f(fk * fk + mj + mi);
    } else {
        // ROF_TV_core.c:157-161, :220-224
        d1 = fj / Mixed<T>::rof_sqrt_f(fj * fj + mi);
        d2 = fi / Mixed<T>::rof_sqrt_f(fi * fi + mj);
        d3 = 0;
    }
}

template <int DIM, int TY, typename T, typename I>
__global__ __launch_bounds__(TX * TY) void rof_iter_kernel(RofArgs<T> a) {
    if (solve_stopped(a.ctrl)) return;
    const Geom& g = a.g;
    const TileCoord tc = tile_of_block(g, TY);
    constexpr int NT = TX * TY, NH = TX + TY;
    const int lx = threadIdx.x, ly = threadIdx.y, tid = ly * TX + lx;
    const int x = tc.x0 + lx, y = tc.y0 + ly;
    const bool own = x < g.nx && y < g.ny;
    // -x halo column (D2, TY points) and -y halo row (D1, TX points) on the
    // low or high waves depending on the block index
    const int hbase = (tc.lin & 1) ? NT - NH : 0;
    const int h = tid - hbase;
    const bool hlane = h >= 0 && h < NH;
    const bool hcol = h < TY;
    const int hx = hcol ? tc.x0 - 1 : tc.x0 + h - TY;
    const int hy = hcol ? tc.y0 + h : tc.y0 - 1;
    const bool hown = hlane && hx >= 0 && hy >= 0 && hx < g.nx && hy < g.ny;
    const int ox = own ? x : 0, oy = own ? y : 0;
    const int hcx = hown ? hx : 0, hcy = hown ? hy : 0;

    __shared__ T D1l[TY + 1][TX];       // [ly + 1][lx]; row 0 = y0 - 1
    __shared__ T D2l[TY][TX + 1];       // [ly][lx + 1]; column 0 = x0 - 1
    __shared__ double red[8];
    const I sxy = (I)g.sxy, nx = (I)g.nx;

    // D3 at the mirrored z-predecessor of the chunk's first plane
    T d3m = 0;
    if (DIM == 3 && g.nz > 1) {
        T t1, t2;
        rof_fluxes<DIM, I>(a, ox, oy, mirror_m(tc.z0, g.nz), t1, t2, d3m);
    }
    double sdd = 0.0, suu = 0.0;
    for (int z = tc.z0; z < tc.z1; ++z) {
        T d1, d2, d3;
        rof_fluxes<DIM, I>(a, ox, oy, z, d1, d2, d3);
        D1l[ly + 1][lx] = d1;
        D2l[ly][lx + 1] = d2;
        if (hlane) {
            T h1, h2, h3;
            rof_fluxes<DIM, I>(a, hcx, hcy, z, h1, h2, h3);
            if (hown) {
                if (hcol) D2l[h][0] = h2;
                else D1l[0][h - TY] = h1;
            }
        }
        __syncthreads();
        if (own) {
            const I i = (I)z * sxy + (I)y * nx + (I)x;
            const int jm = mirror_m(y, g.ny), im = mirror_m(x, g.nx);
            const T dv1 = d1 - D1l[jm - tc.y0 + 1][lx];
            const T dv2 = d2 - D2l[ly][im - tc.x0 + 1];
            T dv = dv1 + dv2;
            if (DIM == 3) {
                // at z = 0 the mirrored predecessor is plane 1 (d3m from the prologue)
                dv = dv + (d3 - d3m);
            }
            const T u = ldx<T>(a.U, i);
            const T un = u + a.tau * (a.lambda * dv - (u - ldx<T>(a.f, i)));
            if (a.check) {
                const T df = un - u;
                sdd += (double)df * (double)df;
                suu += (double)un * (double)un;
            }
            stnt<T>(a.Uout, i, un);
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
