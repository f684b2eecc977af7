// sb.hip -- SB-TV (split Bregman), two kernels per iteration.
//
// Reference algorithm: src/Core/regularisers_CPU/SB_TV_core.c:38-371.  With
// mu' = 1/mu and lam = 2 mu', per iteration:
//   U1 = J(U0), U2 = J(U1)   two Jacobi sweeps (named gauss_seidel in the
//                            reference, :149-173, :254-282) with mirror
//                            boundaries, reading D and B at x and x - e_d
//   v_d = (U2(x + e_d) - U2(x)) + B_d      forward difference, mirrored at the
//                                          last index
//   D = shrink(v, 1/lam)     iso: vector shrinkage; aniso: per component
//   B = B + (U2(x + e_d) - U2(x)) - D      (:233-248, :353-371)
// The residual compares U2 with U1 (:123-133).
//
// Kernel 1 runs the first sweep.  Kernel 2 runs the second sweep on the tile
// plus its +x / +y halo into an LDS plane (double-buffered across z), then the
// shrinkage and Bregman update of the tile.  HBM traffic per voxel-iteration:
// sweep 1 reads U0, f, D(3), B(3) and writes U1 (36 B); kernel 2 reads U1, f,
// D, B and writes U2, D, B (60 B).
#include "solvers.hpp"

namespace cilreg {

template <typename T>
struct SbArgs {
    const T* f;
    const T* Uin;      // sweep input
    T* Uout;           // sweep output (kernel 1) or U2 (kernel 2)
    const T* D[3];
    const T* B[3];
    T* Dout[3];
    T* Bout[3];
    T lam, mu, nc, thr;   // lam = 2/mu_in, mu = 1/mu_in, nc = 1/(mu + (4|6) lam), thr = 1/lam
    int check;
    double* partials;
    const Ctrl* ctrl;
    Geom g;
};

__device__ inline int sb_p(int i, int n) { return i + 1 == n ? (i - 1 >= 0 ? i - 1 : 0) : i + 1; }
__device__ inline int sb_m(int i, int n) { return i - 1 < 0 ? (i + 1 < n ? i + 1 : 0) : i - 1; }

// one Jacobi sweep value at (px, py, z) from a.Uin
template <int DIM, typename I, typename T>
__device__ inline T sb_sweep_at(const SbArgs<T>& a, int px, int py, int z) {
    const Geom& g = a.g;
    const I sxy = (I)g.sxy, nx = (I)g.nx;
    const I zo = (I)z * sxy;
    const I c = zo + (I)py * nx + (I)px;
    const int ip = sb_p(px, g.nx), im = sb_m(px, g.nx), jp = sb_p(py, g.ny), jm = sb_m(py, g.ny);
    const I cim = zo + (I)py * nx + (I)im, cjm = zo + (I)jm * nx + (I)px;
    const I cip = zo + (I)py * nx + (I)ip, cjp = zo + (I)jp * nx + (I)px;
    const T* U = a.Uin;
    auto L = [](const T* p, I e) { return ldx<T>(p, e); };
    T acc;
    if (DIM == 2) {
        acc = L(a.D[0], cim) - L(a.D[0], c) + L(a.D[1], cjm) - L(a.D[1], c) - L(a.B[0], cim) + L(a.B[0], c) -
              L(a.B[1], cjm) + L(a.B[1], c);
        acc += L(U, cip) + L(U, cim) + L(U, cjp) + L(U, cjm);
    } else {
        const int kp = sb_p(z, g.nz), km = sb_m(z, g.nz);
        const I ckm = (I)km * sxy + (I)py * nx + (I)px, ckp = (I)kp * sxy + (I)py * nx + (I)px;
        const T dsum = L(a.D[0], cim) - L(a.D[0], c) + L(a.D[1], cjm) - L(a.D[1], c) + L(a.D[2], ckm) - L(a.D[2], c);
        const T bsum = -L(a.B[0], cim) + L(a.B[0], c) - L(a.B[1], cjm) + L(a.B[1], c) - L(a.B[2], ckm) + L(a.B[2], c);
        acc = dsum + bsum;
        acc += L(U, cip) + L(U, cim) + L(U, cjp) + L(U, cjm) + L(U, ckp) + L(U, ckm);
    }
    acc *= a.lam;
    acc += a.mu * L(a.f, c);
    return a.nc * acc;
}

template <int DIM, int TY, typename T, typename I>
__global__ __launch_bounds__(TX * TY) void sb_sweep_kernel(SbArgs<T> a) {
    if (solve_stopped(a.ctrl)) return;
    const Geom& g = a.g;
    const TileCoord tc = tile_of_block(g, TY);
    const int x = tc.x0 + (int)threadIdx.x, y = tc.y0 + (int)threadIdx.y;
    if (x >= g.nx || y >= g.ny) return;
    for (int z = tc.z0; z < tc.z1; ++z)
        stx<T>(a.Uout, (I)z * (I)g.sxy + (I)y * (I)g.nx + (I)x, sb_sweep_at<DIM, I>(a, x, y, z));
}

template <int DIM, int TY, typename T, bool ANISO, typename I>
__global__ __launch_bounds__(TX * TY) void sb_update_kernel(SbArgs<T> a) {
    if (solve_stopped(a.ctrl)) return;
    const Geom& g = a.g;
    const TileCoord tc = tile_of_block(g, TY);
    constexpr int NT = TX * TY, NH = TX + TY;
    const int lx = threadIdx.x, ly = threadIdx.y, tid = ly * TX + lx;
    const int x = tc.x0 + lx, y = tc.y0 + ly;
    const bool own = x < g.nx && y < g.ny;
    // +x halo column and +y halo row, on the low or high waves by block index
    const int hbase = (tc.lin & 1) ? NT - NH : 0;
    const int h = tid - hbase;
    const bool hlane = h >= 0 && h < NH;
    int hx, hy, hlx, hly;
    if (h < TY) { hx = tc.x0 + TX; hy = tc.y0 + h; hlx = TX; hly = h; }
    else { hx = tc.x0 + h - TY; hy = tc.y0 + TY; hlx = h - TY; hly = TY; }
    const bool hown = hlane && hx < g.nx && hy < g.ny;
    const int ox = own ? x : 0, oy = own ? y : 0;
    const int hcx = hown ? hx : 0, hcy = hown ? hy : 0;

    __shared__ T Ul[2][TY + 1][TX + 1];
    __shared__ double red[8];
    const I sxy = (I)g.sxy, nx = (I)g.nx;

    T below = 0;   // U2 at the own point, plane z-1
    if (DIM == 3 && tc.z0 > 0) below = sb_sweep_at<DIM, I>(a, ox, oy, tc.z0 - 1);
    Ul[0][ly][lx] = sb_sweep_at<DIM, I>(a, ox, oy, tc.z0);
    if (hlane) { const T v = sb_sweep_at<DIM, I>(a, hcx, hcy, tc.z0); if (hown) Ul[0][hly][hlx] = v; }

    double sdd = 0.0, suu = 0.0;
    int cur = 0;
    for (int z = tc.z0; z < tc.z1; ++z) {
        if (DIM == 3 && z + 1 < g.nz) {
            Ul[cur ^ 1][ly][lx] = sb_sweep_at<DIM, I>(a, ox, oy, z + 1);
            if (hlane) {
                const T v = sb_sweep_at<DIM, I>(a, hcx, hcy, z + 1);
                if (hown) Ul[cur ^ 1][hly][hlx] = v;
            }
        }
        __syncthreads();
        T u2 = 0;
        if (own) {
            const I i = (I)z * sxy + (I)y * nx + (I)x;
            u2 = Ul[cur][ly][lx];
            // mirrored forward neighbours: +1, or -1 at the last index
            T ux, uy;
            if (x + 1 < g.nx) ux = Ul[cur][ly][lx + 1];
            else if (g.nx == 1) ux = u2;
            else ux = (lx > 0) ? Ul[cur][ly][lx - 1] : sb_sweep_at<DIM, I>(a, x - 1, y, z);
            if (y + 1 < g.ny) uy = Ul[cur][ly + 1][lx];
            else if (g.ny == 1) uy = u2;
            else uy = (ly > 0) ? Ul[cur][ly - 1][lx] : sb_sweep_at<DIM, I>(a, x, y - 1, z);
            T gr[3], v[3], dn[3], bv[3];
            gr[0] = ux - u2;
            gr[1] = uy - u2;
            gr[2] = 0;
            if (DIM == 3) {
                const T uz = (z + 1 < g.nz) ? Ul[cur ^ 1][ly][lx] : below;
                gr[2] = uz - u2;
            }
            #pragma unroll
            for (int d = 0; d < DIM; ++d) {
                bv[d] = ldx<T>(a.B[d], i);
                v[d] = gr[d] + bv[d];
            }
            if (ANISO) {
                #pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    T s = fabs(v[d]) - a.thr;
                    if (s < (T)0) s = 0;
                    dn[d] = (v[d] != (T)0) ? (v[d] / fabs(v[d])) * s : (T)0;
                }
            } else {
                T n2 = v[0] * v[0] + v[1] * v[1];
                if (DIM == 3) n2 = n2 + v[2] * v[2];
                const T nrm = dsqrt(n2);
                T s = nrm - a.thr;
                if (s < (T)0) s = 0;
                #pragma unroll
                for (int d = 0; d < DIM; ++d) dn[d] = (nrm != (T)0) ? s * (v[d] / nrm) : (T)0;
            }
            #pragma unroll
            for (int d = 0; d < DIM; ++d) {
                stnt<T>(a.Dout[d], i, dn[d]);
                stnt<T>(a.Bout[d], i, bv[d] + (gr[d] - dn[d]));
            }
            stnt<T>(a.Uout, i, u2);
            if (a.check) {
                const T u1 = ldx<T>(a.Uin, i);
                const T df = u2 - u1;
                sdd += (double)df * (double)df;
                suu += (double)u2 * (double)u2;
            }
        }
        __syncthreads();
        below = u2;
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

template <typename T, typename S>
int sb_solve(const T* d_f, T* d_out, T mu_in, int iter, T epsil, int method, int nx, int ny, int nz, const Opts& opts,
             hipStream_t s, int* iters_done, float* re) {
    const int dim = nz <= 1 ? 2 : 3;
    if (dim == 2) nz = 1;
    const size_t n = (size_t)nx * ny * nz;
    *iters_done = 0;
    *re = 0.f;
    if (iter <= 0) {   // Output = Input (SB_TV_core.c:58,101)
        CILREG_CK(hipMemcpyAsync(d_out, d_f, n * sizeof(T), hipMemcpyDeviceToDevice, s));
        CILREG_CK(hipStreamSynchronize(s));
        return 0;
    }
    const bool small = !opts.wide_index && fits_u32(n, sizeof(T));
    // the SB kernels are built for 64 x 4 tiles
    const Geom g = dim == 2 ? make_geom(nx, ny, nz, 4, 1)
                            : plan_geom(small ? (const void*)sb_update_kernel<3, 4, T, false, uint32_t>
                                              : (const void*)sb_update_kernel<3, 4, T, false, uint64_t>,
                                        TX * 4, nx, ny, nz, 4, opts.zchunk);
    Arena ar;
    size_t oU0, oU1, oD[2][3], oB[2][3], oCtrl, oPart;
    oU0 = ar.reserve(n * sizeof(T));
    oU1 = ar.reserve(n * sizeof(T));
    for (int b = 0; b < 2; ++b)
        for (int c = 0; c < dim; ++c) {
            oD[b][c] = ar.reserve(n * sizeof(T));
            oB[b][c] = ar.reserve(n * sizeof(T));
        }
    oCtrl = ar.reserve(sizeof(Ctrl));
    oPart = ar.reserve((size_t)g.nblocks * 2 * sizeof(double));
    if (ar.allocate() != 0) return -1;
    launch_copy(ar.at(oU0), d_f, n * sizeof(T), s);
    for (int c = 0; c < dim; ++c) {
        launch_zero(ar.at(oD[0][c]), n * sizeof(T), s);
        launch_zero(ar.at(oB[0][c]), n * sizeof(T), s);
    }
    Ctrl* ctrl = (Ctrl*)ar.at(oCtrl);
    double* partials = (double*)ar.at(oPart);
    // constants in the working precision (SB_TV_core.c:45-46, :153, :179, :258)
    const T mu = (T)1 / mu_in;
    const T lam = (T)2 * mu;
    const T nc = (T)1 / (mu + (dim == 3 ? (T)6 : (T)4) * lam);
    const T thr = (T)1 / lam;
    auto step = [&](int k, bool check) {
        SbArgs<T> a;
        const int cur = k & 1, nxt = cur ^ 1;
        a.f = d_f;
        for (int c = 0; c < 3; ++c) {
            const int cc = c < dim ? c : 0;
            a.D[c] = (const T*)ar.at(oD[cur][cc]);
            a.B[c] = (const T*)ar.at(oB[cur][cc]);
            a.Dout[c] = (T*)ar.at(oD[nxt][cc]);
            a.Bout[c] = (T*)ar.at(oB[nxt][cc]);
        }
        a.lam = lam;
        a.mu = mu;
        a.nc = nc;
        a.thr = thr;
        a.partials = partials;
        a.ctrl = ctrl;
        a.g = g;
        dim3 grid(g.nblocks), block(TX, 4);
        // sweep 1: U0 -> U1
        a.Uin = (const T*)ar.at(oU0);
        a.Uout = (T*)ar.at(oU1);
        a.check = 0;
        if (small) {
            if (dim == 3) hipLaunchKernelGGL((sb_sweep_kernel<3, 4, T, uint32_t>), grid, block, 0, s, a);
            else hipLaunchKernelGGL((sb_sweep_kernel<2, 4, T, uint32_t>), grid, block, 0, s, a);
        } else {
            if (dim == 3) hipLaunchKernelGGL((sb_sweep_kernel<3, 4, T, uint64_t>), grid, block, 0, s, a);
            else hipLaunchKernelGGL((sb_sweep_kernel<2, 4, T, uint64_t>), grid, block, 0, s, a);
        }
        // sweep 2 + shrink + Bregman: U1 -> U2 (into U0's buffer)
        a.Uin = (const T*)ar.at(oU1);
        a.Uout = (T*)ar.at(oU0);
        a.check = check;
        if (small) {
            if (dim == 3) {
                if (method == 1) hipLaunchKernelGGL((sb_update_kernel<3, 4, T, true, uint32_t>), grid, block, 0, s, a);
                else hipLaunchKernelGGL((sb_update_kernel<3, 4, T, false, uint32_t>), grid, block, 0, s, a);
            } else {
                if (method == 1) hipLaunchKernelGGL((sb_update_kernel<2, 4, T, true, uint32_t>), grid, block, 0, s, a);
                else hipLaunchKernelGGL((sb_update_kernel<2, 4, T, false, uint32_t>), grid, block, 0, s, a);
            }
        } else {
            if (dim == 3) {
                if (method == 1) hipLaunchKernelGGL((sb_update_kernel<3, 4, T, true, uint64_t>), grid, block, 0, s, a);
                else hipLaunchKernelGGL((sb_update_kernel<3, 4, T, false, uint64_t>), grid, block, 0, s, a);
            } else {
                if (method == 1) hipLaunchKernelGGL((sb_update_kernel<2, 4, T, true, uint64_t>), grid, block, 0, s, a);
                else hipLaunchKernelGGL((sb_update_kernel<2, 4, T, false, uint64_t>), grid, block, 0, s, a);
            }
        }
    };
    int done = 0;
    if (drive_iterations(iter, (float)epsil, opts, ctrl, partials, g.nblocks, s, step, &done, re) != 0) return -1;
    *iters_done = done;
    CILREG_CK(hipMemcpyAsync(d_out, ar.at(oU0), n * sizeof(T), hipMemcpyDeviceToDevice, s));
    CILREG_CK(hipStreamSynchronize(s));
    return 0;
}

template int sb_solve<float, float>(const float*, float*, float, int, float, int, int, int, int, const Opts&,
                                    hipStream_t, int*, float*);

}  // namespace cilreg
