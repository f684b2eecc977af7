// dtv.hip -- FGP-dTV (directional total variation, Ehrhardt & Betcke): FGP
// whose dual field and objective gradient are projected onto the orthogonal
// complement of a reference image's normalised gradient, one fused kernel per
// iteration.
//
// Reference algorithm: src/Core/regularisers_CPU/FGP_dTV_core.c:46-380.
// Before the iterations, from the reference image u_ref:
//   B = grad(u_ref) / sqrt(|grad(u_ref)|^2 + eta^2)        (GradNorm, :188-208, :279-306)
// with forward differences that read 0 past the last index.  Per iteration k,
// with m_k = (t_{k-1} - 1) / t_k:
//   R   = P_k + m_k (P_k - P_{k-1})                        (Rupd, :263-274, :368-380)
//   R'  = R - <R, B> B                                      (ProjectVect, :210-223, :308-323)
//   D_k = f - lambda * sum_d (R'_d(x) - R'_d(x - e_d)), R'_d(-1) = 0, max(., 0) if nonneg
//   g_d = D_k(x) - D_k(x + e_d) (0 at the last index);  g' = g - <g, B> B
//   P_{k+1} = proj(R' + g' / (26 lambda))                  (Grad_dfunc, :240-262, :342-367)
// with 1/(8 lambda) in 2D, and t_{k+1} = (1 + sqrt(1 + 4 t_k^2)) / 2 taken
// through a double square root (:95, :149).  The output is D of the last
// iteration; the stopping rule is FGP's (every 5th iteration, four hits).
//
// HBM traffic per voxel-iteration (FP32): read f, B(3), P_k(3), P_{k-1}(3),
// write P_{k+1}(3) = 52 B in 3D (40 B in 2D); R and R' are recomputed, never
// stored; B is computed once per solve.
//
// Fusion: a workgroup owns a 64 x TY column tile and marches through a chunk
// of z.  For plane z+1 it evaluates R' on the tile and on the ring of points
// the divergence reads (-x / -y neighbours of the tile and of its +x / +y halo)
// into two LDS planes, then D on the tile plus its +x / +y halo into an LDS
// plane (double-buffered across z), then updates P on plane z.  R'_3 of the
// plane below and B of the own point stay in registers.
#include "solvers.hpp"

namespace cilreg {

template <typename T>
struct DtvArgs {
    const T* f;
    const T* B[3];      // normalised reference gradient
    const T* P[3];      // P_k
    const T* Pp[3];     // P_{k-1}
    T* Pn[3];           // P_{k+1}
    T* Dout;            // D_k (written when write_d)
    const T* Dprev;     // D_{k-1} (read when check)
    T lambda, step, mult;
    int write_d, check, nonneg;
    double* partials;
    const Ctrl* ctrl;
    Geom g;
};

// B from the reference image; one thread per voxel (runs once per solve).
template <int DIM, typename T>
__global__ __launch_bounds__(256) void dtv_refgrad_kernel(const T* ref, T* B1, T* B2, T* B3, T eta, int nx, int ny,
                                                          int nz) {
    const size_t sxy = (size_t)nx * ny, n = sxy * (size_t)nz;
    const T eta_sq = eta * eta;
    for (size_t c = (size_t)blockIdx.x * 256 + threadIdx.x; c < n; c += (size_t)gridDim.x * 256) {
        const size_t k = c / sxy, r = c - k * sxy;
        const int j = (int)(r / (size_t)nx), i = (int)(r - (size_t)j * nx);
        const T rc = ref[c];
        const T v1 = (i == nx - 1) ? (T)0 : ref[c + 1];
        const T v2 = (j == ny - 1) ? (T)0 : ref[c + nx];
        const T g1 = v1 - rc, g2 = v2 - rc;
        T g3 = 0;
        T m = g1 * g1 + g2 * g2;
        if (DIM == 3) {
            const T v3 = ((int)k == nz - 1) ? (T)0 : ref[c + sxy];
            g3 = v3 - rc;
            m = m + g3 * g3;
        }
        m = (T)sqrt((double)(m + eta_sq));
        B1[c] = g1 / m;
        B2[c] = g2 / m;
        if (DIM == 3) B3[c] = g3 / m;
    }
}

template <int DIM, int TY, typename T, bool ISO, typename I>
__device__ inline void dtv_body(const DtvArgs<T>& a, const TileCoord& tc) {
    const Geom& g = a.g;
    constexpr int NT = TX * TY;
    constexpr int NH = TX + TY;                // D halo: +x column and +y row
    constexpr int NR = 2 * TX + 2 * TY + 2;    // R' ring (includes the D halo points first)
    static_assert(NR <= NT, "ring lanes must fit the workgroup");
    const int lx = threadIdx.x, ly = threadIdx.y, tid = ly * TX + lx;
    const int x = tc.x0 + lx, y = tc.y0 + ly;
    const bool own = x < g.nx && y < g.ny;
    // ring lanes alternate between the low and high waves with the block index
    const int hbase = (tc.lin & 1) ? NT - NR : 0;
    const int h = tid - hbase;
    const bool rlane = h >= 0 && h < NR;
    const bool hlane = h >= 0 && h < NH;
    // ring point (rx, ry): lanes [0, TY) the +x column on the tile's rows,
    // [TY, TY+TX) the +y row -- these also evaluate D --, then the +x column
    // at y0-1, the -x column on rows y0 .. y0+TY and the -y row on x0 .. x0+63
    int rx = 0, ry = 0;
    if (h < TY) { rx = tc.x0 + TX; ry = tc.y0 + h; }
    else if (h < TY + TX) { rx = tc.x0 + h - TY; ry = tc.y0 + TY; }
    else if (h == TY + TX) { rx = tc.x0 + TX; ry = tc.y0 - 1; }
    else if (h < 2 * TY + TX + 2) { rx = tc.x0 - 1; ry = tc.y0 + (h - TY - TX - 1); }
    else { rx = tc.x0 + (h - 2 * TY - TX - 2); ry = tc.y0 - 1; }
    const bool rin = rlane && rx >= 0 && ry >= 0 && rx < g.nx && ry < g.ny;
    const int sx = rx - tc.x0 + 1, sy = ry - tc.y0 + 1;   // slot in the R' planes
    // points outside the volume evaluate a valid stand-in and store nothing;
    // their values are only read where a boundary rule selects 0 instead
    const int ox = own ? x : 0, oy = own ? y : 0;
    const int rcx = rin ? rx : 0, rcy = rin ? ry : 0;

    __shared__ T Rs1[TY + 2][TX + 2];   // R'_1 on [x0-1, x0+64] x [y0-1, y0+TY]
    __shared__ T Rs2[TY + 2][TX + 2];   // R'_2 likewise
    __shared__ T Ds[2][TY + 1][TX + 1];
    __shared__ double red[8];

    const I sxy = (I)g.sxy, nx = (I)g.nx;
    const T lam = a.lambda, mult = a.mult;
    // R' at element i, and B there
    auto rproj = [&](I i, T& r1, T& r2, T& r3, T& b1, T& b2, T& b3) {
        T p = ldx<T>(a.P[0], i), q = ldx<T>(a.Pp[0], i);
        r1 = p + mult * (p - q);
        p = ldx<T>(a.P[1], i);
        q = ldx<T>(a.Pp[1], i);
        r2 = p + mult * (p - q);
        b1 = ldx<T>(a.B[0], i);
        b2 = ldx<T>(a.B[1], i);
        T ip = r1 * b1 + r2 * b2;
        if (DIM == 3) {
            p = ldx<T>(a.P[2], i);
            q = ldx<T>(a.Pp[2], i);
            r3 = p + mult * (p - q);
            b3 = ldx<T>(a.B[2], i);
            ip = ip + r3 * b3;
            r3 = r3 - ip * b3;
        } else {
            r3 = 0;
            b3 = 0;
        }
        r1 = r1 - ip * b1;
        r2 = r2 - ip * b2;
    };
    // D at (px, py, z) from R' there, R'_1 / R'_2 of the -x / -y slots and R'_3 below
    auto dval = [&](int px, int py, int z, int ssx, int ssy, T r1, T r2, T r3, T r3below) -> T {
        const I i = (I)z * sxy + (I)py * nx + (I)px;
        const T fv = ldx<T>(a.f, i);
        const T b1 = px > 0 ? Rs1[ssy][ssx - 1] : (T)0;
        const T b2 = py > 0 ? Rs2[ssy - 1][ssx] : (T)0;
        T d;
        if (DIM == 3) {
            const T b3 = z > 0 ? r3below : (T)0;
            d = fv - lam * (r1 + r2 + r3 - b1 - b2 - b3);
        } else {
            d = fv - lam * (r1 + r2 - b1 - b2);
        }
        if (a.nonneg && d < (T)0) d = (T)0;
        return d;
    };
    // R' of plane z on the tile and the ring into LDS; own values and the
    // D-halo lanes' values are returned for the D stage and the update
    auto rplane = [&](int z, T& r1, T& r2, T& r3, T& b1, T& b2, T& b3, T& h1, T& h2, T& h3) {
        const I base = (I)z * sxy;
        rproj(base + (I)oy * nx + (I)ox, r1, r2, r3, b1, b2, b3);
        Rs1[ly + 1][lx + 1] = r1;
        Rs2[ly + 1][lx + 1] = r2;
        if (rlane) {
            T c1, c2, c3;
            rproj(base + (I)rcy * nx + (I)rcx, h1, h2, h3, c1, c2, c3);
            Rs1[sy][sx] = h1;
            Rs2[sy][sx] = h2;
        }
    };

    T cr1, cr2, cr3, cb1, cb2, cb3;   // own R' and B, plane z
    T hr1 = 0, hr2 = 0, hr3 = 0;      // ring lane's R', plane z
    T r3b = 0, hr3b = 0;              // R'_3 one plane below the chunk start
    if (DIM == 3 && tc.z0 > 0) {
        const I base = (I)(tc.z0 - 1) * sxy;
        T t1, t2, c1, c2, c3;
        rproj(base + (I)oy * nx + (I)ox, t1, t2, r3b, c1, c2, c3);
        if (hlane) rproj(base + (I)rcy * nx + (I)rcx, t1, t2, hr3b, c1, c2, c3);
    }
    rplane(tc.z0, cr1, cr2, cr3, cb1, cb2, cb3, hr1, hr2, hr3);
    __syncthreads();
    Ds[0][ly][lx] = dval(ox, oy, tc.z0, lx + 1, ly + 1, cr1, cr2, cr3, r3b);
    if (hlane) {
        const T d = dval(rcx, rcy, tc.z0, sx, sy, hr1, hr2, hr3, hr3b);
        if (rin) Ds[0][sy - 1][sx - 1] = d;
    }
    __syncthreads();

    double sdd = 0.0, suu = 0.0;
    int cur = 0;
    for (int z = tc.z0; z < tc.z1; ++z) {
        T nr1 = 0, nr2 = 0, nr3 = 0, nb1 = 0, nb2 = 0, nb3 = 0, hn1 = 0, hn2 = 0, hn3 = 0;
        if (DIM == 3 && z + 1 < g.nz) {
            rplane(z + 1, nr1, nr2, nr3, nb1, nb2, nb3, hn1, hn2, hn3);
            __syncthreads();
            Ds[cur ^ 1][ly][lx] = dval(ox, oy, z + 1, lx + 1, ly + 1, nr1, nr2, nr3, cr3);
            if (hlane) {
                const T d = dval(rcx, rcy, z + 1, sx, sy, hn1, hn2, hn3, hr3);
                if (rin) Ds[cur ^ 1][sy - 1][sx - 1] = d;
            }
        }
        __syncthreads();
        if (own) {
            const I i = (I)z * sxy + (I)y * nx + (I)x;
            const T d = Ds[cur][ly][lx];
            T g1 = (x == g.nx - 1) ? (T)0 : d - Ds[cur][ly][lx + 1];
            T g2 = (y == g.ny - 1) ? (T)0 : d - Ds[cur][ly + 1][lx];
            T g3 = 0;
            T ip = g1 * cb1 + g2 * cb2;
            if (DIM == 3) {
                g3 = (z == g.nz - 1) ? (T)0 : d - Ds[cur ^ 1][ly][lx];
                ip = ip + g3 * cb3;
            }
            g1 = g1 - ip * cb1;
            g2 = g2 - ip * cb2;
            T p1 = cr1 + a.step * g1;
            T p2 = cr2 + a.step * g2;
            T p3 = 0;
            if (DIM == 3) {
                g3 = g3 - ip * cb3;
                p3 = cr3 + a.step * g3;
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
        cr1 = nr1; cr2 = nr2; cr3 = nr3;
        cb1 = nb1; cb2 = nb2; cb3 = nb3;
        hr3 = hn3;
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

template <int DIM, int TY, typename T, bool ISO, typename I>
__global__ __launch_bounds__(TX * TY) void dtv_iter_kernel(DtvArgs<T> a) {
    if (solve_stopped(a.ctrl)) return;
    dtv_body<DIM, TY, T, ISO, I>(a, tile_of_block(a.g, TY));
}

template <int DIM, int TY, typename T, typename I>
static const void* dtv_kernel_ptr(bool iso) {
    return iso ? (const void*)dtv_iter_kernel<DIM, TY, T, true, I> : (const void*)dtv_iter_kernel<DIM, TY, T, false, I>;
}

template <int DIM, int TY, typename T>
static void launch_dtv(const DtvArgs<T>& a, bool iso, bool small, hipStream_t s) {
    dim3 block(TX, TY), grid(a.g.nblocks);
    if (small) {
        if (iso) hipLaunchKernelGGL((dtv_iter_kernel<DIM, TY, T, true, uint32_t>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((dtv_iter_kernel<DIM, TY, T, false, uint32_t>), grid, block, 0, s, a);
    } else {
        if (iso) hipLaunchKernelGGL((dtv_iter_kernel<DIM, TY, T, true, uint64_t>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((dtv_iter_kernel<DIM, TY, T, false, uint64_t>), grid, block, 0, s, a);
    }
}

// Device-resident FGP-dTV solve.  d_out receives D of the final iteration
// (left untouched when iter == 0); with epsil != 0 its prior contents seed the
// first residual (FGP_dTV_core.c:78,131).
template <typename T>
int dtv_solve(const T* d_f, const T* d_ref, T* d_out, T lambda, int iter, T epsil, T eta, int method, int nonneg,
              int nx, int ny, int nz, const Opts& opts, hipStream_t s, int* iters_done, float* re) {
    const int dim = nz <= 1 ? 2 : 3;
    if (dim == 2) nz = 1;
    *iters_done = 0;
    *re = 0.f;
    if (iter <= 0) return 0;
    const size_t n = (size_t)nx * ny * nz;
    const bool small = !opts.wide_index && fits_u32(n, sizeof(T));
    const int ty = (opts.tile_y == 4 || dim == 2) ? 4 : 8;
    const bool iso = method == 0;
    Geom g;
    if (dim == 2)
        g = make_geom(nx, ny, nz, 4, 1);
    else if (ty == 8)
        g = plan_geom(small ? dtv_kernel_ptr<3, 8, T, uint32_t>(iso) : dtv_kernel_ptr<3, 8, T, uint64_t>(iso), TX * 8,
                      nx, ny, nz, 8, opts.zchunk);
    else
        g = plan_geom(small ? dtv_kernel_ptr<3, 4, T, uint32_t>(iso) : dtv_kernel_ptr<3, 4, T, uint64_t>(iso), TX * 4,
                      nx, ny, nz, 4, opts.zchunk);

    Arena ar;
    size_t oP[3][3], oB[3], oD[2], oCtrl, oPart;
    for (int b = 0; b < 3; ++b)
        for (int c = 0; c < dim; ++c) oP[b][c] = ar.reserve(n * sizeof(T));
    for (int c = 0; c < dim; ++c) oB[c] = ar.reserve(n * sizeof(T));
    for (int b = 0; b < 2; ++b) oD[b] = ar.reserve(n * sizeof(T));
    oCtrl = ar.reserve(sizeof(Ctrl));
    oPart = ar.reserve((size_t)g.nblocks * 2 * sizeof(double));
    if (ar.allocate() != 0) return -1;
    for (int c = 0; c < dim; ++c) {   // P_0 = P_{-1} = 0
        launch_zero(ar.at(oP[0][c]), n * sizeof(T), s);
        launch_zero(ar.at(oP[2][c]), n * sizeof(T), s);
    }
    auto Pbuf = [&](int j, int c) -> T* { return (T*)ar.at(oP[((j % 3) + 3) % 3][c < dim ? c : 0]); };
    T* B[3];
    for (int c = 0; c < 3; ++c) B[c] = (T*)ar.at(oB[c < dim ? c : 0]);
    {
        size_t blocks = (n + 255) / 256;
        if (blocks > 65536) blocks = 65536;
        if (dim == 3)
            hipLaunchKernelGGL((dtv_refgrad_kernel<3, T>), dim3((unsigned)blocks), dim3(256), 0, s, d_ref, B[0], B[1],
                               B[2], eta, nx, ny, nz);
        else
            hipLaunchKernelGGL((dtv_refgrad_kernel<2, T>), dim3((unsigned)blocks), dim3(256), 0, s, d_ref, B[0], B[1],
                               B[2], eta, nx, ny, nz);
    }
    const bool use_tol = epsil != (T)0;
    T* Dbuf[2] = {(T*)ar.at(oD[0]), (T*)ar.at(oD[1])};
    if (use_tol) launch_copy(Dbuf[1], d_out, n * sizeof(T), s);   // D_{-1} = caller's Output
    Ctrl* ctrl = (Ctrl*)ar.at(oCtrl);
    double* partials = (double*)ar.at(oPart);
    const int ce = opts.check_every;
    auto needs_d = [&](int k) { return k == iter - 1 || (use_tol && (k % ce == 0 || (k + 1) % ce == 0)); };
    const T stepc = (T)1 / ((dim == 3 ? (T)26 : (T)8) * lambda);

    // t_k with a double square root and a double sum rounded once
    // (FGP_dTV_core.c:95,149); mult[k] multiplies (P_k - P_{k-1}) in iteration k
    std::vector<T> mult(iter + 1);
    {
        T tk = 1, m = 0;
        for (int k = 0; k <= iter; ++k) {
            mult[k] = m;
            const T tkp1 = (T)((1.0 + sqrt((double)((T)1 + (T)4 * tk * tk))) * 0.5);
            m = (tk - (T)1) / tkp1;
            tk = tkp1;
        }
    }
    auto step = [&](int k, bool check) {
        DtvArgs<T> a;
        a.f = d_f;
        for (int c = 0; c < 3; ++c) {
            a.B[c] = B[c];
            a.P[c] = Pbuf(k, c);
            a.Pp[c] = Pbuf(k - 1, c);
            a.Pn[c] = Pbuf(k + 1, c);
        }
        a.Dout = Dbuf[k & 1];
        a.Dprev = Dbuf[(k + 1) & 1];
        a.lambda = lambda;
        a.step = stepc;
        a.mult = mult[k];
        a.write_d = needs_d(k) || check;
        a.check = check;
        a.nonneg = nonneg != 0;
        a.partials = partials;
        a.ctrl = ctrl;
        a.g = g;
        if (dim == 3) {
            if (ty == 8) launch_dtv<3, 8>(a, iso, small, s);
            else launch_dtv<3, 4>(a, iso, small, s);
        } else {
            launch_dtv<2, 4>(a, iso, small, s);
        }
    };
    int done = 0;
    if (drive_iterations(iter, (float)epsil, opts, ctrl, partials, g.nblocks, s, step, &done, re) != 0) return -1;
    *iters_done = done;
    const int last = (done < iter) ? done : iter - 1;
    CILREG_CK(hipMemcpyAsync(d_out, Dbuf[last & 1], n * sizeof(T), hipMemcpyDeviceToDevice, s));
    CILREG_CK(hipStreamSynchronize(s));
    return 0;
}
template int dtv_solve<float>(const float*, const float*, float*, float, int, float, float, int, int, int, int, int,
                              const Opts&, hipStream_t, int*, float*);

}  // namespace cilreg
