// tgv_kernels.hpp -- device code of the TGV iteration: argument block,
// per-tile body and the single-volume kernel.  The body is shared by the
// single-volume kernel (tgv.hip) and the grouped multi-volume kernel
// (batch.hip); the algorithm and its reference citations are described at the
// top of tgv.hip.
#pragma once
#include "runtime.hpp"

namespace cilreg {

template <typename T, typename S>
struct TgvArgs {
    const T* f;
    const T* U;
    const S* V[3];
    const S* P[3];
    const S* Q[6];
    T* Uout;
    S* Vout[3];
    S* Pout[3];
    S* Qout[6];
    T lambda, tau, sigma, alpha1, alpha0;
    int check;
    double* partials;
    const Ctrl* ctrl;
    Geom g;
};

// LDS channels shared with the +x neighbour (backward difference along x)
// and the +y neighbour (along y).
enum { L_P1, L_P2, L_Q11, L_Q22, L_Q12, L_Q13, L_Q23, L_N };

// Updated and projected duals at (px, py, z).  u and v (u_bar and v_bar at
// the point) are passed in when the caller carries them from the previous
// plane (carried = true) and loaded otherwise; uz / vz return u_bar and v_bar
// one plane up, which are the next plane's own values.  Every load is
// unconditional: neighbour indices are clamped at the last index and the
// boundary rule is applied with selects, so the loads issue back to back.
template <int DIM, typename I, typename T, typename S>
__device__ inline void tgv_dual_at(const TgvArgs<T, S>& a, int px, int py, int z, bool carried, T& u, T v[3],
                                   T& uz, T vz[3], T p[3], T q[6]) {
    const Geom& g = a.g;
    const I sxy = (I)g.sxy, nx = (I)g.nx;
    const I i = (I)z * sxy + (I)py * nx + (I)px;
    const bool lx = px == g.nx - 1, ly = py == g.ny - 1, lz = (DIM == 2) || z == g.nz - 1;
    const I ix = lx ? i : i + 1, iy = ly ? i : i + nx, iz = lz ? i : i + sxy;
    const T sigma = a.sigma;
    if (!carried) {
        u = ldx<T>(a.U, i);
        #pragma unroll
        for (int c = 0; c < DIM; ++c) v[c] = ldx<T>(a.V[c], i);
    }
    const T ux = ldx<T>(a.U, ix), uy = ldx<T>(a.U, iy);
    uz = DIM == 3 ? ldx<T>(a.U, iz) : (T)0;
    T vx[3], vy[3], po[3], qo[6];
    #pragma unroll
    for (int c = 0; c < DIM; ++c) {
        vx[c] = ldx<T>(a.V[c], ix);
        vy[c] = ldx<T>(a.V[c], iy);
        if (DIM == 3) vz[c] = ldx<T>(a.V[c], iz);
        po[c] = ldx<T>(a.P[c], i);
    }
    const int nq = DIM == 3 ? 6 : 3;
    #pragma unroll
    for (int c = 0; c < nq; ++c) qo[c] = ldx<T>(a.Q[c], i);

    // P (TGV_core.c:360-365): forward difference of u_bar minus v_bar
    p[0] = po[0] + sigma * (lx ? -v[0] : (ux - u) - v[0]);
    p[1] = po[1] + sigma * (ly ? -v[1] : (uy - u) - v[1]);
    T nrm;
    if (DIM == 3) {
        p[2] = po[2] + sigma * (lz ? -v[2] : (uz - u) - v[2]);
        nrm = dsqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]) / a.alpha1;
    } else {
        p[2] = 0;
        nrm = dsqrt(p[0] * p[0] + p[1] * p[1]) / a.alpha1;
    }
    if (nrm > (T)1) {
        p[0] /= nrm;
        p[1] /= nrm;
        if (DIM == 3) p[2] /= nrm;
    }
    // Q (TGV_core.c:398-421): d[a][b] = forward difference of V_a along axis b
    T d[3][3];
    #pragma unroll
    for (int c = 0; c < 3; ++c) d[c][0] = d[c][1] = d[c][2] = 0;
    #pragma unroll
    for (int c = 0; c < DIM; ++c) {
        d[c][0] = lx ? (T)0 : vx[c] - v[c];
        d[c][1] = ly ? (T)0 : vy[c] - v[c];
        if (DIM == 3) d[c][2] = lz ? (T)0 : vz[c] - v[c];
    }
    if (DIM == 3) {
        q[0] = qo[0] + sigma * d[0][0];
        q[1] = qo[1] + sigma * d[1][1];
        q[2] = qo[2] + sigma * d[2][2];
        q[3] = qo[3] + sigma * ((T)0.5 * (d[1][0] + d[0][1]));
        q[4] = qo[4] + sigma * ((T)0.5 * (d[2][0] + d[0][2]));
        q[5] = qo[5] + sigma * ((T)0.5 * (d[2][1] + d[1][2]));
        T m = dsqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + (T)2 * q[3] * q[3] + (T)2 * q[4] * q[4] +
                    (T)2 * q[5] * q[5]);
        m = m / a.alpha0;
        if (m > (T)1) {
            #pragma unroll
            for (int c = 0; c < 6; ++c) q[c] /= m;
        }
    } else {
        q[0] = qo[0] + sigma * d[0][0];
        q[1] = qo[1] + sigma * d[1][1];
        q[2] = qo[2] + sigma * ((T)0.5 * (d[1][0] + d[0][1]));
        q[3] = q[4] = q[5] = 0;
        T m = dsqrt(q[0] * q[0] + q[1] * q[1] + (T)2 * q[2] * q[2]);
        m = m / a.alpha0;
        if (m > (T)1) {
            q[0] /= m;
            q[1] /= m;
            q[2] /= m;
        }
    }
}

// adjoint difference: own value at index 0, minus the predecessor at the last
// index, plain backward difference in between (TGV_core.c:458-466)
template <typename T>
__device__ inline T bdiff(T own, T pred, int pos, int len) {
    if (pos == 0) return own;
    if (pos == len - 1) return -pred;
    return own - pred;
}

template <int DIM, int TY, typename T, typename S, typename I>
__device__ inline void tgv_body(const TgvArgs<T, S>& a, const TileCoord& tc) {
    const Geom& g = a.g;
    constexpr int NT = TX * TY, NH = TX + TY;
    const int lx = threadIdx.x, ly = threadIdx.y, tid = ly * TX + lx;
    const int x = tc.x0 + lx, y = tc.y0 + ly;
    const bool own = x < g.nx && y < g.ny;
    // -x halo column (TY points) and -y halo row (TX points).  The lanes doing
    // the halo pass alternate between the low and the high waves of the block
    // with the block index, so neighbouring workgroups on a CU load different
    // SIMDs with the extra pass.
    const int hbase = (tc.lin & 1) ? NT - NH : 0;
    const int h = tid - hbase;
    const bool hlane = h >= 0 && h < NH;
    int hx, hy, hlx, hly;
    if (h < TY) { hx = tc.x0 - 1; hy = tc.y0 + h; hlx = 0; hly = h + 1; }
    else { hx = tc.x0 + h - TY; hy = tc.y0 - 1; hlx = h - TY + 1; hly = 0; }
    const bool hown = hlane && hx >= 0 && hy >= 0 && hx < g.nx && hy < g.ny;
    // points outside the volume compute on a valid stand-in and store nothing
    const int ox = own ? x : 0, oy = own ? y : 0;
    const int hcx = hown ? hx : 0, hcy = hown ? hy : 0;

    __shared__ T L[L_N][TY + 1][TX + 1];
    __shared__ double red[8];
    const I sxy = (I)g.sxy, nx = (I)g.nx;

    // channels of plane z-1 at the own point: P3, Q33, Q13, Q23
    T zp3 = 0, zq33 = 0, zq13 = 0, zq23 = 0;
    // u_bar, v_bar of the own and the halo point carried from the plane below
    T cu = 0, cv[3] = {0, 0, 0}, hu = 0, hv[3] = {0, 0, 0};
    bool own_carried = false;
    if (DIM == 3 && tc.z0 > 0) {
        T p[3], q[6], u0, v0[3];
        tgv_dual_at<DIM, I>(a, ox, oy, tc.z0 - 1, false, u0, v0, cu, cv, p, q);
        zp3 = p[2]; zq33 = q[2]; zq13 = q[4]; zq23 = q[5];
        own_carried = true;
    }
    double sdd = 0.0, suu = 0.0;
    for (int z = tc.z0; z < tc.z1; ++z) {
        T p[3], q[6], nu = 0, nv[3] = {0, 0, 0};
        tgv_dual_at<DIM, I>(a, ox, oy, z, own_carried || z > tc.z0, cu, cv, nu, nv, p, q);
        {
            const int r = ly + 1, c = lx + 1;
            L[L_P1][r][c] = p[0];
            L[L_P2][r][c] = p[1];
            L[L_Q11][r][c] = q[0];
            L[L_Q22][r][c] = q[1];
            if (DIM == 3) {
                L[L_Q12][r][c] = q[3];
                L[L_Q13][r][c] = q[4];
                L[L_Q23][r][c] = q[5];
            } else {
                L[L_Q12][r][c] = q[2];
            }
        }
        T nhu = 0, nhv[3] = {0, 0, 0};
        if (hlane) {
            T hp[3], hq[6];
            tgv_dual_at<DIM, I>(a, hcx, hcy, z, z > tc.z0, hu, hv, nhu, nhv, hp, hq);
            L[L_P1][hly][hlx] = hp[0];
            L[L_P2][hly][hlx] = hp[1];
            L[L_Q11][hly][hlx] = hq[0];
            L[L_Q22][hly][hlx] = hq[1];
            if (DIM == 3) {
                L[L_Q12][hly][hlx] = hq[3];
                L[L_Q13][hly][hlx] = hq[4];
                L[L_Q23][hly][hlx] = hq[5];
            } else {
                L[L_Q12][hly][hlx] = hq[2];
            }
        }
        __syncthreads();
        if (own) {
            const I i = (I)z * sxy + (I)y * nx + (I)x;
            // x-predecessor at [ly+1][lx], y-predecessor at [ly][lx+1]
            T dv = bdiff(p[0], L[L_P1][ly + 1][lx], x, g.nx) + bdiff(p[1], L[L_P2][ly][lx + 1], y, g.ny);
            if (DIM == 3) dv = dv + bdiff(p[2], zp3, z, g.nz);
            const T ub = cu;
            const T un = (a.lambda * (ub + a.tau * dv) + a.tau * ldx<T>(a.f, i)) / (a.lambda + a.tau);
            const T ubn = (T)2 * un - ub;
            if (a.check) {
                const T df = ubn - ub;
                sdd += (double)df * (double)df;
                suu += (double)ubn * (double)ubn;
            }
            stnt<T>(a.Uout, i, ubn);
            T dq[3];
            if (DIM == 3) {
                dq[0] = bdiff(q[0], L[L_Q11][ly + 1][lx], x, g.nx) + bdiff(q[3], L[L_Q12][ly][lx + 1], y, g.ny) +
                        bdiff(q[4], zq13, z, g.nz);
                dq[1] = bdiff(q[3], L[L_Q12][ly + 1][lx], x, g.nx) + bdiff(q[1], L[L_Q22][ly][lx + 1], y, g.ny) +
                        bdiff(q[5], zq23, z, g.nz);
                dq[2] = bdiff(q[4], L[L_Q13][ly + 1][lx], x, g.nx) + bdiff(q[5], L[L_Q23][ly][lx + 1], y, g.ny) +
                        bdiff(q[2], zq33, z, g.nz);
            } else {
                dq[0] = bdiff(q[0], L[L_Q11][ly + 1][lx], x, g.nx) + bdiff(q[2], L[L_Q12][ly][lx + 1], y, g.ny);
                dq[1] = bdiff(q[2], L[L_Q12][ly + 1][lx], x, g.nx) + bdiff(q[1], L[L_Q22][ly][lx + 1], y, g.ny);
                dq[2] = 0;
            }
            #pragma unroll
            for (int c = 0; c < DIM; ++c) {
                const T vb = cv[c];
                const T vn = vb + a.tau * (p[c] + dq[c]);
                stnt<T>(a.Vout[c], i, (T)2 * vn - vb);
                stnt<T>(a.Pout[c], i, p[c]);
            }
            const int nq = DIM == 3 ? 6 : 3;
            #pragma unroll
            for (int c = 0; c < nq; ++c) stnt<T>(a.Qout[c], i, q[c]);
        }
        __syncthreads();
        zp3 = p[2]; zq33 = q[2]; zq13 = q[4]; zq23 = q[5];
        cu = nu; hu = nhu;
        #pragma unroll
        for (int c = 0; c < 3; ++c) { cv[c] = nv[c]; hv[c] = nhv[c]; }
    }
    if (a.check) {
        block_sum2(sdd, suu, red);
        if (tid == 0) {
            a.partials[2 * tc.lin] = sdd;
            a.partials[2 * tc.lin + 1] = suu;
        }
    }
}

// One iteration on one volume: the workgroup's tile comes from its block index.
template <int DIM, int TY, typename T, typename S, typename I>
__global__ __launch_bounds__(TX * TY) void tgv_iter_kernel(TgvArgs<T, S> a) {
    if (solve_stopped(a.ctrl)) return;
    tgv_body<DIM, TY, T, S, I>(a, tile_of_block(a.g, TY));
}

}  // namespace cilreg
