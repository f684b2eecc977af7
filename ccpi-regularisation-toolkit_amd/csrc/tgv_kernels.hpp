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

// The dual step at one point (TGV_core.c:351-446): P and Q of the next
// iterate from the old P, Q there, u_bar and v_bar there (u, v) and at the
// +x / +y / +z neighbours, and whether the point is the last along each axis
// (the forward difference is zero there).  Shared by every TGV kernel.
template <int DIM, typename T>
__device__ inline void tgv_dual_math(T sigma, T alpha1, T alpha0, bool lx, bool ly, bool lz, T u, const T v[3], T ux,
                                     T uy, T uz, const T vx[3], const T vy[3], const T vz[3], const T po[3],
                                     const T qo[6], T p[3], T q[6]) {
    // P (TGV_core.c:360-365): forward difference of u_bar minus v_bar
    p[0] = po[0] + sigma * (lx ? -v[0] : (ux - u) - v[0]);
    p[1] = po[1] + sigma * (ly ? -v[1] : (uy - u) - v[1]);
    T nrm;
    if (DIM == 3) {
        p[2] = po[2] + sigma * (lz ? -v[2] : (uz - u) - v[2]);
        nrm = dsqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]) / alpha1;
    } else {
        p[2] = 0;
        nrm = dsqrt(p[0] * p[0] + p[1] * p[1]) / alpha1;
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
        m = m / alpha0;
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
        m = m / alpha0;
        if (m > (T)1) {
            q[0] /= m;
            q[1] /= m;
            q[2] /= m;
        }
    }
}

// State in FP32 or 16-bit storage keeps two components of a point side by
// side (pair fields, the first component at the even element).  3D: (V1, V2)
// in V[0], (V3, P1) in V[2], (P2, P3) in P[1], (Q11, Q22) in Q[0], (Q33, Q12)
// in Q[2], (Q13, Q23) in Q[4]; 2D: (V1, V2) in V[0], (P1, P2) in P[0],
// (Q11, Q22) in Q[0] and Q12 alone in Q[2].  The other component pointers
// are unused.  A 3D point then issues 11 loads of a pair instead of 18
// single loads and 6 stores instead of 12 (192x512x512: FP32 0.87 -> 0.79
// ms, BF16 0.71 -> 0.62 ms per iteration).  FP64 state keeps one field per
// component.
template <int DIM, typename S>
constexpr bool tgv_pairs() { return (DIM == 2 || DIM == 3) && (sizeof(S) == 2 || std::is_same<S, float>::value); }

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
    if constexpr (tgv_pairs<DIM, S>() && DIM == 2) {
        if (!carried) {
            u = ldx<T>(a.U, i);
            ldpair<T>(a.V[0], i, v[0], v[1]);
        }
        const T ux = ldx<T>(a.U, ix), uy = ldx<T>(a.U, iy);
        uz = (T)0;
        T vx[3] = {0, 0, 0}, vy[3] = {0, 0, 0}, po[3] = {0, 0, 0}, qo[6] = {0, 0, 0, 0, 0, 0};
        ldpair<T>(a.V[0], ix, vx[0], vx[1]);
        ldpair<T>(a.V[0], iy, vy[0], vy[1]);
        ldpair<T>(a.P[0], i, po[0], po[1]);
        ldpair<T>(a.Q[0], i, qo[0], qo[1]);
        qo[2] = ldx<T>(a.Q[2], i);
        tgv_dual_math<DIM, T>(a.sigma, a.alpha1, a.alpha0, lx, ly, lz, u, v, ux, uy, uz, vx, vy, vz, po, qo, p, q);
        return;
    }
    if constexpr (tgv_pairs<DIM, S>() && DIM == 3) {
        T v3i, p1i, unused;
        ldpair<T>(a.V[2], i, v3i, p1i);
        if (!carried) {
            u = ldx<T>(a.U, i);
            ldpair<T>(a.V[0], i, v[0], v[1]);
            v[2] = v3i;
        }
        const T ux = ldx<T>(a.U, ix), uy = ldx<T>(a.U, iy);
        uz = ldx<T>(a.U, iz);
        T vx[3], vy[3], po[3], qo[6];
        ldpair<T>(a.V[0], ix, vx[0], vx[1]);
        ldpair<T>(a.V[2], ix, vx[2], unused);
        ldpair<T>(a.V[0], iy, vy[0], vy[1]);
        ldpair<T>(a.V[2], iy, vy[2], unused);
        ldpair<T>(a.V[0], iz, vz[0], vz[1]);
        ldpair<T>(a.V[2], iz, vz[2], unused);
        po[0] = p1i;
        ldpair<T>(a.P[1], i, po[1], po[2]);
        ldpair<T>(a.Q[0], i, qo[0], qo[1]);
        ldpair<T>(a.Q[2], i, qo[2], qo[3]);
        ldpair<T>(a.Q[4], i, qo[4], qo[5]);
        tgv_dual_math<DIM, T>(a.sigma, a.alpha1, a.alpha0, lx, ly, lz, u, v, ux, uy, uz, vx, vy, vz, po, qo, p, q);
        return;
    }
    if (!carried) {
        u = ldx<T>(a.U, i);
        #pragma unroll
        for (int c = 0; c < DIM; ++c) v[c] = ldx<T>(a.V[c], i);
    }
    const T ux = ldx<T>(a.U, ix), uy = ldx<T>(a.U, iy);
    uz = DIM == 3 ? ldx<T>(a.U, iz) : (T)0;
    T vx[3] = {0, 0, 0}, vy[3] = {0, 0, 0}, po[3] = {0, 0, 0}, qo[6] = {0, 0, 0, 0, 0, 0};
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
    tgv_dual_math<DIM, T>(a.sigma, a.alpha1, a.alpha0, lx, ly, lz, u, v, ux, uy, uz, vx, vy, vz, po, qo, p, q);
}

// Stores the state of one point (u_bar, v_bar, P, Q) at element i.
template <int DIM, typename I, typename T, typename S, class A>
__device__ inline void tgv_store_state(const A& a, I i, T ubn, const T vbn[3], const T p[3], const T q[6]) {
    stnt<T>(a.Uout, i, ubn);
    if constexpr (tgv_pairs<DIM, S>() && DIM == 2) {
        stpair_nt<T>(a.Vout[0], i, vbn[0], vbn[1]);
        stpair_nt<T>(a.Pout[0], i, p[0], p[1]);
        stpair_nt<T>(a.Qout[0], i, q[0], q[1]);
        stnt<T>(a.Qout[2], i, q[2]);
    } else if constexpr (tgv_pairs<DIM, S>()) {
        stpair_nt<T>(a.Vout[0], i, vbn[0], vbn[1]);
        stpair_nt<T>(a.Vout[2], i, vbn[2], p[0]);
        stpair_nt<T>(a.Pout[1], i, p[1], p[2]);
        stpair_nt<T>(a.Qout[0], i, q[0], q[1]);
        stpair_nt<T>(a.Qout[2], i, q[2], q[3]);
        stpair_nt<T>(a.Qout[4], i, q[4], q[5]);
    } else {
        #pragma unroll
        for (int c = 0; c < DIM; ++c) {
            stnt<T>(a.Vout[c], i, vbn[c]);
            stnt<T>(a.Pout[c], i, p[c]);
        }
        const int nq = DIM == 3 ? 6 : 3;
        #pragma unroll
        for (int c = 0; c < nq; ++c) stnt<T>(a.Qout[c], i, q[c]);
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

// The primal step at one point (TGV_core.c:448-530): u_bar and v_bar of the
// next iterate from the new P, Q there (p, q), their -x / -y / -z
// neighbours (P1, Q11, Q12, Q13 at x-1; P2, Q12, Q22, Q23 at y-1; P3, Q13,
// Q23, Q33 at z-1), the old u_bar, v_bar (ub, vb) and f.  Shared by every
// TGV kernel.
template <int DIM, typename T>
__device__ inline void tgv_primal_math(T lambda, T tau, int x, int y, int z, int nx, int ny, int nz, const T p[3],
                                       const T q[6], const T xm[4], const T ym[4], const T zm[4], T ub,
                                       const T vb[3], T f, T& ubn, T vbn[3]) {
    // xm = {P1, Q11, Q12, Q13}(x-1), ym = {P2, Q12, Q22, Q23}(y-1), zm = {P3, Q13, Q23, Q33}(z-1)
    T dv = bdiff(p[0], xm[0], x, nx) + bdiff(p[1], ym[0], y, ny);
    if (DIM == 3) dv = dv + bdiff(p[2], zm[0], z, nz);
    const T un = (lambda * (ub + tau * dv) + tau * f) / (lambda + tau);
    ubn = (T)2 * un - ub;
    T dq[3];
    if (DIM == 3) {
        dq[0] = bdiff(q[0], xm[1], x, nx) + bdiff(q[3], ym[1], y, ny) + bdiff(q[4], zm[1], z, nz);
        dq[1] = bdiff(q[3], xm[2], x, nx) + bdiff(q[1], ym[2], y, ny) + bdiff(q[5], zm[2], z, nz);
        dq[2] = bdiff(q[4], xm[3], x, nx) + bdiff(q[5], ym[3], y, ny) + bdiff(q[2], zm[3], z, nz);
    } else {
        dq[0] = bdiff(q[0], xm[1], x, nx) + bdiff(q[2], ym[1], y, ny);
        dq[1] = bdiff(q[2], xm[2], x, nx) + bdiff(q[1], ym[2], y, ny);
        dq[2] = 0;
    }
    #pragma unroll
    for (int c = 0; c < DIM; ++c) {
        const T vn = vb[c] + tau * (p[c] + dq[c]);
        vbn[c] = (T)2 * vn - vb[c];
    }
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
    __shared__ double red[kRedSlots];
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
            T xm[4], ym[4], zm[4];
            xm[0] = L[L_P1][ly + 1][lx];
            ym[0] = L[L_P2][ly][lx + 1];
            zm[0] = zp3;
            if (DIM == 3) {
                xm[1] = L[L_Q11][ly + 1][lx]; xm[2] = L[L_Q12][ly + 1][lx]; xm[3] = L[L_Q13][ly + 1][lx];
                ym[1] = L[L_Q12][ly][lx + 1]; ym[2] = L[L_Q22][ly][lx + 1]; ym[3] = L[L_Q23][ly][lx + 1];
                zm[1] = zq13; zm[2] = zq23; zm[3] = zq33;
            } else {
                xm[1] = L[L_Q11][ly + 1][lx]; xm[2] = L[L_Q12][ly + 1][lx]; xm[3] = 0;
                ym[1] = L[L_Q12][ly][lx + 1]; ym[2] = L[L_Q22][ly][lx + 1]; ym[3] = 0;
                zm[1] = zm[2] = zm[3] = 0;
            }
            const T ub = cu;
            T ubn, vbn[3] = {0, 0, 0};
            tgv_primal_math<DIM, T>(a.lambda, a.tau, x, y, z, g.nx, g.ny, g.nz, p, q, xm, ym, zm, ub, cv,
                                    ldx<T>(a.f, i), ubn, vbn);
            if (a.check) {
                const T df = ubn - ub;
                sdd += (double)df * (double)df;
                suu += (double)ubn * (double)ubn;
            }
            tgv_store_state<DIM, I, T, S>(a, i, ubn, vbn, p, q);
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

// ---------------------------------------------------------------------------
// Two iterations per pass (temporal blocking of depth 2, 3D)
// ---------------------------------------------------------------------------
// The TGV state (u_bar, v_bar, P, Q) does not involve older iterates, so a
// pass that advances k -> k+2 reads the 14 input fields once and writes the
// 13 state fields once: 108 B per two voxel-iterations instead of 216.  Each
// value is computed by tgv_dual_math / tgv_primal_math with the inputs the
// one-iteration kernel would use (intermediate fields rounded through the
// storage type, as they would be stored and read back), so the results are
// bit-identical.
//
// The workgroup owns the 64 x TY tile T and evaluates iteration k on
// B = [x0-2, x0+64] x [y0-2, y0+TY] (67 x (TY+3) points; slot 0 of every
// thread is its own point of T, slot 1 of 3*67 + 3*TY ring lanes a point of
// B \ T).  Validity shrinks stage by stage: the duals of iteration k on B,
// its primal from x0-1 / y0-1 on, the duals of k+1 on [x0-1, x0+63] x
// [y0-1, y0+TY-1], the primal of k+1 on T.
//
// z pipeline for plane zi of the march over [z0 - 2, z1]:
//   S1  P, Q of k+1 at zi        (u_bar, v_bar of k at zi and zi+1)
//   S2  u_bar, v_bar of k+1 at zi (P, Q of k+1 at zi, x-1, y-1 and zi-1)
//   S3  P, Q of k+2 at zi-1      (u_bar, v_bar of k+1 at zi-1, +x, +y, zi)
//   S4  u_bar, v_bar of k+2 at zi-1 on T; all 13 state fields stored.
// In the one-iteration mode (single) S2 stores the state of k+1 on T, as the
// one-iteration kernel does, and S3 / S4 do nothing.
template <typename T, typename S>
struct Tgv2Args {
    const T* f;
    const T* U;
    const S* V[3];
    const S* P[3];
    const S* Q[6];
    T* Uout;
    S* Vout[3];
    S* Pout[3];
    S* Qout[6];
    T* Ualt;            // u_bar after iteration k (written when check0 and non-null: the output of a stop at k)
    T lambda, tau, sigma, alpha1, alpha0;
    int check0, check1;
    int single;         // advance one iteration (k -> k+1) only: S2 stores the state
    double* partials0;
    double* partials1;
    const Ctrl* ctrl;
    Geom g;
};

template <int TY>
struct Tgv2Shape {
    static constexpr int BW = TX + 3, BH = TY + 3;
    static constexpr int NR = 3 * BW + 3 * TY;
    static constexpr int NT = TX * TY;
};

// Ring point h (0 <= h < 3 (TX+3) + 3 TY) of the region, in region coordinates.
template <int TY>
__device__ inline void tgv2_ring_point(int h, int& rx, int& ry) {
    constexpr int BW = TX + 3;
    if (h < 3 * BW) {
        const int row = h / BW;
        rx = h - row * BW;
        ry = row < 2 ? row : TY + 2;           // rows y0-2, y0-1, y0+TY
    } else {
        const int k = h - 3 * BW, col = k / TY;
        rx = col < 2 ? col : TX + 2;           // columns x0-2, x0-1, x0+TX
        ry = k - col * TY + 2;
    }
}

// RW = 0: TX x TY threads, each the own tile point (slot 0) and NR of them
// also a ring point (slot 1, on the low or high waves by block index).
// RW > 0: TX x TY + 64 RW threads, the first TX x TY the tile, the RW waves
// after them the ring, one point per thread.  SINGLE: whether the
// one-iteration mode is compiled in (the grouped kernel's odd last
// iterations; the single-volume solver runs those on tgv_body).
template <int TY, typename T, typename S, typename I, int RW = 0, bool SINGLE = true>
__device__ inline void tgv2_body(const Tgv2Args<T, S>& a, const TileCoord& tc) {
    using Sh = Tgv2Shape<TY>;
    constexpr int BW = Sh::BW, BH = Sh::BH, NR = Sh::NR, NT = Sh::NT;
    constexpr int NS = RW ? 1 : 2;   // slots per thread
    static_assert(RW == 0 || 64 * RW >= NR, "ring waves must hold the ring");
    const Geom& g = a.g;
    const int lx = threadIdx.x, ly = threadIdx.y, tid = ly * TX + lx;

    int bx[2] = {0, 0}, by[2] = {0, 0};
    int nslot;
    bool tile0;   // slot 0 is a tile point
    if (RW) {
        tile0 = tid < NT;
        const int h = tid - NT;
        const bool ring = !tile0 && h < NR;
        int rx = 0, ry = 0;
        if (ring) tgv2_ring_point<TY>(h, rx, ry);
        bx[0] = tile0 ? lx + 2 : rx;
        by[0] = tile0 ? ly + 2 : ry;
        nslot = (tile0 || ring) ? 1 : 0;
    } else {
        tile0 = true;
        bx[0] = lx + 2;
        by[0] = ly + 2;
        const int hbase = (tc.lin & 1) ? NT - NR : 0;
        const int h = tid - hbase;
        const bool ring = h >= 0 && h < NR;
        int rx = 0, ry = 0;
        if (ring) tgv2_ring_point<TY>(h, rx, ry);
        bx[1] = rx;
        by[1] = ry;
        nslot = ring ? 2 : 1;
    }
    int X[2], Y[2], cx[2], cy[2];
    #pragma unroll
    for (int s = 0; s < 2; ++s) {
        X[s] = tc.x0 + bx[s] - 2;
        Y[s] = tc.y0 + by[s] - 2;
        const bool in = X[s] >= 0 && Y[s] >= 0 && X[s] < g.nx && Y[s] < g.ny;
        cx[s] = in ? X[s] : 0;
        cy[s] = in ? Y[s] : 0;
    }
    const bool own = tile0 && X[0] < g.nx && Y[0] < g.ny;

    __shared__ T PQ[L_N][BH + 1][BW + 1];      // P, Q channels read at x-1 / y-1
    __shared__ T UV[2][4][BH + 1][BW + 1];     // u_bar, v_bar of k+1, planes by parity
    __shared__ double red[kRedSlots];

    const I sxy = (I)g.sxy, nx = (I)g.nx;
    const int nz = g.nz, z0 = tc.z0, z1 = tc.z1;
    const T sigma = a.sigma, tau = a.tau, lam = a.lambda;

    // views of the input state for tgv_dual_at
    TgvArgs<T, S> in;
    in.f = a.f;
    in.U = a.U;
    #pragma unroll
    for (int c = 0; c < 3; ++c) { in.V[c] = a.V[c]; in.P[c] = a.P[c]; }
    #pragma unroll
    for (int c = 0; c < 6; ++c) in.Q[c] = a.Q[c];
    in.sigma = sigma;
    in.alpha1 = a.alpha1;
    in.alpha0 = a.alpha0;
    in.g = g;

    T cu[2] = {0, 0}, cv[2][3] = {};     // u_bar, v_bar of k at the current plane (carried)
    T p1v[2][3] = {}, q1v[2][6] = {};    // P, Q of k+1 at the previous plane
    T zp2[4] = {0, 0, 0, 0};             // P3, Q13, Q23, Q33 of k+2 at the previous plane (slot 0)
    T f0 = 0;                            // f at the previous plane (slot 0)
    double sdd0 = 0.0, suu0 = 0.0, sdd1 = 0.0, suu1 = 0.0;
    // With ring waves, values only tile threads use live in LDS slots of
    // their own (thirteen registers fewer under the 80-register cap): the
    // four residual sums, added to in the same order, and the slot-0 z
    // pipeline (zp2, f0).
    constexpr int NL = RW ? NT : 1;
    __shared__ double acc[4][NL];
    __shared__ T zl[5][NL];
    const int lt = (RW && tile0) ? tid : 0;
    if (RW && tile0) {
        #pragma unroll
        for (int c = 0; c < 4; ++c) acc[c][lt] = 0.0;
        #pragma unroll
        for (int c = 0; c < 5; ++c) zl[c][lt] = (T)0;
    }
    auto add_acc = [&](int c0, double& r0, double& r1, double v0, double v1) {
        if (RW) {
            acc[c0][lt] += v0;
            acc[c0 + 1][lt] += v1;
        } else {
            r0 += v0;
            r1 += v1;
        }
    };
    const int zs = z0 >= 2 ? z0 - 2 : 0;
    for (int zi = zs; zi <= z1; ++zi) {
        const bool s12 = zi < nz;
        T p1[2][3] = {}, q1[2][6] = {}, nu[2] = {0, 0}, nv[2][3] = {};
        // ---- S1: P, Q of k+1 at zi
        if (s12) {
            #pragma unroll
            for (int s = 0; s < NS; ++s) {
                if (s < nslot) {
                    tgv_dual_at<3, I>(in, cx[s], cy[s], zi, zi > zs, cu[s], cv[s], nu[s], nv[s], p1[s], q1[s]);
                    PQ[L_P1][by[s]][bx[s]] = p1[s][0];
                    PQ[L_P2][by[s]][bx[s]] = p1[s][1];
                    PQ[L_Q11][by[s]][bx[s]] = q1[s][0];
                    PQ[L_Q22][by[s]][bx[s]] = q1[s][1];
                    PQ[L_Q12][by[s]][bx[s]] = q1[s][3];
                    PQ[L_Q13][by[s]][bx[s]] = q1[s][4];
                    PQ[L_Q23][by[s]][bx[s]] = q1[s][5];
                }
            }
        }
        __syncthreads();
        // ---- S2: u_bar, v_bar of k+1 at zi
        T fcur = 0;
        if (s12 && zi + 1 >= z0) {
            #pragma unroll
            for (int s = 0; s < NS; ++s) {
                if (s < nslot) {
                    const int xb = bx[s] > 0 ? bx[s] - 1 : 0, yb = by[s] > 0 ? by[s] - 1 : 0;
                    T xm[4], ym[4], zm[4];
                    xm[0] = PQ[L_P1][by[s]][xb];
                    xm[1] = PQ[L_Q11][by[s]][xb];
                    xm[2] = PQ[L_Q12][by[s]][xb];
                    xm[3] = PQ[L_Q13][by[s]][xb];
                    ym[0] = PQ[L_P2][yb][bx[s]];
                    ym[1] = PQ[L_Q12][yb][bx[s]];
                    ym[2] = PQ[L_Q22][yb][bx[s]];
                    ym[3] = PQ[L_Q23][yb][bx[s]];
                    zm[0] = p1v[s][2];
                    zm[1] = q1v[s][4];
                    zm[2] = q1v[s][5];
                    zm[3] = q1v[s][2];
                    const I i = (I)zi * sxy + (I)cy[s] * nx + (I)cx[s];
                    const T fv = ldx<T>(a.f, i);
                    if (s == 0) fcur = fv;
                    T ubn, vbn[3];
                    tgv_primal_math<3, T>(lam, tau, X[s], Y[s], zi, g.nx, g.ny, nz, p1[s], q1[s], xm, ym, zm,
                                          cu[s], cv[s], fv, ubn, vbn);
                    if (s == 0 && own && zi >= z0 && zi < z1) {
                        if (SINGLE && a.single)   // the state of k+1, as tgv_body stores it
                            tgv_store_state<3, I, T, S>(a, (I)zi * sxy + (I)Y[0] * nx + (I)X[0], ubn, vbn, p1[0],
                                                        q1[0]);
                        if (a.check0) {
                            const T df = ubn - cu[0];
                            add_acc(0, sdd0, suu0, (double)df * (double)df, (double)ubn * (double)ubn);
                            if (a.Ualt) stnt<T>(a.Ualt, (I)zi * sxy + (I)Y[0] * nx + (I)X[0], ubn);
                        }
                    }
                    UV[zi & 1][0][by[s]][bx[s]] = ubn;
                    #pragma unroll
                    for (int c = 0; c < 3; ++c) UV[zi & 1][1 + c][by[s]][bx[s]] = through_storage<T, S>(vbn[c]);
                }
            }
        }
        __syncthreads();
        // ---- S3: P, Q of k+2 at zc
        const int zc = zi - 1;
        T p2[2][3] = {}, q2[2][6] = {};
        const bool s3 = !(SINGLE && a.single) && zc >= (z0 >= 1 ? z0 - 1 : 0) && zc <= z1 - 1;
        if (s3) {
            #pragma unroll
            for (int s = 0; s < NS; ++s) {
                if (s < nslot) {
                    const int b = zc & 1;
                    const int xp = bx[s] + 1, yp = by[s] + 1;
                    const T u = UV[b][0][by[s]][bx[s]];
                    const T ux = UV[b][0][by[s]][xp], uy = UV[b][0][yp][bx[s]], uz = UV[b ^ 1][0][by[s]][bx[s]];
                    T v[3], vx[3], vy[3], vz[3], po[3], qo[6];
                    #pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        v[c] = UV[b][1 + c][by[s]][bx[s]];
                        vx[c] = UV[b][1 + c][by[s]][xp];
                        vy[c] = UV[b][1 + c][yp][bx[s]];
                        vz[c] = UV[b ^ 1][1 + c][by[s]][bx[s]];
                        po[c] = through_storage<T, S>(p1v[s][c]);
                    }
                    #pragma unroll
                    for (int c = 0; c < 6; ++c) qo[c] = through_storage<T, S>(q1v[s][c]);
                    const bool bl_x = X[s] == g.nx - 1, bl_y = Y[s] == g.ny - 1, bl_z = zc == nz - 1;
                    tgv_dual_math<3, T>(sigma, a.alpha1, a.alpha0, bl_x, bl_y, bl_z, u, v, ux, uy, uz, vx, vy, vz,
                                        po, qo, p2[s], q2[s]);
                    PQ[L_P1][by[s]][bx[s]] = p2[s][0];
                    PQ[L_P2][by[s]][bx[s]] = p2[s][1];
                    PQ[L_Q11][by[s]][bx[s]] = q2[s][0];
                    PQ[L_Q22][by[s]][bx[s]] = q2[s][1];
                    PQ[L_Q12][by[s]][bx[s]] = q2[s][3];
                    PQ[L_Q13][by[s]][bx[s]] = q2[s][4];
                    PQ[L_Q23][by[s]][bx[s]] = q2[s][5];
                }
            }
        }
        __syncthreads();
        // ---- S4: u_bar, v_bar of k+2 at zc on T; store the state
        if (s3 && zc >= z0 && own) {
            const int b = zc & 1;
            const int xb = bx[0] - 1, yb = by[0] - 1;
            T xm[4], ym[4];
            xm[0] = PQ[L_P1][by[0]][xb];
            xm[1] = PQ[L_Q11][by[0]][xb];
            xm[2] = PQ[L_Q12][by[0]][xb];
            xm[3] = PQ[L_Q13][by[0]][xb];
            ym[0] = PQ[L_P2][yb][bx[0]];
            ym[1] = PQ[L_Q12][yb][bx[0]];
            ym[2] = PQ[L_Q22][yb][bx[0]];
            ym[3] = PQ[L_Q23][yb][bx[0]];
            const T ub = UV[b][0][by[0]][bx[0]];
            T vb[3];
            #pragma unroll
            for (int c = 0; c < 3; ++c) vb[c] = UV[b][1 + c][by[0]][bx[0]];
            T ubn, vbn[3];
            T zpv[4], f0v = f0;
            #pragma unroll
            for (int c = 0; c < 4; ++c) zpv[c] = zp2[c];
            if (RW) {
                #pragma unroll
                for (int c = 0; c < 4; ++c) zpv[c] = zl[c][lt];
                f0v = zl[4][lt];
            }
            tgv_primal_math<3, T>(lam, tau, X[0], Y[0], zc, g.nx, g.ny, nz, p2[0], q2[0], xm, ym, zpv, ub, vb, f0v,
                                  ubn, vbn);
            if (a.check1) {
                const T df = ubn - ub;
                add_acc(2, sdd1, suu1, (double)df * (double)df, (double)ubn * (double)ubn);
            }
            const I io = (I)zc * sxy + (I)Y[0] * nx + (I)X[0];
            tgv_store_state<3, I, T, S>(a, io, ubn, vbn, p2[0], q2[0]);
        }
        __syncthreads();
        // ---- shift the pipeline
        if (s3) {
            if (RW) {
                if (tile0) {
                    zl[0][lt] = p2[0][2];
                    zl[1][lt] = q2[0][4];
                    zl[2][lt] = q2[0][5];
                    zl[3][lt] = q2[0][2];
                }
            } else {
                zp2[0] = p2[0][2];
                zp2[1] = q2[0][4];
                zp2[2] = q2[0][5];
                zp2[3] = q2[0][2];
            }
        }
        if (s12) {
            #pragma unroll
            for (int s = 0; s < NS; ++s) {
                cu[s] = nu[s];
                #pragma unroll
                for (int c = 0; c < 3; ++c) { cv[s][c] = nv[s][c]; p1v[s][c] = p1[s][c]; }
                #pragma unroll
                for (int c = 0; c < 6; ++c) q1v[s][c] = q1[s][c];
            }
            if (RW) {
                if (tile0) zl[4][lt] = fcur;
            } else {
                f0 = fcur;
            }
        }
    }
    if (RW && tile0) {
        sdd0 = acc[0][lt];
        suu0 = acc[1][lt];
        sdd1 = acc[2][lt];
        suu1 = acc[3][lt];
    }
    if (a.check0) {
        block_sum2(sdd0, suu0, red);
        if (tid == 0) {
            a.partials0[2 * tc.lin] = sdd0;
            a.partials0[2 * tc.lin + 1] = suu0;
        }
        __syncthreads();
    }
    if (a.check1) {
        block_sum2(sdd1, suu1, red);
        if (tid == 0) {
            a.partials1[2 * tc.lin] = sdd1;
            a.partials1[2 * tc.lin + 1] = suu1;
        }
    }
}

// FP32 arithmetic, 32-bit offsets.  On 64 x 8 tiles the ring runs on four
// waves of its own (768 threads, one point per thread) and registers are
// capped at 80, so two workgroups fit a CU with six waves per SIMD; on 64 x 4
// tiles each thread also carries a ring point and the cap is 128 (two
// 256-thread workgroups per SIMD pair).
template <int TY>
constexpr int tgv2_ring_waves() { return TY == 8 ? 4 : 0; }
template <int TY>
constexpr int tgv2_threads() { return TX * TY + 64 * tgv2_ring_waves<TY>(); }

template <int TY, typename T, typename S>
__global__ __launch_bounds__(tgv2_threads<TY>()) __attribute__((amdgpu_waves_per_eu(tgv2_ring_waves<TY>() ? 6 : 4)))
void tgv2_iter_kernel(Tgv2Args<T, S> a) {
    if (solve_stopped(a.ctrl)) return;
    tgv2_body<TY, T, S, uint32_t, tgv2_ring_waves<TY>(), false>(a, tile_of_block(a.g, TY));
}

// No register cap: 64-bit offsets (fields of 2 GiB and more) and FP64, whose
// bodies do not fit 128 registers without spilling.
template <int TY, typename T, typename S, typename I>
__global__ __launch_bounds__(TX * TY) void tgv2_iter_kernel_uncapped(Tgv2Args<T, S> a) {
    if (solve_stopped(a.ctrl)) return;
    tgv2_body<TY, T, S, I, 0, false>(a, tile_of_block(a.g, TY));
}

}  // namespace cilreg
