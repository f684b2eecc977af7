// common.hpp -- shared device/host machinery of the MI355X regulariser library.
//
// Layout of every field in HBM: one dense array per channel, index
// (nx*ny)*z + nx*y + x with x fastest (the reference's layout,
// src/Core/regularisers_CPU/FGP_TV_core.c:228).  Kernels march a 64 x TY
// (x, y) tile of threads through a chunk of z planes ("2.5D" marching): each
// thread owns one (x, y) column, values a plane needs from its z-predecessor
// stay in registers, and the per-plane intermediate results that neighbouring
// threads need (the updated dual field, the FGP objective, the ROF fluxes) go
// through LDS.  Inputs are read straight from HBM with neighbour re-reads
// served by L1/L2; blocks are remapped so each XCD walks a contiguous run of
// tiles (its L2 then holds the halos of its own tiles).
//
// Arithmetic is kept operation-for-operation identical to the reference C-OMP
// code (the library is compiled with -ffp-contract=off, IEEE division and
// sqrt), including its double promotions, so FP32 results match the CPU
// oracle bit for bit; only the stopping-rule norms differ (FP64 tree sums here,
// a serial FP32 sum in the reference).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <type_traits>

#pragma clang fp contract(off)

namespace cilreg {

constexpr int TX = 64;          // tile width in x = one wavefront
constexpr int NXCD = 8;         // MI355X: 8 XCDs, workgroups dealt round-robin

// ---------------------------------------------------------------------------
// storage types for dual / auxiliary fields
// ---------------------------------------------------------------------------
struct bf16_t { uint16_t bits; };

template <typename S> struct Store;
template <> struct Store<float> {
    __device__ static inline float ld(const float* p, size_t i) { return p[i]; }
    __device__ static inline void st(float* p, size_t i, float v) { p[i] = v; }
};
template <> struct Store<double> {
    __device__ static inline double ld(const double* p, size_t i) { return p[i]; }
    __device__ static inline void st(double* p, size_t i, double v) { p[i] = v; }
};
template <> struct Store<_Float16> {
    __device__ static inline float ld(const _Float16* p, size_t i) { return (float)p[i]; }
    __device__ static inline void st(_Float16* p, size_t i, float v) { p[i] = (_Float16)v; }  // RNE
};
template <> struct Store<bf16_t> {
    __device__ static inline float ld(const bf16_t* p, size_t i) {
        return __uint_as_float(((uint32_t)p[i].bits) << 16);
    }
    // round to nearest even (gfx950 converts in hardware, v_cvt_pk_bf16_f32)
    __device__ static inline uint16_t bits_of(float v) { return __builtin_bit_cast(uint16_t, (__bf16)v); }
    __device__ static inline void st(bf16_t* p, size_t i, float v) { p[i].bits = bits_of(v); }
};

template <typename T, typename S>
__device__ inline T ldT(const S* p, size_t i) { return (T)Store<S>::ld(p, i); }
template <typename T, typename S>
__device__ inline void stT(S* p, size_t i, T v) { Store<S>::st(p, i, v); }

// A value as it reads back after a store in S (identity for S = T): what a
// later kernel sees of a field this one stores, for passes that keep using
// the value themselves.
template <typename T, typename S>
__device__ inline T through_storage(T v) {
    if constexpr (std::is_same<S, bf16_t>::value) {
        return (T)__uint_as_float(((uint32_t)Store<bf16_t>::bits_of((float)v)) << 16);
    } else if constexpr (std::is_same<S, _Float16>::value) {
        return (T)(float)(_Float16)(float)v;
    } else {
        return v;
    }
}

// Element access with an index type I.  With I = uint32_t the byte offset is
// formed in 32 bits (the host guarantees it fits), so loads use the
// scalar-base + 32-bit vector-offset addressing mode instead of 64-bit
// address arithmetic per access.
template <typename I> struct IdxT;
template <> struct IdxT<uint32_t> {
    template <typename S> __device__ static inline const S* at(const S* p, uint32_t e) {
        return (const S*)((const char*)p + (uint32_t)(e * (uint32_t)sizeof(S)));
    }
    template <typename S> __device__ static inline S* at(S* p, uint32_t e) {
        return (S*)((char*)p + (uint32_t)(e * (uint32_t)sizeof(S)));
    }
};
template <> struct IdxT<uint64_t> {
    template <typename S> __device__ static inline const S* at(const S* p, uint64_t e) { return p + e; }
    template <typename S> __device__ static inline S* at(S* p, uint64_t e) { return p + e; }
};

// Every field lives in device global memory: the accesses go through
// global-address-space pointers, so they compile to global (not flat)
// memory instructions even where the pointer was read from memory (the
// grouped kernels' job descriptors) rather than passed as a kernel argument.
template <typename S>
__device__ inline const __attribute__((address_space(1))) S* gmem(const S* p) {
    return (const __attribute__((address_space(1))) S*)p;
}
template <typename S>
__device__ inline __attribute__((address_space(1))) S* gmem(S* p) {
    return (__attribute__((address_space(1))) S*)p;
}

template <typename T, typename I, typename S>
__device__ inline T ldx(const S* p, I e) {
    if constexpr (std::is_same<S, bf16_t>::value) {
        return (T)__uint_as_float(((uint32_t)gmem(IdxT<I>::at(p, e))->bits) << 16);
    } else {
        return (T)*gmem(IdxT<I>::at(p, e));
    }
}
template <typename T, typename I, typename S>
__device__ inline void stx(S* p, I e, T v) {
    if constexpr (std::is_same<S, bf16_t>::value) {
        gmem(IdxT<I>::at(p, e))->bits = Store<bf16_t>::bits_of((float)v);
    } else if constexpr (std::is_same<S, _Float16>::value) {
        *gmem(IdxT<I>::at(p, e)) = (_Float16)(float)v;
    } else {
        *gmem(IdxT<I>::at(p, e)) = (S)v;
    }
}
template <typename T, typename I, typename S>
__device__ inline void stnt(S* p, I e, T v) {
    if constexpr (std::is_same<S, bf16_t>::value) {
        __builtin_nontemporal_store(Store<bf16_t>::bits_of((float)v), &gmem(IdxT<I>::at(p, e))->bits);
    } else {
        __builtin_nontemporal_store((S)v, gmem(IdxT<I>::at(p, e)));
    }
}

// Pair fields: two values of one point side by side (elements 2e and 2e + 1),
// moved by one access of 2 sizeof(S) bytes.  The host keeps 2 n sizeof(S)
// within the 32-bit offset range when I = uint32_t.
template <typename S> struct PairV { typedef S v __attribute__((ext_vector_type(2))); };
template <> struct PairV<bf16_t> { typedef uint32_t v; };

template <typename T, typename I, typename S>
__device__ inline void ldpair(const S* p, I e, T& lo, T& hi) {
    using V = typename PairV<S>::v;
    const V w = *gmem(IdxT<I>::at((const V*)p, e));
    if constexpr (std::is_same<S, bf16_t>::value) {
        lo = (T)__uint_as_float(w << 16);
        hi = (T)__uint_as_float(w & 0xffff0000u);
    } else {
        lo = (T)w.x;
        hi = (T)w.y;
    }
}
template <typename T, typename I, typename S>
__device__ inline void stpair_nt(S* p, I e, T lo, T hi) {
    using V = typename PairV<S>::v;
    V w;
    if constexpr (std::is_same<S, bf16_t>::value) {
        w = (uint32_t)Store<bf16_t>::bits_of((float)lo) | ((uint32_t)Store<bf16_t>::bits_of((float)hi) << 16);
    } else if constexpr (std::is_same<S, _Float16>::value) {
        w.x = (_Float16)(float)lo;
        w.y = (_Float16)(float)hi;
    } else {
        w.x = (S)lo;
        w.y = (S)hi;
    }
    __builtin_nontemporal_store(w, gmem(IdxT<I>::at((V*)p, e)));
}

// Largest element count addressed with 32-bit byte offsets for element size es.
inline bool fits_u32(size_t n_elems, size_t es) { return n_elems * es < ((size_t)1 << 31); }

// ---------------------------------------------------------------------------
// volume geometry and the block -> tile mapping
// ---------------------------------------------------------------------------
struct Geom {
    int nx, ny, nz;       // volume extent (nz = 1 for 2D)
    int tiles_x, tiles_y; // xy tiles of TX x ty
    int zc;               // planes per z chunk
    int nzc;              // number of z chunks
    int nblocks;          // tiles_x * tiles_y * nzc
    long long sxy;        // nx * ny
};

inline Geom make_geom(int nx, int ny, int nz, int ty, int zc) {
    Geom g;
    g.nx = nx; g.ny = ny; g.nz = nz;
    g.tiles_x = (nx + TX - 1) / TX;
    g.tiles_y = (ny + ty - 1) / ty;
    if (zc < 1) zc = 1;
    if (zc > nz) zc = nz;
    g.zc = zc;
    g.nzc = (nz + zc - 1) / zc;
    g.nblocks = g.tiles_x * g.tiles_y * g.nzc;
    g.sxy = (long long)nx * ny;
    return g;
}

// Hardware deals workgroup b to XCD b % 8.  Remap so XCD x receives the
// contiguous logical range [x*q + min(x, r), ...): its tiles are neighbours and
// their halos hit its own L2.
__device__ inline int xcd_logical_block(int b, int nb) {
    int q = nb / NXCD, r = nb % NXCD;
    int x = b % NXCD, s = b / NXCD;
    return x * q + (x < r ? x : r) + s;
}

struct TileCoord { int x0, y0, z0, z1; int lin; };

// Tile of logical block L (tiles in x fastest, then y, then z chunks).
__device__ inline TileCoord tile_of_logical(const Geom& g, int ty, int L) {
    TileCoord t;
    t.lin = L;
    int txi = L % g.tiles_x;
    int rem = L / g.tiles_x;
    int tyi = rem % g.tiles_y;
    int zci = rem / g.tiles_y;
    t.x0 = txi * TX;
    t.y0 = tyi * ty;
    t.z0 = zci * g.zc;
    t.z1 = t.z0 + g.zc < g.nz ? t.z0 + g.zc : g.nz;
    return t;
}

__device__ inline TileCoord tile_of_block(const Geom& g, int ty) {
    return tile_of_logical(g, ty, xcd_logical_block((int)blockIdx.x, g.nblocks));
}

// ---------------------------------------------------------------------------
// stopping rule: per-block FP64 partial sums, a fixed-order final reduction
// ---------------------------------------------------------------------------
// Device control block of one solve.  The reference's rule
// (ROF_TV_core.c:79-89): every check, re = |u - u_prev| / |u|; count the
// checks with re < epsil (cumulative) and stop when count > 3.
struct Ctrl {
    int stopped;     // 1 once the rule fired; every later kernel returns at once
    int stop_iter;   // the iteration index at which it fired
    int count;       // cumulative number of sub-tolerance checks
    float re;        // last residual, as the reference reports it
    double sum_d, sum_u;  // last check's sums (diagnostics)
};

// LDS scratch of block_sum2: two doubles per wave of the largest workgroup
// (1024 threads).
constexpr int kRedSlots = 2 * (1024 / 64);

// Sum (a, b) over the workgroup in a fixed order (within each wave, then over
// the waves in index order); result valid in thread 0.  lds holds two doubles
// per wave of the workgroup (kRedSlots covers every launch shape).
__device__ inline void block_sum2(double& a, double& b, double* lds) {
    #pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
    }
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    const int w = tid >> 6, nw = (blockDim.x * blockDim.y) >> 6;
    if ((tid & 63) == 0) { lds[2 * w] = a; lds[2 * w + 1] = b; }
    __syncthreads();
    if (tid == 0) {
        double sa = 0, sb = 0;
        for (int i = 0; i < nw; ++i) { sa += lds[2 * i]; sb += lds[2 * i + 1]; }
        a = sa; b = sb;
    }
}

// The flag is written by an earlier kernel of the same stream, so a plain
// load at kernel start sees it.
__device__ inline bool solve_stopped(const Ctrl* c) { return c != nullptr && c->stopped != 0; }

// ---------------------------------------------------------------------------
// correctly rounded single-precision square root
// ---------------------------------------------------------------------------
// For x in [2^-96, FLT_MAX], v_sqrt_f32 is within one ulp of sqrt(x) and the
// neighbours r -/+ 1 ulp of its result r are normal and finite; the rounded
// root is r - 1 ulp when x - (r - 1 ulp) r <= 0, r + 1 ulp when
// x - (r + 1 ulp) r > 0, and r otherwise (both residuals by fma, exact).
// This is the selection the compiler's IEEE expansion of sqrtf makes after
// it has scaled tiny arguments and before it patches zeros and infinities,
// so on that interval it returns sqrtf's bits (checked over every float of
// it by cilreg_check_sqrt_rn).  A wave with any lane outside the interval
// (zero, denormal, negative, infinite, NaN) calls sqrtf.
//
// The two decisions are read from the sign bits of (r -/+ 1 ulp) r - x, the
// negated residuals (an exact zero is +0, so "x - (r - 1 ulp) r <= 0" is a
// clear sign bit): the root is the bit pattern of r - 1 ulp plus one for
// each neighbour whose square lies below x (at most r + 1 ulp: the second
// implies the first).  Integer shifts and one add instead of compares and
// selects, which on this target wait two cycles each between the compare and
// its use.
__device__ inline float sqrt_rn(float x) {
    const bool fast = x >= 0x1p-96f && x <= 0x1.fffffep127f;   // false for NaN
    if (__builtin_amdgcn_ballot_w64(!fast)) return sqrtf(x);
    const float r = __builtin_amdgcn_sqrtf(x);
    const uint32_t ri = __float_as_uint(r);
    const float rm = __uint_as_float(ri - 1u), rp = __uint_as_float(ri + 1u);
    const uint32_t above_m = __float_as_uint(__builtin_fmaf(rm, r, -x)) >> 31;   // 1: x - (r - 1 ulp) r > 0
    const uint32_t above_p = __float_as_uint(__builtin_fmaf(rp, r, -x)) >> 31;   // 1: x - (r + 1 ulp) r > 0
    return __uint_as_float(ri - 1u + above_m + above_p);
}

// Correctly rounded double square root for x in [2^-1000, 2^1000] (no
// scaling needed there): an estimate g = x rsq(x), h = rsq(x) / 2, one
// coupled Newton step on both (e = 1/2 - g h), then two residual
// corrections g += (x - g^2) h, every step a fused multiply-add.
__device__ inline double sqrt_dn(double x) {
    const double r = __builtin_amdgcn_rsq(x);
    double g = x * r, h = r * 0.5;
    const double e = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, e, g);
    h = __builtin_fma(h, e, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    return __builtin_fma(d, h, g);
}

// ---------------------------------------------------------------------------
// exact restatements of the reference's mixed-precision expressions
// ---------------------------------------------------------------------------
template <typename T> struct Mixed;
template <> struct Mixed<float> {
    // sqrtf((float)((double)s + eps)): ROF_TV_core.c:160,199,223,262 (eps 1e-8;
    // LLT_ROF_core.c:256,296,320,359 with 1e-12)
    __device__ static inline float rof_sqrt_f(float s, double eps = 1.0e-8) { return sqrt_rn((float)((double)s + eps)); }
    // (float)sqrt((double)s + eps): the 3D D1, ROF_TV_core.c:136, LLT_ROF_core.c:232.
    // s >= 0 (squares and squared minmods), so x = s + eps lies in
    // [eps, 2^129) unless s is infinite or NaN: in range, sqrt_dn gives sqrt's
    // bits without its scaling and special-value patches (checked for every
    // float s with both eps values by cilreg_check_rof_sqrt_d); a wave with an
    // out-of-range lane calls sqrt.
    __device__ static inline float rof_sqrt_d(float s, double eps = 1.0e-8) {
        const double x = (double)s + eps;
        if (__builtin_amdgcn_ballot_w64(!(x < 0x1p+1000))) return (float)sqrt(x);   // also NaN
        return (float)sqrt_dn(x);
    }
    // (float)((double)num / (1.0 + lt)): PD_TV_core.c:191,249
    //
    // Without the double division: q = num * RN(1/c), c = 1 + lt, is within
    // 2^-52 |num/c| of the quotient and d = 2^-50 |q| (exact) exceeds that
    // bound plus half an ulp, so RN_d(num/c) lies in [RN_d(q-d), RN_d(q+d)];
    // rounding to float is monotone, so when both ends round to the same
    // float the double quotient does too.  A wave holding any lane where
    // they differ (a float rounding boundary within 2^-50 |q| of q, about one
    // lane in 2^26, or a non-finite num) divides instead.
    __device__ static inline float pd_div(float num, float lt) {
        const double c = 1.0 + (double)lt;
        const double q = (double)num * (1.0 / c);
        const double d = fabs(q) * 0x1p-50;
        const float lo = (float)(q - d), hi = (float)(q + d);
        float r = lo;
        if (__builtin_amdgcn_ballot_w64(!(lo == hi))) r = (float)((double)num / c);
        return r;
    }
    // squared minmod, ROF_TV_core.c:132-133.  The reference multiplies in
    // double (fabs promotes) and rounds the product to float; the factor is
    // 0, +-0.5 or +-1, so the double product is exact and its rounding to
    // float equals the correctly rounded float product: float arithmetic
    // gives the same bits.
    __device__ static inline float minmod_sq(float a, float b) {
        int sa = (a > 0.f) - (a < 0.f), sb = (b > 0.f) - (b < 0.f);
        const float fa = fabsf(a), fb = fabsf(b);
        const float m = (0.5f * (float)(sa + sb)) * (fa < fb ? fa : fb);
        return m * m;
    }
    // N squared minmods at once.  For finite a, b the reference's product is
    // +-min(|a|,|b|) when the signs agree and 0 otherwise (a factor 0.5
    // arises only when one of them is 0, which makes the minimum 0): the
    // median of (a, b, 0).  It differs only when a or b is NaN or when a and
    // b are infinities of opposite sign; a*b is then NaN or -inf, and a wave
    // holding any lane whose a*b is not above -inf (overflow to -inf
    // included) takes the reference's expression instead.
    template <int N>
    __device__ static inline void minmod_sq_n(const float (&a)[N], const float (&b)[N], float (&m)[N]) {
        bool odd = false;
        #pragma unroll
        for (int q = 0; q < N; ++q) odd = odd || !(a[q] * b[q] > -__builtin_inff());
        if (__builtin_amdgcn_ballot_w64(odd)) {
            #pragma unroll
            for (int q = 0; q < N; ++q) m[q] = minmod_sq(a[q], b[q]);
        } else {
            #pragma unroll
            for (int q = 0; q < N; ++q) {
                const float v = __builtin_amdgcn_fmed3f(a[q], b[q], 0.f);
                m[q] = v * v;
            }
        }
    }
};
template <> struct Mixed<double> {
    __device__ static inline double rof_sqrt_f(double s, double eps = 1.0e-8) { return sqrt(s + eps); }
    __device__ static inline double rof_sqrt_d(double s, double eps = 1.0e-8) { return sqrt(s + eps); }
    __device__ static inline double pd_div(double num, double lt) { return num / (1.0 + lt); }
    __device__ static inline double minmod_sq(double a, double b) {
        int sa = (a > 0.0) - (a < 0.0), sb = (b > 0.0) - (b < 0.0);
        double fa = fabs(a), fb = fabs(b);
        double m = (0.5 * (double)(sa + sb)) * (fa < fb ? fa : fb);
        return m * m;
    }
    template <int N>
    __device__ static inline void minmod_sq_n(const double (&a)[N], const double (&b)[N], double (&m)[N]) {
        #pragma unroll
        for (int q = 0; q < N; ++q) m[q] = minmod_sq(a[q], b[q]);
    }
};

template <typename T> __device__ inline T dsqrt(T x);
template <> __device__ inline float dsqrt<float>(float x) { return sqrt_rn(x); }
template <> __device__ inline double dsqrt<double>(double x) { return sqrt(x); }

}  // namespace cilreg
