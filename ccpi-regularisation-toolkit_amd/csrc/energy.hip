// energy.hip -- the ROF objective in FP64 on the device: the "final objective"
// of the precision study.
//
// Reference: TV_energy2D / TV_energy3D, src/Core/regularisers_CPU/utils.c:64-117:
//   E = sum_x 2 lambda |grad+ u(x)| + sum_x (u(x) - u0(x))^2      (type 1)
//   E = sum_x 2 lambda |grad+ u(x)|                                (type 2)
// with forward differences and a replicate boundary (the difference along an
// axis is 0 at its last index).  The reference sums serially in float; here
// every term is formed and summed in double, in a fixed order that does not
// depend on the device: a fixed grid of EB workgroups, each thread striding
// over the voxels, a fixed-order workgroup reduction, then one workgroup
// summing the EB partials.  The same call on the same data gives the same
// bits.  HBM traffic: u read once (its +x/+y/+z neighbours come from cache)
// and u0 once, i.e. 8 B per voxel in FP32.
#include "runtime.hpp"
#include "../../include/cilreg_ext.h"

namespace cilreg {

namespace {

constexpr int EB = 2048;   // workgroups of the partial-sum kernel
constexpr int ET = 256;    // threads per workgroup

__device__ inline void block_sum2_256(double& a, double& b, double* lds) {
    lds[threadIdx.x] = a;
    lds[ET + threadIdx.x] = b;
    __syncthreads();
    for (int s = ET / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            lds[threadIdx.x] += lds[threadIdx.x + s];
            lds[ET + threadIdx.x] += lds[ET + threadIdx.x + s];
        }
        __syncthreads();
    }
    a = lds[0];
    b = lds[ET];
}

template <typename T>
__global__ __launch_bounds__(ET) void energy_partials_kernel(const T* __restrict__ U, const T* __restrict__ U0,
                                                             double two_lambda, int nx, int ny, int nz,
                                                             double* partials) {
    __shared__ double lds[2 * ET];
    const long long sxy = (long long)nx * ny, n = sxy * nz;
    double eg = 0.0, ed = 0.0;
    for (long long c = (long long)blockIdx.x * ET + threadIdx.x; c < n; c += (long long)EB * ET) {
        const int i = (int)(c % nx);
        const long long r = c / nx;
        const int j = (int)(r % ny);
        const int k = (int)(r / ny);
        const double u = (double)U[c];
        const double gx = (i + 1 < nx ? (double)U[c + 1] : u) - u;
        const double gy = (j + 1 < ny ? (double)U[c + nx] : u) - u;
        const double gz = (k + 1 < nz ? (double)U[c + sxy] : u) - u;
        eg += two_lambda * sqrt(gx * gx + gy * gy + gz * gz);
        const double dd = u - (double)U0[c];
        ed += dd * dd;
    }
    block_sum2_256(eg, ed, lds);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = eg;
        partials[2 * blockIdx.x + 1] = ed;
    }
}

__global__ __launch_bounds__(ET) void energy_final_kernel(const double* partials, double* out) {
    __shared__ double lds[2 * ET];
    double eg = 0.0, ed = 0.0;
    for (int b = threadIdx.x; b < EB; b += ET) {   // fixed order per thread
        eg += partials[2 * b];
        ed += partials[2 * b + 1];
    }
    block_sum2_256(eg, ed, lds);
    if (threadIdx.x == 0) {
        out[0] = eg;
        out[1] = ed;
    }
}

template <typename T>
int tv_energy(const T* d_U, const T* d_U0, double lambda, int type, int nx, int ny, int nz, double* energy,
              void* stream) {
    if (!d_U || !d_U0 || !energy) { set_error("null pointer argument"); return -1; }
    if (nx <= 0 || ny <= 0 || nz < 0) { set_error("invalid dimensions (%d, %d, %d)", nx, ny, nz); return -1; }
    if (type != 1 && type != 2) { set_error("energy type must be 1 or 2, got %d", type); return -1; }
    if (nz == 0) nz = 1;
    hipStream_t s = caller_stream(stream);
    if (!s) { set_error("no stream"); return -1; }
    Arena ar;
    const size_t op = ar.reserve((size_t)EB * 2 * sizeof(double));
    const size_t oo = ar.reserve(2 * sizeof(double));
    if (ar.allocate() != 0) return -1;
    double* part = (double*)ar.at(op);
    double* out = (double*)ar.at(oo);
    hipLaunchKernelGGL((energy_partials_kernel<T>), dim3(EB), dim3(ET), 0, s, d_U, d_U0, 2.0 * lambda, nx, ny, nz,
                       part);
    hipLaunchKernelGGL(energy_final_kernel, dim3(1), dim3(ET), 0, s, (const double*)part, out);
    double h[2] = {0, 0};
    CILREG_CK(hipGetLastError());
    CILREG_CK(hipMemcpyAsync(h, out, sizeof h, hipMemcpyDeviceToHost, s));
    CILREG_CK(hipStreamSynchronize(s));
    *energy = type == 2 ? h[0] : h[0] + h[1];
    return 0;
}

}  // namespace

}  // namespace cilreg

extern "C" {

int cilreg_tv_energy_dev(const float* d_U, const float* d_U0, double lambda, int type, int dimX, int dimY, int dimZ,
                         double* energy, void* stream) {
    return cilreg::tv_energy<float>(d_U, d_U0, lambda, type, dimX, dimY, dimZ, energy, stream);
}

int cilreg_tv_energy_dev_f64(const double* d_U, const double* d_U0, double lambda, int type, int dimX, int dimY,
                             int dimZ, double* energy, void* stream) {
    return cilreg::tv_energy<double>(d_U, d_U0, lambda, type, dimX, dimY, dimZ, energy, stream);
}

}  // extern "C"
