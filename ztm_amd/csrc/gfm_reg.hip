// Smoothness term of the regularised misfit
//     f(rho) = 0.5 ||G rho - d||^2 + 0.5 lambda ||grad rho||^2
// with grad = forward differences along i, j and k inside the grid (no
// differences across the boundary). Its gradient is lambda * L^T L rho, with
//     (L^T L rho)_n = deg(n) rho_n - sum of rho over the axis neighbours of n,
// deg(n) = number of axis neighbours inside the grid (2..6). One thread per
// cell; the kernel is HBM-bound (7 reads, 1 write per cell, mostly cached).
// No reference counterpart: ZTM's inversion uses a finite-difference
// Jacobian (TFM.cpp:124-135) and no smoothness prior.
#include "gfm_internal.h"

namespace ztm {

template <typename T>
__global__ void ltl_kernel(const T* __restrict__ rho, T* __restrict__ out, int nx, int ny,
                           int nz, int64_t n_cells) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= n_cells) return;
    const int i = (int)(n % nx);
    const int64_t rw = n / nx;
    const int j = (int)(rw % ny);
    const int k = (int)(rw / ny);
    const int64_t sy = nx, sz = (int64_t)nx * ny;
    const T c = rho[n];
    T acc = T(0);
    if (i > 0) acc += c - rho[n - 1];
    if (i < nx - 1) acc += c - rho[n + 1];
    if (j > 0) acc += c - rho[n - sy];
    if (j < ny - 1) acc += c - rho[n + sy];
    if (k > 0) acc += c - rho[n - sz];
    if (k < nz - 1) acc += c - rho[n + sz];
    out[n] = acc;
}

int ltl_launch(int precision, const void* rho, void* out, int nx, int ny, int nz,
               hipStream_t stream) {
    const int64_t n = (int64_t)nx * ny * nz;
    if (n == 0) return ZTM_OK;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    if (precision == ZTM_FP64)
        hipLaunchKernelGGL(ltl_kernel<double>, dim3(blocks), dim3(256), 0, stream,
                           (const double*)rho, (double*)out, nx, ny, nz, n);
    else
        hipLaunchKernelGGL(ltl_kernel<float>, dim3(blocks), dim3(256), 0, stream,
                           (const float*)rho, (float*)out, nx, ny, nz, n);
    ZTM_HIP(hipGetLastError());
    return ZTM_OK;
}

}  // namespace ztm
