// Stored-matrix variant: G is materialised once per plan as FP32 [M][ld]
// (station-major, row stride ld = N rounded up to 4 with zero padding, 4 B per
// station-cell pair), each entry evaluated in FP64 --
// the point-mass term or the exact prism integral -- and rounded once, so a
// stored entry carries no FP32 geometry error and needs no near-field
// correction. Forward and adjoint then stream G from HBM:
//   forward  g[m]  = sum_n G[m][n] rho[n]   (rows streamed, 8 stations per
//                                            workgroup share each rho load)
//   adjoint  out[n] = sum_m G[m][n] r[m]    (columns: a lane owns 4 cells and
//                                            walks the rows; coalesced 16 B)
// Both are HBM-bound: 4 B of G per pair per pass; accumulation is FP64.
// This is the layout that lets one geometry's G be reused across the
// hundreds of forward models of an inversion step (TFM.cpp:124-135), where
// recomputing the pair terms would repeat the same work.
#include "gfm_internal.h"
#include "gfm_pair.h"

#include <algorithm>

namespace ztm {

// ---------------------------------------------------------- materialise
// grid = (ceil(N / 256), M); one thread per entry of row m.
template <bool PRISM>
__global__ __launch_bounds__(256) void store_kernel(const double4* __restrict__ st, int nx,
                                                    int ny, int64_t n_cells, int64_t ld,
                                                    double sy, double sz, double scale,
                                                    float* __restrict__ G) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= n_cells) return;
    const int m = blockIdx.y;
    const double4 p = st[m];
    const int i = (int)(n % nx);
    const int64_t rw = n / nx;
    const int j = (int)(rw % ny);
    const int k = (int)(rw / ny);
    double v;
    if constexpr (PRISM) {
        v = prism_cell(p.x, p.y, p.z, i, j, k, sy, sz);
    } else {
        const double a = p.x - i, dy = p.y - j * sy, dz = p.z - k * sz;
        const double r2 = a * a + dy * dy + dz * dz + kEpsD;
        v = dz / (r2 * sqrt(r2));
    }
    G[(size_t)m * ld + n] = (float)(scale * v);
}

// --------------------------------------------------------------- forward
constexpr int kRows = 8;   // stations per workgroup

// grid = ceil(M / kRows); block = 256; each thread walks float4 columns.
__global__ __launch_bounds__(256) void stored_fwd_kernel(const float* __restrict__ G,
                                                         const float* __restrict__ rho,
                                                         int m, int64_t n_cells, int64_t ld,
                                                         float thr, double* __restrict__ g) {
    __shared__ double red[kRows][4];
    const int m0 = blockIdx.x * kRows;
    const int nrow = min(kRows, m - m0);
    double acc[kRows];
#pragma unroll
    for (int r = 0; r < kRows; ++r) acc[r] = 0.0;
    const int64_t n4 = n_cells / 4;
    const float4* rho4 = reinterpret_cast<const float4*>(rho);
    for (int64_t q = threadIdx.x; q < n4; q += blockDim.x) {
        float4 x = rho4[q];
        x.x = fabsf(x.x) > thr ? x.x : 0.f;
        x.y = fabsf(x.y) > thr ? x.y : 0.f;
        x.z = fabsf(x.z) > thr ? x.z : 0.f;
        x.w = fabsf(x.w) > thr ? x.w : 0.f;
#pragma unroll
        for (int r = 0; r < kRows; ++r) {
            if (r < nrow) {
                const float4 gv = reinterpret_cast<const float4*>(G + (size_t)(m0 + r) * ld)[q];
                acc[r] = fma((double)gv.x, (double)x.x, acc[r]);
                acc[r] = fma((double)gv.y, (double)x.y, acc[r]);
                acc[r] = fma((double)gv.z, (double)x.z, acc[r]);
                acc[r] = fma((double)gv.w, (double)x.w, acc[r]);
            }
        }
    }
    // ragged tail (n_cells % 4 values)
    for (int64_t n = n4 * 4 + threadIdx.x; n < n_cells; n += blockDim.x) {
        float x = rho[n];
        x = fabsf(x) > thr ? x : 0.f;
        for (int r = 0; r < nrow; ++r)
            acc[r] = fma((double)G[(size_t)(m0 + r) * ld + n], (double)x, acc[r]);
    }
    // wave reduction (fixed butterfly), then across the 4 waves
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
        double v = acc[r];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (lane == 0) red[r][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x < nrow) {
        const int r = threadIdx.x;
        g[m0 + r] = red[r][0] + red[r][1] + red[r][2] + red[r][3];
    }
}

// --------------------------------------------------------------- adjoint
// grid = (ceil(ceil(N/4) / 256), splits); a lane owns 4 consecutive cells
// (one aligned float4 of each padded row); station range split over
// blockIdx.y, FP64 partials [splits][N].
__global__ __launch_bounds__(256) void stored_adj_kernel(const float* __restrict__ G,
                                                         const float* __restrict__ r, int m,
                                                         int64_t n_cells, int64_t ld,
                                                         int st_per_split,
                                                         double* __restrict__ part) {
    const int64_t n = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (n >= n_cells) return;
    const int mb = blockIdx.y * st_per_split;
    const int me = min(mb + st_per_split, m);
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    for (int mm = mb; mm < me; ++mm) {
        const double w = (double)r[mm];
        const float4 gv = *reinterpret_cast<const float4*>(G + (size_t)mm * ld + n);
        a0 = fma((double)gv.x, w, a0);
        a1 = fma((double)gv.y, w, a1);
        a2 = fma((double)gv.z, w, a2);
        a3 = fma((double)gv.w, w, a3);
    }
    double* out = part + (size_t)blockIdx.y * n_cells;
    out[n] = a0;
    if (n + 1 < n_cells) out[n + 1] = a1;
    if (n + 2 < n_cells) out[n + 2] = a2;
    if (n + 3 < n_cells) out[n + 3] = a3;
}

__global__ void stored_adj_reduce(const double* __restrict__ part, int splits, int64_t n_cells,
                                  float* __restrict__ out) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= n_cells) return;
    double s = 0.0;
    for (int k = 0; k < splits; ++k) s += part[(size_t)k * n_cells + n];
    out[n] = (float)s;
}

// --------------------------------------------------------------- launchers
int64_t stored_ld(const PmShape& s) { return (s.n_cells + 3) / 4 * 4; }

int stored_build_launch(const PmShape& s, bool prism, const double4* st, double sy, double sz,
                        double scale, float* G, hipStream_t stream) {
    if (s.m == 0 || s.n_cells == 0) return ZTM_OK;
    if (s.m > 65535) return fail(ZTM_EINVAL, "stored matrix supports at most 65535 stations");
    const int64_t ld = stored_ld(s);
    if (int rc = zero_fill(G, sizeof(float) * (size_t)s.m * ld, stream)) return rc;
    const dim3 grid((unsigned)((s.n_cells + 255) / 256), (unsigned)s.m);
    {
        ProfScope prof("stored_build", stream);
        if (prism)
            hipLaunchKernelGGL(store_kernel<true>, grid, dim3(256), 0, stream, st, s.nx, s.ny,
                               s.n_cells, ld, sy, sz, scale, G);
        else
            hipLaunchKernelGGL(store_kernel<false>, grid, dim3(256), 0, stream, st, s.nx, s.ny,
                               s.n_cells, ld, sy, sz, scale, G);
        ZTM_HIP(hipGetLastError());
    }
    return ZTM_OK;
}

int stored_fwd_launch(const PmShape& s, const float* G, const float* rho, double thr,
                      double* g, hipStream_t stream) {
    if (s.m == 0) return ZTM_OK;
    if (((uintptr_t)rho & 15) != 0) return fail(ZTM_EINVAL, "rho must be 16-byte aligned");
    ProfScope prof("stored_fwd", stream);
    hipLaunchKernelGGL(stored_fwd_kernel, dim3((s.m + kRows - 1) / kRows), dim3(256), 0, stream,
                       G, rho, s.m, s.n_cells, stored_ld(s), (float)thr, g);
    ZTM_HIP(hipGetLastError());
    return ZTM_OK;
}

int stored_adj_splits(const PmShape& s) {
    const int64_t wgs = ((s.n_cells + 3) / 4 + 255) / 256;
    const int sp = (int)std::max<int64_t>(1, (2048 + wgs - 1) / wgs);
    return std::max(1, std::min(sp, std::max(1, s.m / 64)));
}

int stored_adj_launch(const PmShape& s, const float* G, const float* r, double* part, float* out,
                      hipStream_t stream) {
    if (s.n_cells == 0) return ZTM_OK;
    if (s.m == 0) return zero_fill(out, sizeof(float) * (size_t)s.n_cells, stream);
    const int splits = stored_adj_splits(s);
    const int per = (s.m + splits - 1) / splits;
    {
        ProfScope prof("stored_adj", stream);
        hipLaunchKernelGGL(stored_adj_kernel,
                           dim3((unsigned)(((s.n_cells + 3) / 4 + 255) / 256), splits), dim3(256),
                           0, stream, G, r, s.m, s.n_cells, stored_ld(s), per, part);
        ZTM_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(stored_adj_reduce, dim3((unsigned)((s.n_cells + 255) / 256)), dim3(256), 0,
                       stream, part, splits, s.n_cells, out);
    ZTM_HIP(hipGetLastError());
    return ZTM_OK;
}

}  // namespace ztm
