// Right-rectangular-prism cell formula (Plouff 1976; Nagy, Papp & Benedek
// 2000): each cell contributes the exact volume integral of (z_m - z)/r^3
// instead of the point-mass term V dz/r^3 of GFM_GPU.cu:22. No reference
// counterpart; the FP64 oracle is oracle_prism_* (gfm_oracle.c).
//
// FP32 path (hybrid). The 8-corner closed form is a third difference of
// F ~ r ln r, so in FP32 it loses ~r^2 ln r of relative accuracy per cell
// (1e-3 at 8 cells). Instead:
//   far field (beyond kMid cells): the point-mass kernel of the main kernels,
//     plus the quadrupole term for non-cubic cells (error O((a/r)^4), below
//     2e-5 of the cell's own 1/r^2 scale at 8 cells);
//   mid zone (kNear < d <= kMid cells): 3x3x3 Gauss-Legendre in FP32;
//   near field (d <= kNear cells): the closed form in FP64.
// The main kernels run the far-field kernel on every pair; the mid-zone and
// near-field pairs are then corrected to their better value (corr_pair in
// gfm_pair.h): per station here for the forward, per cell in the adjoint's
// reduce kernel. Distances d are per-axis cell-index distances to the
// station's nearest cell.
//
// FP64 path (the FP64 GPU path of the prism). The closed form on every pair,
// with the four x-corner planes of a row shared between neighbouring cells:
// cell i's value is G(i+1/2) - G(i-1/2), G = four-corner sum over y and z.
#include "gfm_internal.h"
#include "gfm_pair.h"

#include <algorithm>

namespace ztm {

// ------------------------------------------------ FP32 mid-zone correction
// One wave per station; lanes stride over the correction box (radii rx, ry,
// rz, clamped to the grid) and then the cells of clamped rows outside it;
// fixed butterfly reduction, so corr[m] is deterministic.
template <bool QUAD>
__global__ __launch_bounds__(256) void prism_near_fwd_kernel(NearArgs na,
                                                             const float* __restrict__ rho,
                                                             int m, double* __restrict__ corr) {
    const int lane = threadIdx.x & 63;
    const int st = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (st >= m) return;   // uniform per wave
    const float4 pf = na.st_f[st];
    const double4 pd = na.st_d[st];
    const int4 c = na.st_c[st];
    const int i0 = max(c.x - na.rx, 0), i1 = min(c.x + na.rx, na.nx - 1);
    const int j0 = max(c.y - na.ry, 0), j1 = min(c.y + na.ry, na.ny - 1);
    const int k0 = max(c.z - na.rz, 0), k1 = min(c.z + na.rz, na.nz - 1);
    const int bi = i1 - i0 + 1, bj = j1 - j0 + 1, bk = k1 - k0 + 1;
    const int total = (bi > 0 && bj > 0 && bk > 0) ? bi * bj * bk : 0;
    double acc = 0.0;
    for (int t = lane; t < total; t += 64) {
        const int ii = i0 + t % bi;
        const int rest = t / bi;
        const int jj = j0 + rest % bj;
        const int kk = k0 + rest / bj;
        const float w = rho[((size_t)kk * na.ny + jj) * na.nx + ii];
        if (!(fabsf(w) > na.thr)) continue;
        acc += corr_pair<QUAD>(kCorrPrism, w, pf, pd, c, ii, jj, kk, na);
    }
    acc += clamped_rows_corr<QUAD>(kCorrPrism, pf, pd, c, rho, na, lane, 64);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) corr[st] = acc;
}

int prism_near_fwd_launch(const PmShape& s, const NearArgs& near, const float* rho,
                          double* corr, hipStream_t stream) {
    if (s.m == 0) return ZTM_OK;
    const dim3 grid((s.m + 3) / 4);
    ProfScope prof("prism_near_fwd", stream);
    if (near.quad)
        hipLaunchKernelGGL(prism_near_fwd_kernel<true>, grid, dim3(256), 0, stream, near, rho,
                           s.m, corr);
    else
        hipLaunchKernelGGL(prism_near_fwd_kernel<false>, grid, dim3(256), 0, stream, near, rho,
                           s.m, corr);
    ZTM_HIP(hipGetLastError());
    return ZTM_OK;
}

// ZTM_FP64_ZONED: the same walk over the correction box in FP64 (FP64
// densities and threshold, corr_pair64: FP64 Gauss-Legendre mid zone, closed
// form near zone); FP64 rows are never clamped, so there is no second part.
template <bool QUAD>
__global__ __launch_bounds__(256) void prism_near_fwd64_kernel(NearArgs na,
                                                               const double* __restrict__ rho,
                                                               double thr, int m,
                                                               double* __restrict__ corr) {
    const int lane = threadIdx.x & 63;
    const int st = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (st >= m) return;   // uniform per wave
    const double4 pd = na.st_d[st];
    const int4 c = na.st_c[st];
    const int i0 = max(c.x - na.rx, 0), i1 = min(c.x + na.rx, na.nx - 1);
    const int j0 = max(c.y - na.ry, 0), j1 = min(c.y + na.ry, na.ny - 1);
    const int k0 = max(c.z - na.rz, 0), k1 = min(c.z + na.rz, na.nz - 1);
    const int bi = i1 - i0 + 1, bj = j1 - j0 + 1, bk = k1 - k0 + 1;
    const int total = (bi > 0 && bj > 0 && bk > 0) ? bi * bj * bk : 0;
    double acc = 0.0;
    for (int t = lane; t < total; t += 64) {
        const int ii = i0 + t % bi;
        const int rest = t / bi;
        const int jj = j0 + rest % bj;
        const int kk = k0 + rest / bj;
        const double w = rho[((size_t)kk * na.ny + jj) * na.nx + ii];
        if (!(fabs(w) > thr)) continue;
        acc += corr_pair64<QUAD>(w, pd, c, ii, jj, kk, na);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) corr[st] = acc;
}

int prism_near_fwd64_launch(const PmShape& s, const NearArgs& near, const double* rho,
                            double thr, double* corr, hipStream_t stream) {
    if (s.m == 0) return ZTM_OK;
    const dim3 grid((s.m + 3) / 4);
    ProfScope prof("prism_near_fwd64", stream);
    if (near.quad)
        hipLaunchKernelGGL(prism_near_fwd64_kernel<true>, grid, dim3(256), 0, stream, near, rho,
                           thr, s.m, corr);
    else
        hipLaunchKernelGGL(prism_near_fwd64_kernel<false>, grid, dim3(256), 0, stream, near, rho,
                           thr, s.m, corr);
    ZTM_HIP(hipGetLastError());
    return ZTM_OK;
}

// ----------------------------------------------------- FP64 exact kernels
// Four-corner sum over (y, z) at one x-corner plane, station-relative.
__device__ __forceinline__ double corner_plane(double u, double v0, double v1, double w0,
                                               double w1) {
    return prism_F(u, v0, w0) - prism_F(u, v0, w1) - prism_F(u, v1, w0) + prism_F(u, v1, w1);
}

// Forward: grid = (m_pad / 64, n_slices); block = 256; one station per lane,
// the 4 waves split the slice's rows (staged in LDS) and reduce across waves.
__global__ __launch_bounds__(256) void prism64_fwd_kernel(
    const double4* __restrict__ st, const double* __restrict__ rho, int nx, int ny, int n_rows,
    double sy, double sz, double thr, int rows_per_slice, double* __restrict__ part,
    int m_pad) {
    extern __shared__ __align__(16) unsigned char smem[];
    double* lrho = reinterpret_cast<double*>(smem);
    const int row0 = blockIdx.y * rows_per_slice;
    const int nr = min(rows_per_slice, n_rows - row0);
    const size_t base = (size_t)row0 * nx;
    for (int t = threadIdx.x; t < nr * nx; t += blockDim.x) {
        double v = rho[base + t];
        if (!(fabs(v) > thr)) v = 0.0;
        lrho[t] = v;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int mi = blockIdx.x * 64 + lane;
    const double4 p = st[mi];
    const double hy = 0.5 * fabs(sy), hz = 0.5 * fabs(sz);
    double acc = 0.0;
    for (int r = wave; r < nr; r += 4) {
        const int row = row0 + r;
        const int k = row / ny, j = row - k * ny;
        const double v0 = j * sy - hy - p.y, v1 = j * sy + hy - p.y;
        const double w0 = k * sz - hz - p.z, w1 = k * sz + hz - p.z;
        const double* lr = lrho + (size_t)r * nx;
        double gprev = corner_plane(-0.5 - p.x, v0, v1, w0, w1);
        double ar = 0.0;
        for (int i = 0; i < nx; ++i) {
            const double gnext = corner_plane((double)i + 0.5 - p.x, v0, v1, w0, w1);
            ar = fma(lr[i], gnext - gprev, ar);
            gprev = gnext;
        }
        acc += ar;
    }
    __syncthreads();
    double* red = reinterpret_cast<double*>(smem);
    red[wave * 64 + lane] = acc;
    __syncthreads();
    if (wave == 0)
        part[(size_t)blockIdx.y * m_pad + mi] =
            red[lane] + red[64 + lane] + red[128 + lane] + red[192 + lane];
}

__global__ void prism64_fwd_reduce(const double* __restrict__ part, int n_slices, int m_pad,
                                   int m, double scale, double* __restrict__ g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    double s = 0.0;
    for (int k = 0; k < n_slices; ++k) s += part[(size_t)k * m_pad + i];
    g[i] = scale * s;
}

// Adjoint: a lane owns kC64 consecutive cells of one row (C + 1 x-corner
// planes per station); station tiles of 256 in LDS; 4 waves split the tile
// and reduce across waves; one split over stations per blockIdx.y.
constexpr int kC64 = 4;

__global__ __launch_bounds__(256) void prism64_adj_kernel(
    const double4* __restrict__ st, const double* __restrict__ r, int m, int nx, int ny,
    double sy, double sz, int segs_per_row, int64_t n_segs, int st_per_split, int m_pad,
    double* __restrict__ part, int64_t n_cells) {
    __shared__ double4 tile[256];
    __shared__ double red[4 * kC64 * 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t seg = (int64_t)blockIdx.x * 64 + lane;
    const int64_t segc = seg < n_segs ? seg : 0;
    const int64_t row = segc / segs_per_row;
    const int i0 = (int)(segc - row * segs_per_row) * kC64;
    const int k = (int)(row / ny), j = (int)(row - (int64_t)k * ny);
    const double hy = 0.5 * fabs(sy), hz = 0.5 * fabs(sz);
    double acc[kC64];
#pragma unroll
    for (int c = 0; c < kC64; ++c) acc[c] = 0.0;
    const int mb = blockIdx.y * st_per_split;
    const int me = min(mb + st_per_split, m_pad);
    for (int t0 = mb; t0 < me; t0 += 256) {
        __syncthreads();
        {
            const int mm = t0 + threadIdx.x;
            double4 p = st[mm];
            p.w = mm < m ? r[mm] : 0.0;
            tile[threadIdx.x] = p;
        }
        __syncthreads();
        for (int s = wave * 64; s < wave * 64 + 64; ++s) {
            const double4 p = tile[s];
            if (p.w == 0.0) continue;   // uniform across the wave
            const double v0 = j * sy - hy - p.y, v1 = j * sy + hy - p.y;
            const double w0 = k * sz - hz - p.z, w1 = k * sz + hz - p.z;
            double gprev = corner_plane((double)i0 - 0.5 - p.x, v0, v1, w0, w1);
#pragma unroll
            for (int c = 0; c < kC64; ++c) {
                const double gnext = corner_plane((double)(i0 + c) + 0.5 - p.x, v0, v1, w0, w1);
                acc[c] = fma(p.w, gnext - gprev, acc[c]);
                gprev = gnext;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < kC64; ++c) red[(wave * kC64 + c) * 64 + lane] = acc[c];
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * kC64; e += blockDim.x) {
        const int c = e >> 6, l = e & 63;
        const double v = red[c * 64 + l] + red[(kC64 + c) * 64 + l] +
                         red[(2 * kC64 + c) * 64 + l] + red[(3 * kC64 + c) * 64 + l];
        const int64_t sg = (int64_t)blockIdx.x * 64 + l;
        if (sg >= n_segs) continue;
        const int64_t rw = sg / segs_per_row;
        const int ii = (int)(sg - rw * segs_per_row) * kC64 + c;
        if (ii >= nx) continue;
        part[(size_t)blockIdx.y * n_cells + rw * nx + ii] = v;
    }
}

__global__ void prism64_adj_reduce(const double* __restrict__ part, int n_splits,
                                   int64_t n_cells, double scale, double* __restrict__ out) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= n_cells) return;
    double s = 0.0;
    for (int k = 0; k < n_splits; ++k) s += part[(size_t)k * n_cells + n];
    out[n] = scale * s;
}

namespace {
constexpr int kPrismTargetWgs = 2048;

int p64_rows_per_slice(const PmShape& s) {
    const int lds_rows = std::max(1, (int)(32768 / ((size_t)s.nx * sizeof(double))));
    const int tiles = s.m_pad / 64;
    const int64_t want = std::max<int64_t>(1, (kPrismTargetWgs + tiles - 1) / tiles);
    int64_t rps = (s.n_rows + want - 1) / want;
    rps = std::max<int64_t>(1, std::min<int64_t>(rps, lds_rows));
    return (int)rps;
}

int p64_splits(const PmShape& s) {
    const int64_t segs = s.n_rows * ((s.nx + kC64 - 1) / kC64);
    const int64_t wgs = (segs + 63) / 64;
    const int sp = (int)std::max<int64_t>(1, (kPrismTargetWgs + wgs - 1) / wgs);
    return std::max(1, std::min(sp, s.m_pad / 256));
}
}  // namespace

size_t prism64_workspace(const PmShape& s, const PmTuning& t, bool adjoint) {
    (void)t;
    if (adjoint) return (size_t)p64_splits(s) * (size_t)s.n_cells;
    const int rps = p64_rows_per_slice(s);
    return (size_t)((s.n_rows + rps - 1) / rps) * (size_t)s.m_pad;
}

int prism64_fwd_launch(const PmShape& s, const PmTuning& t, const double4* st,
                       const double* rho, double sy, double sz, double thr, double scale,
                       double* part, double* g, hipStream_t stream) {
    (void)t;
    if (s.m == 0) return ZTM_OK;
    if ((size_t)s.nx * sizeof(double) > 65536)
        return fail(ZTM_EINVAL, "nx=%d too large for one LDS row", s.nx);
    const int rps = p64_rows_per_slice(s);
    const int n_slices = (int)((s.n_rows + rps - 1) / rps);
    if (n_slices > 65535) return fail(ZTM_EINVAL, "too many row slices (%d)", n_slices);
    const size_t lds = std::max((size_t)rps * s.nx * sizeof(double), (size_t)256 * sizeof(double));
    {
        ProfScope prof("prism64_fwd", stream);
        hipLaunchKernelGGL(prism64_fwd_kernel, dim3(s.m_pad / 64, n_slices), dim3(256), lds,
                           stream, st, rho, s.nx, s.ny, (int)s.n_rows, sy, sz, thr, rps, part,
                           s.m_pad);
        ZTM_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(prism64_fwd_reduce, dim3((s.m + 255) / 256), dim3(256), 0, stream, part,
                       n_slices, s.m_pad, s.m, scale, g);
    ZTM_HIP(hipGetLastError());
    return ZTM_OK;
}

int prism64_adj_launch(const PmShape& s, const PmTuning& t, const double4* st,
                       const double* r, double sy, double sz, double scale, double* part,
                       double* out, hipStream_t stream) {
    (void)t;
    if (s.n_cells == 0) return ZTM_OK;
    const int splits = p64_splits(s);
    const int segs_per_row = (s.nx + kC64 - 1) / kC64;
    const int64_t n_segs = s.n_rows * segs_per_row;
    const int tiles = s.m_pad / 256;
    const int per_split = ((tiles + splits - 1) / splits) * 256;
    {
        ProfScope prof("prism64_adj", stream);
        hipLaunchKernelGGL(prism64_adj_kernel, dim3((unsigned)((n_segs + 63) / 64), splits),
                           dim3(256), 0, stream, st, r, s.m, s.nx, s.ny, sy, sz, segs_per_row,
                           n_segs, per_split, s.m_pad, part, s.n_cells);
        ZTM_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(prism64_adj_reduce, dim3((unsigned)((s.n_cells + 255) / 256)), dim3(256),
                       0, stream, part, splits, s.n_cells, scale, out);
    ZTM_HIP(hipGetLastError());
    return ZTM_OK;
}

}  // namespace ztm
