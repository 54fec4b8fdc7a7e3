// FP64 ensemble: Y = G P and X = G^T R for B models or residual vectors with
// G (stations x cells, point mass) generated once per tile in FP64 and
// multiplied on FP64 MFMA (v_mfma_f64_16x16x4_f64). It is the split-BF16/FP16
// ensemble's algorithm -- G formed on the fly and reused across the whole
// batch -- with double everywhere: inputs, G, products and accumulation.
//
// Work decomposition (forward; the adjoint swaps stations and cells): a
// workgroup = 4 waves; a wave owns 32 rows (two 16-row MFMA tiles) and the
// workgroup's model group (NT <= 8 16-column tiles, 4 doubles per lane each;
// a batch above 128 runs as two groups on grid z, each forming its own G, so
// that two workgroups fit a CU). The K loop walks the
// slice 16 values at a time: the workgroup stages the chunk of P (16 K values
// x Bp models, FP64, thresholded) in LDS, prefetching the next chunk into
// registers meanwhile; each wave forms its A operands -- one G value per lane
// per row tile and K step, lane = row lane%16, K value lane/16 -- and issues
// 2 x NT MFMAs per K step, the B operand of model tile t read from LDS once
// for both row tiles.
// MFMA f64 16x16x4 lane layout (measured on gfx950): A[i][k] and B[k][j] at
// lane i + 16k (resp. j + 16k); D[i][j] in register i / 4 of lane
// j + 16 (i % 4).
//
// Layouts (HBM): P [B][N] and R [B][M] FP64 (model-major, as the caller holds
// them); Y [B][M], X [B][N] FP64; K-slice partials [KS][rows_pad][Bp] FP64,
// summed in a fixed order by the reduce.
#include "gfm_internal.h"

#include <algorithm>
#include <cmath>

namespace ztm {

typedef double f64x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int kChunk64 = 16;               // K values per LDS chunk (4 MFMA K steps)
constexpr int kRows64 = 32;                // rows per wave (two 16-row tiles)
constexpr int kTile64 = 4 * kRows64;       // rows per workgroup
constexpr int kSlots64 = 512;              // resident workgroups: 2 per CU

// q = n / d, r = n % d for 0 <= n < 2^31 (double reciprocal, one correction)
__device__ __forceinline__ void divmod64(int n, int d, double inv, int& q, int& r) {
    q = (int)((double)n * inv);
    r = n - q * d;
    const bool lo = r < 0, hi = r >= d;
    q += hi ? 1 : (lo ? -1 : 0);
    r += hi ? -d : (lo ? d : 0);
}

// dz / r^3 in cell units, as the FP64 single-model kernels form it: the row
// term dy^2 + dz^2 floored by kEpsD, then rsqrt(a^2 + R) cubed.
__device__ __forceinline__ double g64(double a, double dy, double dz) {
    const double v = rsqrt(fma(a, a, fma(dy, dy, fma(dz, dz, kEpsD))));
    return dz * (v * v * v);
}

__device__ __forceinline__ f64x4 mfma64(double a, double b, f64x4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Chunk k0 .. k0 + 15 of X [B][K] (model-major) in flight to registers:
// element e = threadIdx.x + 256 u holds K value e % 16 of model e / 16. The
// loads are unconditional (entries outside the chunk or beyond B read X[0])
// and nothing reads their registers until chunk_store, after the next chunk's
// MFMAs, so the whole chunk is in flight at once; ok has bit u set where the
// entry is real (otherwise it is staged as 0).
template <int NT>
struct Chunk64 {
    double v[NT];
    unsigned ok;

    __device__ __forceinline__ void load(const double* __restrict__ X, int B, int64_t K,
                                         int64_t k0, int64_t k_end) {
        ok = 0u;
#pragma unroll
        for (int u = 0; u < NT; ++u) {
            const int e = threadIdx.x + 256 * u;
            const int c = e & 15, b = e >> 4;
            const int64_t k = k0 + c;
            const bool in = b < B && k < k_end;
            ok |= in ? 1u << u : 0u;
            v[u] = X[in ? (size_t)b * K + k : 0];
        }
    }

    // Writes the chunk to LDS [16][Bp + 1]: FWD applies the threshold (NaN or
    // |rho| <= thr: 0); otherwise the values are multiplied by mul (the
    // adjoint takes the plan's scale in here, so that its accumulators go to
    // memory untouched and stay in the MFMA registers).
    template <bool FWD>
    __device__ __forceinline__ void store(double* lp, double thr, double mul) const {
        constexpr int LD = 16 * NT + 1;   // keeps 16 consecutive K values on distinct banks
#pragma unroll
        for (int u = 0; u < NT; ++u) {
            const int e = threadIdx.x + 256 * u;
            double x = (ok >> u) & 1u ? v[u] : 0.0;
            if (FWD) {
                if (!(fabs(x) > thr)) x = 0.0;
            } else {
                x *= mul;
            }
            lp[(e & 15) * LD + (e >> 4)] = x;
        }
    }
};

template <int NT> constexpr int lds_ld64() { return 16 * NT + 1; }

// Forward. grid = (m_pad / 128, KS); block = 256.
template <int NT>
__global__ __launch_bounds__(256) void batched64_fwd_kernel(
    const double4* __restrict__ st, const double* __restrict__ P, int B, int64_t N, int nx,
    int ny, double inv_nx, double inv_ny, double sy, double sz, double thr,
    int64_t cells_per_slice, int m_pad, int bp, double* __restrict__ part) {
    constexpr int LD = lds_ld64<NT>();
    __shared__ double lp[kChunk64 * LD];
    const int model0 = blockIdx.z * 16 * NT;   // this workgroup's model group
    P += (size_t)model0 * N;
    B -= model0;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int row0 = blockIdx.x * kTile64 + wave * kRows64;
    const double4 p0 = st[row0 + li], p1 = st[row0 + 16 + li];
    const int64_t n_begin = (int64_t)blockIdx.y * cells_per_slice;
    const int64_t n_end = min(N, n_begin + cells_per_slice);
    f64x4 acc[2][NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[0][t] = acc[1][t] = f64x4{0.0, 0.0, 0.0, 0.0};
    __shared__ double4 lc[kChunk64];   // the chunk's cell centres (cell units)
    Chunk64<NT> pre;
    pre.load(P, B, N, n_begin, n_end);
    for (int64_t n0 = n_begin; n0 < n_end; n0 += kChunk64) {
        __syncthreads();   // every wave is done with the previous chunk
        pre.template store<true>(lp, thr, 1.0);
        if (threadIdx.x < kChunk64) {
            int rw, i, j, k;
            divmod64((int)(n0 + threadIdx.x), nx, inv_nx, rw, i);
            divmod64(rw, ny, inv_ny, k, j);
            lc[threadIdx.x] = double4{(double)i, (double)j * sy, (double)k * sz, 0.0};
        }
        __syncthreads();
        if (n0 + kChunk64 < n_end) pre.load(P, B, N, n0 + kChunk64, n_end);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int kq = 4 * q + lk;   // this lane's K value in the chunk
            const double4 c = lc[kq];
            const double a0 = g64(p0.x - c.x, p0.y - c.y, p0.z - c.z);
            const double a1 = g64(p1.x - c.x, p1.y - c.y, p1.z - c.z);
            const double* lrow = lp + kq * LD + li;
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const double b = lrow[16 * t];
                acc[0][t] = mfma64(a0, b, acc[0][t]);
                acc[1][t] = mfma64(a1, b, acc[1][t]);
            }
        }
    }
    double* out = part + (size_t)blockIdx.y * m_pad * bp + model0;
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                out[(size_t)(row0 + 16 * s + 4 * r + lk) * bp + 16 * t + li] = acc[s][t][r];
}

// Adjoint. grid = (n_pad / 128, KS); block = 256. The residuals are staged
// times scale. One K-slice writes X directly; several write partials for the
// reduce (which then sums them with scale 1).
template <int NT>
__global__ __launch_bounds__(256) void batched64_adj_kernel(
    const double4* __restrict__ st, const double* __restrict__ R, int B, int M, int m_pad,
    int64_t N, int nx, int ny, double inv_nx, double inv_ny, double sy, double sz,
    int st_per_slice, double scale, double* __restrict__ X, double* __restrict__ part,
    int64_t n_pad, int bp) {
    constexpr int LD = lds_ld64<NT>();
    __shared__ double lp[kChunk64 * LD];
    const int model0 = blockIdx.z * 16 * NT;   // this workgroup's vector group
    R += (size_t)model0 * M;
    B -= model0;
    __shared__ double4 ls[kChunk64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int64_t row0 = (int64_t)blockIdx.x * kTile64 + wave * kRows64;
    double cx[2], cy[2], cz[2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {   // this lane's two cells (rows of A)
        int rw, i, j, k;
        divmod64((int)(row0 + 16 * s + li), nx, inv_nx, rw, i);
        divmod64(rw, ny, inv_ny, k, j);
        cx[s] = (double)i;
        cy[s] = (double)j * sy;
        cz[s] = (double)k * sz;
    }
    const int m_begin = blockIdx.y * st_per_slice;
    const int m_end = min(m_pad, m_begin + st_per_slice);
    f64x4 acc[2][NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[0][t] = acc[1][t] = f64x4{0.0, 0.0, 0.0, 0.0};
    Chunk64<NT> pre;
    pre.load(R, B, M, m_begin, min(m_end, M));
    // the chunk's 16 station records, in flight to threads 0 .. 15 likewise
    double4 pst = st[m_begin + (threadIdx.x & 15)];
    for (int m0 = m_begin; m0 < m_end; m0 += kChunk64) {
        __syncthreads();
        pre.template store<false>(lp, 0.0, scale);
        if (threadIdx.x < kChunk64) ls[threadIdx.x] = pst;
        __syncthreads();
        if (m0 + kChunk64 < m_end) {
            pre.load(R, B, M, m0 + kChunk64, min(m_end, M));
            pst = st[m0 + kChunk64 + (threadIdx.x & 15)];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int kq = 4 * q + lk;   // this lane's station in the chunk
            const double4 p = ls[kq];
            const double a0 = g64(p.x - cx[0], p.y - cy[0], p.z - cz[0]);
            const double a1 = g64(p.x - cx[1], p.y - cy[1], p.z - cz[1]);
            const double* lrow = lp + kq * LD + li;
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const double b = lrow[16 * t];
                acc[0][t] = mfma64(a0, b, acc[0][t]);
                acc[1][t] = mfma64(a1, b, acc[1][t]);
            }
        }
    }
    if (part) {
        double* out = part + (size_t)blockIdx.y * n_pad * bp + model0;
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int t = 0; t < NT; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    out[(size_t)(row0 + 16 * s + 4 * r + lk) * bp + 16 * t + li] = acc[s][t][r];
    } else {
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int t = 0; t < NT; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int64_t n = row0 + 16 * s + 4 * r + lk;
                    const int b = 16 * t + li;
                    if (n < N && b < B) X[(size_t)(model0 + b) * N + n] = acc[s][t][r];
                }
    }
}

// out[b][row] = scale * sum over KS slices of part[slice][row][b], rows < n_rows.
__global__ __launch_bounds__(256) void batched64_reduce_kernel(const double* __restrict__ part,
                                                               int ks, int64_t rows_pad, int bp,
                                                               int B, int64_t n_rows,
                                                               double scale,
                                                               double* __restrict__ out) {
    const int64_t total = n_rows * bp;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * 256) {
        const int b = (int)(e % bp);
        const int64_t row = e / bp;
        if (b >= B) continue;
        double s = 0.0;
        for (int k = 0; k < ks; ++k) s += part[((size_t)k * rows_pad + row) * bp + b];
        out[(size_t)b * n_rows + row] = scale * s;
    }
}

int tiles_of(int B) {   // 16-column model tiles over the batch: 1, 2, 4, 8 or 16
    int nt = 1;
    while (16 * nt < B) nt *= 2;
    return nt;
}

// Model tiles per workgroup: at most 8 (128 models), which keeps a wave's
// accumulators to 128 registers and two workgroups per CU; a batch of 256
// runs as two model groups (grid z), each forming its own G.
int group_tiles(int B) { return std::min(tiles_of(B), 8); }

// K-slices for `wgs` workgroups per slice: the fewest (at most max_ks) whose
// workgroups fill the resident slots in whole rounds to within 5% -- a last
// round that is mostly empty would cost a full workgroup's time -- else the
// best fill found.
int pick_slices(int64_t wgs, int64_t max_ks) {
    max_ks = std::max<int64_t>(1, std::min<int64_t>(max_ks, 64));
    int best = 1;
    double best_fill = 0.0;
    for (int ks = 1; ks <= max_ks; ++ks) {
        const double rounds = (double)(wgs * ks) / kSlots64;
        const double fill = rounds / std::ceil(rounds);
        if (fill >= 0.95) return ks;
        if (fill > best_fill + 1e-9) {
            best_fill = fill;
            best = ks;
        }
    }
    return best;
}

}  // namespace

size_t batched64_fwd_workspace(const PmShape& s, int B, int* ks_out) {
    const int64_t wgs = s.m_pad / kTile64 * (tiles_of(B) / group_tiles(B));
    const int64_t chunks = (s.n_cells + kChunk64 - 1) / kChunk64;
    const int ks = pick_slices(wgs, chunks);
    if (ks_out) *ks_out = ks;
    return sizeof(double) * (size_t)ks * (size_t)s.m_pad * (size_t)(16 * tiles_of(B));
}

size_t batched64_adj_workspace(const PmShape& s, int B, int* ks_out) {
    const int64_t n_pad = (s.n_cells + kTile64 - 1) / kTile64 * kTile64;
    const int64_t wgs = n_pad / kTile64 * (tiles_of(B) / group_tiles(B));
    const int64_t chunks = s.m_pad / kChunk64;
    const int ks = pick_slices(wgs, chunks);
    if (ks_out) *ks_out = ks;
    return ks == 1 ? 0 : sizeof(double) * (size_t)ks * (size_t)n_pad * (size_t)(16 * tiles_of(B));
}

int batched64_fwd_launch(const PmShape& s, int B, const double4* st, const double* P,
                         double thr, double sy, double sz, double scale, double* ws, double* Y,
                         hipStream_t stream) {
    if (B <= 0 || s.m == 0) return ZTM_OK;
    if (s.n_cells >= ((int64_t)1 << 31) - kChunk64)
        return fail(ZTM_EINVAL, "the FP64 ensemble needs fewer than 2^31 cells");
    if (s.m_pad % kTile64 != 0)
        return fail(ZTM_EINVAL, "padded station count %d is not a multiple of %d", s.m_pad,
                    kTile64);
    int ks = 1;
    batched64_fwd_workspace(s, B, &ks);
    const int64_t chunks = (s.n_cells + kChunk64 - 1) / kChunk64;
    const int64_t cps = (chunks + ks - 1) / ks * kChunk64;   // whole chunks per slice
    ks = (int)((s.n_cells + cps - 1) / cps);
    const int nt = tiles_of(B), gt = group_tiles(B);
    const dim3 grid(s.m_pad / kTile64, ks, nt / gt);
    const double inx = 1.0 / s.nx, iny = 1.0 / s.ny;
    {
        ProfScope prof("batched64_fwd", stream);
#define ZTM_B64F(NT)                                                                        \
    hipLaunchKernelGGL((batched64_fwd_kernel<NT>), grid, dim3(256), 0, stream, st, P, B,    \
                       (int64_t)s.n_cells, s.nx, s.ny, inx, iny, sy, sz, thr, cps, s.m_pad,  \
                       16 * nt, ws)
        switch (gt) {
            case 1: ZTM_B64F(1); break;
            case 2: ZTM_B64F(2); break;
            case 4: ZTM_B64F(4); break;
            default: ZTM_B64F(8); break;
        }
#undef ZTM_B64F
        ZTM_HIP(hipGetLastError());
    }
    {
        ProfScope prof("batched64_reduce", stream);
        const int64_t total = (int64_t)s.m * 16 * nt;
        const unsigned blocks = (unsigned)std::min<int64_t>((total + 255) / 256, 1 << 16);
        hipLaunchKernelGGL(batched64_reduce_kernel, dim3(blocks), dim3(256), 0, stream, ws, ks,
                           (int64_t)s.m_pad, 16 * nt, B, (int64_t)s.m, scale, Y);
        ZTM_HIP(hipGetLastError());
    }
    return ZTM_OK;
}

int batched64_adj_launch(const PmShape& s, int B, const double4* st, const double* R, double sy,
                         double sz, double scale, double* ws, double* X, hipStream_t stream) {
    if (B <= 0 || s.n_cells == 0) return ZTM_OK;
    if (s.n_cells >= ((int64_t)1 << 31) - kTile64)
        return fail(ZTM_EINVAL, "the FP64 ensemble needs fewer than 2^31 cells");
    if (s.m_pad % kChunk64 != 0)
        return fail(ZTM_EINVAL, "padded station count %d is not a multiple of %d", s.m_pad,
                    kChunk64);
    int ks = 1;
    batched64_adj_workspace(s, B, &ks);
    const int chunks = s.m_pad / kChunk64;
    const int sps = (chunks + ks - 1) / ks * kChunk64;   // whole chunks per slice
    ks = (s.m_pad + sps - 1) / sps;
    const int nt = tiles_of(B), gt = group_tiles(B);
    const int64_t n_pad = (s.n_cells + kTile64 - 1) / kTile64 * kTile64;
    const dim3 grid((unsigned)(n_pad / kTile64), ks, nt / gt);
    const double inx = 1.0 / s.nx, iny = 1.0 / s.ny;
    double* part = ks > 1 ? ws : nullptr;
    {
        ProfScope prof("batched64_adj", stream);
#define ZTM_B64A(NT)                                                                         \
    hipLaunchKernelGGL((batched64_adj_kernel<NT>), grid, dim3(256), 0, stream, st, R, B, s.m, \
                       s.m_pad, (int64_t)s.n_cells, s.nx, s.ny, inx, iny, sy, sz, sps, scale, \
                       X, part, n_pad, 16 * nt)
        switch (gt) {
            case 1: ZTM_B64A(1); break;
            case 2: ZTM_B64A(2); break;
            case 4: ZTM_B64A(4); break;
            default: ZTM_B64A(8); break;
        }
#undef ZTM_B64A
        ZTM_HIP(hipGetLastError());
    }
    if (ks > 1) {
        ProfScope prof("batched64_reduce", stream);
        const int64_t total = (int64_t)s.n_cells * 16 * nt;
        const unsigned blocks = (unsigned)std::min<int64_t>((total + 255) / 256, 1 << 16);
        hipLaunchKernelGGL(batched64_reduce_kernel, dim3(blocks), dim3(256), 0, stream, ws, ks,
                           n_pad, 16 * nt, B, (int64_t)s.n_cells, 1.0, X);
        ZTM_HIP(hipGetLastError());
    }
    return ZTM_OK;
}

}  // namespace ztm
