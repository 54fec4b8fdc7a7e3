// Point-mass gravity forward (g = G rho) and adjoint (G^T r) kernels for
// gfx950, FP32 and FP64.
//
// Cell formula of GFM_GPU.cu:22 and TFM.cpp:98: a cell of density rho at
// distance r below/above a station adds  kappa * rho * V * dz / r^3  to g_z.
// The reference scatters one thread per cell over all stations with a float
// atomicAdd per pair (GFM_GPU.cu:17-22). Here both directions are gathers:
//
//   forward  (pm_fwd): a workgroup owns 64*S stations (S per lane, kept in
//     registers) and a slice of grid rows (fixed j,k); the slice's densities
//     are staged in LDS once (thresholded, GFM_GPU.cu:19) and read back as
//     wave-uniform broadcasts. The four waves split the slice's rows and their
//     partial sums are reduced across wavefronts through LDS at the end. Per
//     row, dy^2+dz^2 is computed once per station, so a pair costs
//     sub, fma, rsq, 2 mul, fma.
//   adjoint  (pm_adj): a lane owns C consecutive cells of one row; station
//     tiles of 256 (coordinates + residual) are staged in LDS and each wave
//     walks a quarter of the tile; cross-wave LDS reduction at the end.
//
// Precision: coordinates are relative to the centre of cell (0,0,0) in cell
// units, so station-cell offsets are differences of small numbers (exact for
// the x offset). Inner sums run over one grid row (forward) or one 256-station
// tile (adjoint) in the working precision and are then accumulated in FP64;
// all cross-wave and cross-slice sums are FP64 in a fixed order, so results
// are bitwise reproducible run to run (no atomics).
#include "gfm_internal.h"
#include "gfm_pair.h"

#include <algorithm>

namespace ztm {

template <typename T> struct Vec4;
template <> struct Vec4<float>  { using type = float4; };
template <> struct Vec4<double> { using type = double4; };

__device__ __forceinline__ float  rsq(float x)  { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double rsq(double x) { return rsqrt(x); }
// Row term dy^2 + dz^2: clamped in FP32 (row_R32, gfm_pair.h; the pairs of a
// clamped row are corrected in FP64 afterwards), floored by kEpsD in FP64.
__device__ __forceinline__ float row_R(float dy, float dz) { return row_R32(dy, dz); }
__device__ __forceinline__ double row_R(double dy, double dz) {
    return fma(dy, dy, fma(dz, dz, kEpsD));
}

// Far-field kernel of one pair: v^3 (point mass), or with the quadrupole term
// of a non-cubic cell (FP32 prism path only; see gfm_pair.h).
// Scheduling hint for the FP32 pair loops: N transcendental (rsq) groups,
// each followed by KV other VALU instructions. Issuing each rsq among
// independent VALU work, instead of next to its own consumers, measured
// faster on gfx950 than the compiler's default order: adjoint C8/G4 5% with
// KV = 4, forward S2/G4 2-3% with KV = 6 (KV = 4 there is 4% slower).
template <int N, int KV>
__device__ __forceinline__ void interleave_trans() {
#pragma unroll
    for (int u = 0; u < N; ++u) {
        __builtin_amdgcn_sched_group_barrier(0x400, 1, 0);   // TRANS
        __builtin_amdgcn_sched_group_barrier(0x002, KV, 0);  // VALU
    }
}

template <typename T, bool QUAD>
__device__ __forceinline__ T kterm(T a, T R, T qrow, const QuadArgs& q) {
    if constexpr (sizeof(T) == sizeof(float)) {
        return pair_k32<QUAD>(a, R, qrow, q);
    } else {
        return pair_k64<QUAD>(a, R, qrow, q);
    }
}

// ---------------------------------------------------------------- forward
// grid = (m_pad / (64*S), n_slices); block = 256.
// st:   prepared stations [m_pad] (x, y, z in cell units; see gfm_internal.h)
// part: [n_slices][m_pad] FP64 partial sums  sum_n rho_n dz/r^3
// pm_fwd_block: one workgroup's work -- station tile bx, row slice by -- so
// that the single-problem kernel and the grouped (multi-problem) kernel share
// it; smem must hold max(rows_per_slice * nxp * sizeof(T), 4*S*64*8) bytes.
// Rows a wave processes together in the FP32 point-mass forward, sharing the
// per-column offset a = X - i across them (fewer for more stations per lane,
// to bound the registers); 1 for FP64 and the quadrupole kernel. Matches
// fwd_group().
template <typename T, int S, bool QUAD>
constexpr int kFwdGroup = QUAD ? 1
                          : sizeof(T) == sizeof(float) ? (S <= 1 ? 8 : S == 2 ? 4 : S == 4 ? 2 : 1)
                                                       : (S == 2 ? 4 : 1);

// F > 1 (FP32 row groups only): F density fields, field f at rho + f *
// field_stride, share every pair's r^-3; their partials go to
// part[(by * F + f) * m_pad + m] and smem holds F density tiles. Beyond 4
// fields the row group is 2 rows, which bounds the registers the per-field
// accumulators and density vectors take. Matches fields_group().
template <int F>
constexpr int kFieldsGroup = F > 4 ? 2 : 4;

template <typename T, int S, bool QUAD, int F = 1>
__device__ __forceinline__ void pm_fwd_block(
    const typename Vec4<T>::type* __restrict__ st, const T* __restrict__ rho,
    int nx, int ny, int n_rows, T sy, T sz, T thr, int rows_per_slice,
    double* __restrict__ part, int m_pad, const QuadArgs& qa, int bx, int by,
    unsigned char* smem, int64_t field_stride = 0) {
    using V4 = typename Vec4<T>::type;
    static_assert(F == 1 || (S == 2 && !QUAD && sizeof(T) == sizeof(float)),
                  "several fields need the FP32 row-group path");
    constexpr int G = F > 1 ? kFieldsGroup<F> : kFwdGroup<T, S, QUAD>;
    T* lrho = reinterpret_cast<T*>(smem);

    const int nxp = (nx + 3) & ~3;
    const int row0 = by * rows_per_slice;
    const int nr = min(rows_per_slice, n_rows - row0);
    const int nr_pad = (nr + G - 1) / G * G;   // whole row groups; extra rows are 0
    const size_t base = (size_t)row0 * (size_t)nx;

    // Stage this slice's densities, thresholded and zero-padded to nxp. A
    // thread's element t = r nxp + i steps by 256 each pass; (r, i) follow it
    // without a division per element. With row groups (G > 1) the tile is
    // laid out [row group][column quad][field][row in group][4]: the F x G
    // quads a wave reads for one column quad are contiguous, so that one
    // address and constant offsets reach them all. Without groups it is
    // [field][row][nxp].
    const int tile = nr_pad * nxp;   // one field's tile
    const int nx4 = nxp >> 2;
    {
        const int dr = 256 / nxp, di = 256 - dr * nxp;
        int r = threadIdx.x / nxp, i = threadIdx.x - r * nxp;
        for (int t = threadIdx.x; t < tile; t += 256) {
            int o = t;   // float offset of field 0's element
            if constexpr (G > 1) {
                const int rgl = r / G, g = r - rgl * G;
                o = (((rgl * nx4 + (i >> 2)) * F) * G + g) * 4 + (i & 3);
            }
#pragma unroll
            for (int f = 0; f < F; ++f) {
                T v = T(0);
                if (i < nx && r < nr) {
                    v = rho[(size_t)f * field_stride + base + (size_t)r * nx + i];
                    if (!(fabs(v) > thr)) v = T(0);   // NaN or |rho| <= thr: skipped
                }
                lrho[G > 1 ? o + f * G * 4 : f * tile + t] = v;
            }
            r += dr;
            i += di;
            if (i >= nxp) {
                i -= nxp;
                ++r;
            }
        }
    }
    __syncthreads();

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int m0 = bx * (64 * S);

    T X[S], Y[S], Z[S];
    double acc[F][S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const V4 p = st[m0 + s * 64 + lane];
        X[s] = p.x; Y[s] = p.y; Z[s] = p.z;
#pragma unroll
        for (int f = 0; f < F; ++f) acc[f][s] = 0.0;
    }

    if constexpr (G > 1) {
        // Groups of G consecutive rows per wave: per station and cell column
        // a = X - i is formed once and shared by the G rows, so a pair costs
        // r^2 = fma(a, a, R), rsq, v*v*v and the accumulate (the same
        // arithmetic as pair_k32, which the near-field corrections replicate).
        // Padded rows hold 0.
        for (int rg = wave * G; rg < nr; rg += 4 * G) {
            T R[G][S], ar[F][G][S], xm[S];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const int row = row0 + rg + g;
                const int k = row / ny;
                const int j = row - k * ny;
                const T yj = T(j) * sy;
                const T zk = T(k) * sz;
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    R[g][s] = row_R(Y[s] - yj, Z[s] - zk);
#pragma unroll
                    for (int f = 0; f < F; ++f) ar[f][g][s] = T(0);
                }
            }
#pragma unroll
            for (int s = 0; s < S; ++s) xm[s] = X[s];
            const V4* lgrp = reinterpret_cast<const V4*>(lrho) + (rg / G) * nx4 * F * G;
            for (int i4 = 0; i4 < nx4; ++i4) {
                V4 q[F][G];
                const V4* lq = lgrp + i4 * (F * G);
#pragma unroll
                for (int f = 0; f < F; ++f)
#pragma unroll
                    for (int g = 0; g < G; ++g) q[f][g] = lq[f * G + g];   // wave-uniform
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const T a0 = xm[s], a1 = xm[s] - T(1), a2 = xm[s] - T(2), a3 = xm[s] - T(3);
#pragma unroll
                    for (int g = 0; g < G; ++g) {
                        const T v0 = rsq(fma(a0, a0, R[g][s]));
                        const T v1 = rsq(fma(a1, a1, R[g][s]));
                        const T v2 = rsq(fma(a2, a2, R[g][s]));
                        const T v3 = rsq(fma(a3, a3, R[g][s]));
                        const T k0 = (v0 * v0) * v0, k1 = (v1 * v1) * v1;
                        const T k2 = (v2 * v2) * v2, k3 = (v3 * v3) * v3;
#pragma unroll
                        for (int f = 0; f < F; ++f) {
                            T t = ar[f][g][s];
                            t = fma(q[f][g].x, k0, t);
                            t = fma(q[f][g].y, k1, t);
                            t = fma(q[f][g].z, k2, t);
                            t = fma(q[f][g].w, k3, t);
                            ar[f][g][s] = t;
                        }
                    }
                    xm[s] -= T(4);
                }
                if constexpr (sizeof(T) == sizeof(float)) interleave_trans<S * G * 4, 5 + F>();
            }
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const int row = row0 + rg + g;
                const T zk = T(row / ny) * sz;
#pragma unroll
                for (int s = 0; s < S; ++s)
#pragma unroll
                    for (int f = 0; f < F; ++f)
                        acc[f][s] += (double)((Z[s] - zk) * ar[f][g][s]);
            }
        }
    } else
    for (int r = wave; r < nr; r += 4) {
        const int row = row0 + r;
        const int k = row / ny;
        const int j = row - k * ny;
        const T yj = T(j) * sy;
        const T zk = T(k) * sz;
        T R[S], dz[S], ar[S], xm[S], QR[S];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const T dy = Y[s] - yj;
            dz[s] = Z[s] - zk;
            R[s] = row_R(dy, dz[s]);
            QR[s] = QUAD ? fma(T(qa.q2y), dy * dy, T(qa.q2z) * (dz[s] * dz[s])) : T(0);
            ar[s] = T(0);
            xm[s] = X[s];
        }
        const V4* lrow = reinterpret_cast<const V4*>(lrho + r * nxp);
        for (int i4 = 0; i4 < (nxp >> 2); ++i4) {
            const V4 q = lrow[i4];   // wave-uniform LDS broadcast
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const T a0 = xm[s];
                const T a1 = xm[s] - T(1);
                const T a2 = xm[s] - T(2);
                const T a3 = xm[s] - T(3);
                const T k0 = kterm<T, QUAD>(a0, R[s], QR[s], qa);
                const T k1 = kterm<T, QUAD>(a1, R[s], QR[s], qa);
                const T k2 = kterm<T, QUAD>(a2, R[s], QR[s], qa);
                const T k3 = kterm<T, QUAD>(a3, R[s], QR[s], qa);
                T t = ar[s];
                t = fma(q.x, k0, t);
                t = fma(q.y, k1, t);
                t = fma(q.z, k2, t);
                t = fma(q.w, k3, t);
                ar[s] = t;
                xm[s] -= T(4);
            }
        }
#pragma unroll
        for (int s = 0; s < S; ++s) acc[0][s] += (double)(dz[s] * ar[s]);
    }

    // Cross-wavefront reduction (fixed order w0+w1+w2+w3).
    __syncthreads();
    double* red = reinterpret_cast<double*>(smem);
#pragma unroll
    for (int f = 0; f < F; ++f)
#pragma unroll
        for (int s = 0; s < S; ++s) red[((wave * F + f) * S + s) * 64 + lane] = acc[f][s];
    __syncthreads();
    for (int t = threadIdx.x; t < 64 * S * F; t += blockDim.x) {
        const int f = t / (64 * S), u = t - f * (64 * S);
        const double* rf = red + f * S * 64 + u;
        const double v = rf[0] + rf[F * S * 64] + rf[2 * F * S * 64] + rf[3 * F * S * 64];
        part[((size_t)by * F + f) * m_pad + m0 + u] = v;
    }
}

// F FP32 density fields through one pass (S = 2, row groups of
// kFieldsGroup<F> rows). grid = (m_pad / 128, n_slices); block = 256.
template <int F>
__global__ __launch_bounds__(256) void pm_fwd_fields_kernel(
    const float4* __restrict__ st, const float* __restrict__ rho, int64_t field_stride, int nx,
    int ny, int n_rows, float sy, float sz, float thr, int rows_per_slice,
    double* __restrict__ part, int m_pad) {
    extern __shared__ __align__(16) unsigned char smem[];
    pm_fwd_block<float, 2, false, F>(st, rho, nx, ny, n_rows, sy, sz, thr, rows_per_slice, part,
                                     m_pad, QuadArgs{}, blockIdx.x, blockIdx.y, smem,
                                     field_stride);
}

// grid = (m_pad / (64*S), n_slices); block = 256.
template <typename T, int S, bool QUAD>
__global__ __launch_bounds__(256) void pm_fwd_kernel(
    const typename Vec4<T>::type* __restrict__ st, const T* __restrict__ rho,
    int nx, int ny, int n_rows, T sy, T sz, T thr, int rows_per_slice,
    double* __restrict__ part, int m_pad, QuadArgs qa) {
    extern __shared__ __align__(16) unsigned char smem[];
    pm_fwd_block<T, S, QUAD>(st, rho, nx, ny, n_rows, sy, sz, thr, rows_per_slice, part, m_pad,
                             qa, blockIdx.x, blockIdx.y, smem);
}

// g[m] = scale * (sum_slices part[slice][m] + correction)
//   (general API, FP64 out), or gref[map[m]] += (float)(...) (reference ABI).
// kCorrPmNear: FP64 near field over the station's correction box and its
// clamped rows; kCorrPrism: the per-station correction computed beforehand.
// One wave per station: the lanes stride over the slice partials and the
// correction pairs, then a fixed butterfly sums them (deterministic; every
// lane returns the result).
// Slice k's partial of station i is part[k * stride + i].
__device__ __forceinline__ double pm_fwd_reduce_wave(const double* __restrict__ part,
                                                    int n_slices, int stride, int i, double scale,
                                                    const float* __restrict__ rho_f,
                                                    const NearArgs& na, int lane) {
    double s = 0.0;
    for (int k = lane; k < n_slices; k += 64) s += part[(size_t)k * stride + i];
    if (na.mode == kCorrPmNear) {
        const float4 pf = na.st_f[i];
        const double4 pd = na.st_d[i];
        const int4 c = na.st_c[i];
        const int i0 = max(c.x - na.rx, 0), i1 = min(c.x + na.rx, na.nx - 1);
        const int j0 = max(c.y - na.ry, 0), j1 = min(c.y + na.ry, na.ny - 1);
        const int k0 = max(c.z - na.rz, 0), k1 = min(c.z + na.rz, na.nz - 1);
        const int bi = i1 - i0 + 1, bj = j1 - j0 + 1, bk = k1 - k0 + 1;
        const int total = (bi > 0 && bj > 0 && bk > 0) ? bi * bj * bk : 0;
        for (int t = lane; t < total; t += 64) {
            const int ii = i0 + t % bi;
            const int rest = t / bi;
            const int jj = j0 + rest % bj;
            const int kk = k0 + rest / bj;
            const float w = rho_f[((size_t)kk * na.ny + jj) * na.nx + ii];
            if (!(fabsf(w) > na.thr)) continue;
            s += corr_pair<false>(kCorrPmNear, w, pf, pd, c, ii, jj, kk, na);
        }
        s += clamped_rows_corr<false>(kCorrPmNear, pf, pd, c, rho_f, na, lane, 64);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (na.mode == kCorrPrism || na.mode == kCorrPrism64) s += na.corr[i];
    return s * scale;
}

// block = 256 = 4 stations (one wave each)
__global__ __launch_bounds__(256) void pm_fwd_reduce_kernel(const double* __restrict__ part,
                                     int n_slices, int stride, int m,
                                     double scale, double* __restrict__ g,
                                     float* __restrict__ gref,
                                     const int* __restrict__ out_map,
                                     const float* __restrict__ rho_f,
                                     NearArgs na) {
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= m) return;   // uniform per wave
    const double v = pm_fwd_reduce_wave(part, n_slices, stride, i, scale, rho_f, na, lane);
    if (lane != 0) return;
    if (out_map)
        gref[out_map[i]] += (float)v;
    else
        g[i] = v;
}

// The F-field forward's reduces in one launch: grid = (ceil(m / 4), F);
// block = 256. Field f = blockIdx.y reads its partials at part + f * m_pad
// (slice stride F * m_pad), its densities at rho_f + f * field_stride for the
// near-field correction, and writes g[f * m + i]; per field the arithmetic of
// pm_fwd_reduce_kernel.
__global__ __launch_bounds__(256) void pm_fwd_fields_reduce_kernel(
    const double* __restrict__ part, int n_slices, int stride, int m, int m_pad, double scale,
    double* __restrict__ g, const float* __restrict__ rho_f, int64_t field_stride, NearArgs na) {
    const int f = blockIdx.y;
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= m) return;   // uniform per wave
    const double v = pm_fwd_reduce_wave(part + (size_t)f * m_pad, n_slices, stride, i, scale,
                                        rho_f + f * field_stride, na, lane);
    if (lane == 0) g[(size_t)f * m + i] = v;
}

// ---------------------------------------------------------------- adjoint
// Rows a lane takes together in the point-mass adjoint (same k, G
// consecutive j): per station, dz, q = r dz and the column offsets a = X - i
// are formed once for all G rows. FP64 uses half the FP32 group at C = 8 (its
// accumulators take twice the registers); the quadrupole kernel uses 1.
// Matches adj_group().
template <typename T, int C, bool QUAD>
constexpr int kAdjGroup = QUAD ? 1
                          : sizeof(T) == sizeof(float) ? (C >= 16 ? 2 : 4)
                                                       : (C >= 8 ? 2 : 4);

// Rows of the row-term table of the FP32 point-mass adjoint (see pm_adj_block).
constexpr int kAdjRows = 16;

// grid = (ceil(n_segs/64), n_splits); block = 256; a segment is C
// consecutive cells of G consecutive rows of one k plane (the last segment of
// a row and the last row group of a plane may be ragged); segments run
// i fastest, then the row group, then k.
// pm_adj_block: one workgroup's work -- segment block bx, station split by;
// smem must hold adj_lds() bytes.
//
// Row-term table (FP32 point mass, G > 1): when the workgroup's segments lie
// in one k plane and span at most kAdjRows rows, the station staging also
// forms q = r dz and dz^2 once per station, and R = max(dy^2 + dz^2, 1/4) once
// per (station, row) into an LDS table, instead of once per lane. Lanes then
// read their G row terms with one LDS load. The values are those of the
// general path, operation for operation, so both give identical results.
template <typename T, int C, bool QUAD>
__device__ __forceinline__ void pm_adj_block(
    const typename Vec4<T>::type* __restrict__ st, const T* __restrict__ r,
    int m, int nx, int ny, T sy, T sz, int segs_per_row, int64_t n_segs,
    int st_per_split, int m_pad, double* __restrict__ part,
    T* __restrict__ out, double scale, int64_t n_cells, int direct, const QuadArgs& qa,
    int bx, int by, unsigned char* smem) {
    using V4 = typename Vec4<T>::type;
    constexpr int G = kAdjGroup<T, C, QUAD>;
    V4* tile = reinterpret_cast<V4*>(smem);

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int rgroups = (ny + G - 1) / G;
    const int64_t seg = (int64_t)bx * 64 + lane;
    const int64_t segc = seg < n_segs ? seg : n_segs - 1;   // past the end: computed, not written
    const int64_t rowg = segc / segs_per_row;
    const int i0 = (int)(segc - rowg * segs_per_row) * C;
    const int k = (int)(rowg / rgroups);
    const int j0 = (int)(rowg - (int64_t)k * rgroups) * G;
    const T zk = T(k) * sz;
    const T xi0 = T(i0);
    T yj[G];
#pragma unroll
    for (int g = 0; g < G; ++g) yj[g] = T(j0 + g) * sy;

    double accd[G][C];
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int c = 0; c < C; ++c) accd[g][c] = 0.0;

    // workgroup-uniform choice of the row-term table
    constexpr bool kTable = sizeof(T) == sizeof(float) && !QUAD && G > 1;
    bool table = false;
    int rowoff = 0;
    T yfirst = T(0);
    if constexpr (kTable) {
        const int64_t s0 = (int64_t)bx * 64;
        const int64_t s1 = min(s0 + 63, n_segs - 1);
        const int64_t rg0 = s0 / segs_per_row, rg1 = s1 / segs_per_row;
        const int kw = (int)(rg0 / rgroups);
        table = kw == (int)(rg1 / rgroups) && (rg1 - rg0 + 1) * G <= kAdjRows;
        const int jfirst = (int)(rg0 - (int64_t)kw * rgroups) * G;
        rowoff = j0 - jfirst;   // in [0, kAdjRows - G] when table
        yfirst = T(jfirst);
    }
    T* rt = reinterpret_cast<T*>(smem + 256 * sizeof(V4));   // [256][kAdjRows]

    // the pairs of one station with this lane's G x C cells
    auto pairs = [&](T (&acc)[G][C], T q, T xm, const T (&R)[G], const T (&qrow)[G]) {
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const T a = xm - T(c);
#pragma unroll
            for (int g = 0; g < G; ++g)
                acc[g][c] = fma(q, kterm<T, QUAD>(a, R[g], qrow[g], qa), acc[g][c]);
        }
        if constexpr (sizeof(T) == sizeof(float) && !QUAD) interleave_trans<C * G, 4>();
    };

    const int mb = by * st_per_split;
    const int me = min(mb + st_per_split, m_pad);
    for (int t0 = mb; t0 < me; t0 += 256) {
        __syncthreads();
        {
            const int mm = t0 + threadIdx.x;
            V4 p = st[mm];
            p.w = mm < m ? r[mm] : T(0);
            if (kTable && table) {
                const T dz = p.z - zk;
                p = V4{p.x, p.w * dz, dz * dz, p.y};   // (x, q, dz^2, y)
            }
            tile[threadIdx.x] = p;
        }
        __syncthreads();
        T acc[G][C];
#pragma unroll
        for (int g = 0; g < G; ++g)
#pragma unroll
            for (int c = 0; c < C; ++c) acc[g][c] = T(0);
        const int sb = wave * 64;
        if (kTable && table) {
            {   // R for kAdjRows rows from the workgroup's first row: thread ->
                // row t % kAdjRows, stations t / kAdjRows + (256 / kAdjRows) u
                const int rho = threadIdx.x % kAdjRows;
                const T yr = (yfirst + T(rho)) * sy;
#pragma unroll 4
                for (int u = 0; u < kAdjRows; ++u) {
                    const int s = threadIdx.x / kAdjRows + (256 / kAdjRows) * u;
                    const V4 p = tile[s];
                    const T dy = p.w - yr;
                    rt[s * kAdjRows + rho] = fmax(fma(dy, dy, p.z), T(kRmin2F));
                }
            }
            __syncthreads();
            const T qrow[G] = {};
            for (int s = sb; s < sb + 64; ++s) {
                const V4 p = tile[s];    // wave-uniform LDS broadcast
                T R[G];
#pragma unroll
                for (int g = 0; g < G; ++g) R[g] = rt[s * kAdjRows + rowoff + g];
                pairs(acc, p.y, p.x - xi0, R, qrow);
            }
        } else {
            for (int s = sb; s < sb + 64; ++s) {
                const V4 p = tile[s];    // wave-uniform LDS broadcast
                const T dz = p.z - zk;
                const T q = p.w * dz;
                const T xm = p.x - xi0;
                T R[G], qrow[G];
#pragma unroll
                for (int g = 0; g < G; ++g) {
                    const T dy = p.y - yj[g];
                    R[g] = row_R(dy, dz);
                    qrow[g] = QUAD ? fma(T(qa.q2y), dy * dy, T(qa.q2z) * (dz * dz)) : T(0);
                }
                pairs(acc, q, xm, R, qrow);
            }
        }
#pragma unroll
        for (int g = 0; g < G; ++g)
#pragma unroll
            for (int c = 0; c < C; ++c) accd[g][c] += (double)acc[g][c];
    }

    // Cross-wavefront reduction (fixed order), one row of the group at a time.
    double* red = reinterpret_cast<double*>(smem);
#pragma unroll
    for (int g = 0; g < G; ++g) {
        __syncthreads();
#pragma unroll
        for (int c = 0; c < C; ++c) red[(wave * C + c) * 64 + lane] = accd[g][c];
        __syncthreads();
        for (int e = threadIdx.x; e < 64 * C; e += blockDim.x) {
            const int c = e >> 6;
            const int l = e & 63;
            const double v = red[c * 64 + l] + red[(C + c) * 64 + l] +
                             red[(2 * C + c) * 64 + l] + red[(3 * C + c) * 64 + l];
            const int64_t sg = (int64_t)bx * 64 + l;
            if (sg >= n_segs) continue;
            const int64_t rg = sg / segs_per_row;
            const int ii = (int)(sg - rg * segs_per_row) * C + c;
            const int kk = (int)(rg / rgroups);
            const int jj = (int)(rg - (int64_t)kk * rgroups) * G + g;
            if (ii >= nx || jj >= ny) continue;
            const int64_t cell = ((int64_t)kk * ny + jj) * nx + ii;
            if (direct) {
                out[cell] = (T)(scale * v);
            } else {
                part[(size_t)by * (size_t)n_cells + cell] = v;
            }
        }
    }
}

// grid = (ceil(n_segs/64), n_splits); block = 256.
template <typename T, int C, bool QUAD>
__global__ __launch_bounds__(256) void pm_adj_kernel(
    const typename Vec4<T>::type* __restrict__ st, const T* __restrict__ r,
    int m, int nx, int ny, T sy, T sz, int segs_per_row, int64_t n_segs,
    int st_per_split, int m_pad, double* __restrict__ part,
    T* __restrict__ out, double scale, int64_t n_cells, int direct, QuadArgs qa) {
    extern __shared__ __align__(16) unsigned char smem[];
    pm_adj_block<T, C, QUAD>(st, r, m, nx, ny, sy, sz, segs_per_row, n_segs, st_per_split, m_pad,
                             part, out, scale, n_cells, direct, qa, blockIdx.x, blockIdx.y, smem);
}

// One half-wave (32 lanes) per cell of the near-field layers: out[n] =
// scale * (sum over splits of part[split][n] + correction). The lanes stride
// over the split partials, over the station columns within (rx, ry) of the
// cell's column (stations of a column in CSR order) and over the row keys of
// stations with clamped FP32 row terms; a fixed butterfly within the half
// then sums them, so the result is deterministic. Every lane of the half
// returns it. MODE (kCorrPmNear or kCorrPrism) is na.mode fixed at compile
// time, which keeps the FP64 prism formula out of the point-mass kernel's
// registers.
constexpr int kNearLanes = 32;   // lanes per near-field cell

template <bool QUAD, int MODE>
__device__ __forceinline__ double pm_adj_reduce_wave(const double* __restrict__ part,
                                                     int n_splits, int64_t n_cells, int64_t n,
                                                     const float* __restrict__ r_f,
                                                     const NearArgs& na, int lane) {
    double s = 0.0;
    for (int k = lane; k < n_splits; k += kNearLanes) s += part[(size_t)k * n_cells + n];
    const int ii = (int)(n % na.nx);
    const int64_t rw = n / na.nx;
    const int jj = (int)(rw % na.ny);
    const int kk = (int)(rw / na.ny);
    // (a) stations whose correction box holds this cell: via the column index
    if (kk >= na.kmin - na.rz && kk <= na.kmax + na.rz) {
        const int wcols = na.nx + 2 * na.cmx;
        const int bw = 2 * na.rx + 1;
        const int total = bw * (2 * na.ry + 1);
        for (int t = lane; t < total; t += kNearLanes) {
            const int ci = ii - na.rx + t % bw;
            const int cj = jj - na.ry + t / bw;
            if (cj < -na.cmy || cj >= na.ny + na.cmy) continue;
            if (ci < -na.cmx || ci >= na.nx + na.cmx) continue;
            const int col = (cj + na.cmy) * wcols + (ci + na.cmx);
            for (int q = na.col_start[col]; q < na.col_start[col + 1]; ++q) {
                const int m = na.col_st[q];
                const int4 c = na.st_c[m];
                if (c.z < kk - na.rz || c.z > kk + na.rz) continue;
                s += corr_pair<QUAD>(MODE, r_f[m], na.st_f[m], na.st_d[m], c, ii, jj, kk, na);
            }
        }
    }
    // (b) stations whose FP32 row term for this cell's row was clamped,
    //     outside (a): via the row-key index
    if (na.row_start && kk >= na.rkmin - na.kr && kk <= na.rkmax + na.kr) {
        const int wkey = na.ny + 2 * na.rmy;
        const int bj = 2 * na.jr + 1;
        const int total = bj * (2 * na.kr + 1);
        for (int t = lane; t < total; t += kNearLanes) {
            const int cj = jj - na.jr + t % bj;
            const int ck = kk - na.kr + t / bj;
            if (ck < -na.rmz || ck >= na.nz + na.rmz) continue;
            if (cj < -na.rmy || cj >= na.ny + na.rmy) continue;
            const int key = (ck + na.rmz) * wkey + (cj + na.rmy);
            for (int q = na.row_start[key]; q < na.row_start[key + 1]; ++q) {
                const int m = na.row_st[q];
                const int4 c = na.st_c[m];
                if (abs(ii - c.x) <= na.rx && abs(jj - c.y) <= na.ry && abs(kk - c.z) <= na.rz)
                    continue;   // done in (a)
                const float4 pf = na.st_f[m];
                if (!row_clamped(pf.y - (float)jj * na.syf, pf.z - (float)kk * na.szf)) continue;
                s += corr_pair<QUAD>(MODE, r_f[m], pf, na.st_d[m], c, ii, jj, kk, na);
            }
        }
    }
#pragma unroll
    for (int off = kNearLanes / 2; off > 0; off >>= 1) s += __shfl_xor(s, off);
    return s;
}

// ZTM_FP64_ZONED (kCorrPrism64): the same walk over the stations whose
// correction box holds the cell, with FP64 residuals and the FP64
// correction (no clamped rows in FP64).
template <bool QUAD>
__device__ __forceinline__ double pm_adj_reduce_wave64(const double* __restrict__ part,
                                                       int n_splits, int64_t n_cells, int64_t n,
                                                       const double* __restrict__ r_d,
                                                       const NearArgs& na, int lane) {
    double s = 0.0;
    for (int k = lane; k < n_splits; k += kNearLanes) s += part[(size_t)k * n_cells + n];
    const int ii = (int)(n % na.nx);
    const int64_t rw = n / na.nx;
    const int jj = (int)(rw % na.ny);
    const int kk = (int)(rw / na.ny);
    if (kk >= na.kmin - na.rz && kk <= na.kmax + na.rz) {
        const int wcols = na.nx + 2 * na.cmx;
        const int bw = 2 * na.rx + 1;
        const int total = bw * (2 * na.ry + 1);
        for (int t = lane; t < total; t += kNearLanes) {
            const int ci = ii - na.rx + t % bw;
            const int cj = jj - na.ry + t / bw;
            if (cj < -na.cmy || cj >= na.ny + na.cmy) continue;
            if (ci < -na.cmx || ci >= na.nx + na.cmx) continue;
            const int col = (cj + na.cmy) * wcols + (ci + na.cmx);
            for (int q = na.col_start[col]; q < na.col_start[col + 1]; ++q) {
                const int m = na.col_st[q];
                const int4 c = na.st_c[m];
                if (c.z < kk - na.rz || c.z > kk + na.rz) continue;
                if (abs(ii - c.x) > na.rx || abs(jj - c.y) > na.ry) continue;
                s += corr_pair64<QUAD>(r_d[m], na.st_d[m], c, ii, jj, kk, na);
            }
        }
    }
#pragma unroll
    for (int off = kNearLanes / 2; off > 0; off >>= 1) s += __shfl_xor(s, off);
    return s;
}

template <bool QUAD>
__global__ __launch_bounds__(256) void pm_adj_reduce64_kernel(const double* __restrict__ part,
                                                              int n_splits, int64_t n_cells,
                                                              int64_t n0, int64_t count,
                                                              double scale,
                                                              double* __restrict__ out,
                                                              const double* __restrict__ r_d,
                                                              NearArgs na) {
    const int lane = threadIdx.x % kNearLanes;
    constexpr int kPerBlock = 256 / kNearLanes;
    for (int64_t t = (int64_t)blockIdx.x * kPerBlock + threadIdx.x / kNearLanes; t < count;
         t += (int64_t)gridDim.x * kPerBlock) {   // uniform per half-wave
        const double v = pm_adj_reduce_wave64<QUAD>(part, n_splits, n_cells, n0 + t, r_d, na,
                                                    lane);
        if (lane == 0) out[n0 + t] = scale * v;
    }
}

// Cells outside the near-field layers [klo, khi]: the split-K sum only. The
// near-field walk is kept out of this kernel so that it runs at full
// occupancy; pm_adj_reduce_kernel does the layers klo..khi.
template <typename T>
__device__ __forceinline__ void pm_adj_sum_one(const double* __restrict__ part, int n_splits,
                                               int64_t n_cells, int64_t n, double scale,
                                               T* __restrict__ out) {
    double s = 0.0;
    for (int k = 0; k < n_splits; ++k) s += part[(size_t)k * n_cells + n];
    out[n] = (T)(scale * s);
}

template <typename T>
__global__ __launch_bounds__(256) void pm_adj_sum_kernel(const double* __restrict__ part,
                                                         int n_splits, int64_t n_cells,
                                                         int64_t plane, int klo, int khi,
                                                         double scale, T* __restrict__ out) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= n_cells) return;
    const int64_t kk = n / plane;
    if (kk >= klo && kk <= khi) return;
    pm_adj_sum_one<T>(part, n_splits, n_cells, n, scale, out);
}

// Cells n0 .. n0+count-1 (the near-field layers), one half-wave each (block
// = 256 = 8 cells): sum plus corrections.
template <bool QUAD, int MODE>
__global__ __launch_bounds__(256) void pm_adj_reduce_kernel(const double* __restrict__ part,
                                     int n_splits, int64_t n_cells, int64_t n0, int64_t count,
                                     double scale, float* __restrict__ out,
                                     const float* __restrict__ r_f, NearArgs na) {
    const int lane = threadIdx.x % kNearLanes;
    constexpr int kPerBlock = 256 / kNearLanes;
    for (int64_t t = (int64_t)blockIdx.x * kPerBlock + threadIdx.x / kNearLanes; t < count;
         t += (int64_t)gridDim.x * kPerBlock) {   // uniform per half-wave
        const double v = pm_adj_reduce_wave<QUAD, MODE>(part, n_splits, n_cells, n0 + t, r_f, na,
                                                        lane);
        if (lane == 0) out[n0 + t] = (float)(scale * v);
    }
}


// ------------------------------------------------------ grouped launches
// Index of the problem owning item `b` in a prefix array start[0..np].
__device__ __forceinline__ int find_problem(const int64_t* __restrict__ start, int np, int64_t b) {
    int lo = 0, hi = np - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= b) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Next ticket of a persistent workgroup, uniform across the workgroup.
__device__ __forceinline__ int next_ticket(unsigned* ticket, int* slot) {
    if (threadIdx.x == 0) *slot = (int)atomicAdd(ticket, 1u);
    __syncthreads();
    const int t = __builtin_amdgcn_readfirstlane(*slot);
    __syncthreads();   // the slot is rewritten only after every lane has read it
    return t;
}

// Persistent forward: each workgroup takes tickets (one ticket = one tile of
// 64*S stations x one row slice of one problem) until the list is exhausted.
// Problems are ordered by decreasing block cost, so the cheap blocks fill the
// tail. MAXS bounds the station-per-lane variants compiled in (and so the
// registers the kernel needs).
template <int MAXS>
__global__ __launch_bounds__(256) void pm_fwd_grouped_kernel(
    const GFwdDesc* __restrict__ d, const int64_t* __restrict__ blk_start, int np, int total,
    const float* __restrict__ rho, unsigned* __restrict__ ticket) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int slot;
    const QuadArgs qa{};
    for (;;) {
        const int b = next_ticket(ticket, &slot);
        if (b >= total) break;
        const int p = find_problem(blk_start, np, b);
        const GFwdDesc& q = d[p];
        const int local = b - (int)blk_start[p];
        const int bx = local % q.tiles_x, by = local / q.tiles_x;
        const float* rp = rho + q.cell_off;
        if (MAXS >= 8 && q.S == 8)
            pm_fwd_block<float, (MAXS >= 8 ? 8 : 4), false>(q.st, rp, q.nx, q.ny, q.n_rows, q.sy, q.sz,
                                                         q.thr, q.rps, q.part, q.m_pad, qa, bx, by, smem);
        else if (q.S == 2)
            pm_fwd_block<float, 2, false>(q.st, rp, q.nx, q.ny, q.n_rows, q.sy, q.sz, q.thr, q.rps,
                                          q.part, q.m_pad, qa, bx, by, smem);
        else
            pm_fwd_block<float, 4, false>(q.st, rp, q.nx, q.ny, q.n_rows, q.sy, q.sz, q.thr, q.rps,
                                          q.part, q.m_pad, qa, bx, by, smem);
        __syncthreads();   // LDS is reused by the next ticket
    }
}

// Per station: slice sum + FP64 near-field correction; writes g (FP64) and/or
// gf (FP32) at the problem's station offset. Resets the forward ticket for
// the next pass (the persistent kernel before it has completed).
__global__ __launch_bounds__(256) void pm_fwd_reduce_grouped_kernel(
    const GFwdDesc* __restrict__ d, const int64_t* __restrict__ st_start, int np, int64_t total,
    const float* __restrict__ rho, double* __restrict__ g, float* __restrict__ gf,
    unsigned* __restrict__ ticket) {
    const int lane = threadIdx.x & 63;
    const int64_t t = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);   // one wave per station
    if (blockIdx.x == 0 && threadIdx.x == 0) *ticket = 0u;
    if (t >= total) return;
    const int p = find_problem(st_start, np, t);
    const GFwdDesc& q = d[p];
    const int i = (int)(t - st_start[p]);
    const double v = pm_fwd_reduce_wave(q.part, q.n_slices, q.m_pad, i, q.scale,
                                        rho + q.cell_off, q.na, lane);
    if (lane != 0) return;
    if (g) g[q.st_off + i] = v;
    if (gf) gf[q.st_off + i] = (float)v;
}

// Persistent adjoint: one ticket = 64 segments of C cells x one station split.
template <int MAXC>
__global__ __launch_bounds__(256) void pm_adj_grouped_kernel(
    const GAdjDesc* __restrict__ d, const int64_t* __restrict__ blk_start, int np, int total,
    const float* __restrict__ r, unsigned* __restrict__ ticket) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int slot;
    const QuadArgs qa{};
    for (;;) {
        const int b = next_ticket(ticket, &slot);
        if (b >= total) break;
        const int p = find_problem(blk_start, np, b);
        const GAdjDesc& q = d[p];
        const int local = b - (int)blk_start[p];
        const int bx = local % q.blocks_x, by = local / q.blocks_x;
        const float* rp = r + q.st_off;
        if (MAXC >= 16 && q.C == 16)
            pm_adj_block<float, (MAXC >= 16 ? 16 : 8), false>(
                q.st, rp, q.m, q.nx, q.ny, q.sy, q.sz, q.segs_per_row, q.n_segs, q.st_per_split,
                q.m_pad, q.part, nullptr, q.scale, q.n_cells, 0, qa, bx, by, smem);
        else
            pm_adj_block<float, 8, false>(q.st, rp, q.m, q.nx, q.ny, q.sy, q.sz, q.segs_per_row,
                                          q.n_segs, q.st_per_split, q.m_pad, q.part, nullptr,
                                          q.scale, q.n_cells, 0, qa, bx, by, smem);
        __syncthreads();
    }
}

// Grouped adjoint reduce, in two launches like the single-problem one: the
// split-K sum over every cell outside its problem's near-field layers (this
// launch also resets the ticket counter), then the near-field layers.
__global__ __launch_bounds__(256) void pm_adj_sum_grouped_kernel(
    const GAdjDesc* __restrict__ d, const int64_t* __restrict__ cell_start, int np,
    int64_t total, float* __restrict__ out, unsigned* __restrict__ ticket) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) *ticket = 0u;
    if (t >= total) return;
    const int p = find_problem(cell_start, np, t);
    const GAdjDesc& q = d[p];
    const int64_t n = t - cell_start[p];
    const int64_t kk = n / ((int64_t)q.nx * q.ny);
    if (kk >= q.klo && kk <= q.khi) return;
    pm_adj_sum_one<float>(q.part, q.splits, q.n_cells, n, q.scale, out + q.cell_off);
}

// One half-wave per near-field cell (block = 256 = 8 cells).
__global__ __launch_bounds__(256) void pm_adj_reduce_grouped_kernel(
    const GAdjDesc* __restrict__ d, const int64_t* __restrict__ near_start, int np,
    int64_t total, const float* __restrict__ r, float* __restrict__ out) {
    const int lane = threadIdx.x % kNearLanes;
    constexpr int kPerBlock = 256 / kNearLanes;
    for (int64_t t = (int64_t)blockIdx.x * kPerBlock + threadIdx.x / kNearLanes; t < total;
         t += (int64_t)gridDim.x * kPerBlock) {   // uniform per half-wave
        const int p = find_problem(near_start, np, t);
        const GAdjDesc& q = d[p];
        const int64_t n = q.near_off + (t - near_start[p]);
        const double v = pm_adj_reduce_wave<false, kCorrPmNear>(q.part, q.splits, q.n_cells, n,
                                                                r + q.st_off, q.na, lane);
        if (lane == 0) out[q.cell_off + n] = (float)(q.scale * v);
    }
}

// ------------------------------------------------------- station prep
__global__ void prep_ref_kernel(const float* __restrict__ s, int bfr, int di,
                                int fh, int fv, int ni_count, int m_int,
                                int m_pad, float4* __restrict__ st,
                                double4* __restrict__ st_d, int4* __restrict__ st_c,
                                int* __restrict__ out_map) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m_pad) return;
    if (t < m_int) {
        const int ni = bfr + t % ni_count;
        const int nj = bfr + t / ni_count;
        const int m = ni + nj * di;
        // Cell units (L = 1/fh): station (ni, nj, s) -> (ni*fh, nj*fh, s*fh);
        // cell (i,j,k) sits at (i, j, k*fh/fv).
        const double zd = (double)s[m] * fh;
        st[t] = make_float4((float)(ni * fh), (float)(nj * fh), (float)zd, 0.f);
        st_d[t] = make_double4((double)(ni * fh), (double)(nj * fh), zd, 0.0);
        st_c[t] = make_int4(ni * fh, nj * fh, (int)rint(zd * fv / fh), 0);
        out_map[t] = m;
    } else {
        st[t] = make_float4(0.f, 0.f, 0.f, 0.f);
        st_d[t] = make_double4(0.0, 0.0, 0.0, 0.0);
        st_c[t] = make_int4(0, 0, 0, 0);
    }
}

int prep_ref_stations(const float* s, int bfr, int di, int dj, int fh, int fv,
                      float4* st, double4* st_d, int4* st_c, int* out_map,
                      int m_int, int m_pad, hipStream_t stream) {
    (void)dj;
    const int ni_count = di - 2 * bfr;
    hipLaunchKernelGGL(prep_ref_kernel, dim3((m_pad + 255) / 256), dim3(256), 0,
                       stream, s, bfr, di, fh, fv, ni_count, m_int, m_pad, st,
                       st_d, st_c, out_map);
    ZTM_HIP(hipGetLastError());
    return ZTM_OK;
}

// ------------------------------------------------------------ launchers
constexpr int kTargetWgs = 2048;   // 8 workgroups per CU on 256 CUs

int fwd_group(int S, size_t elem, bool quad) {
    if (quad) return 1;
    if (elem != sizeof(float)) return S == 2 ? 4 : 1;
    return S <= 1 ? 8 : S == 2 ? 4 : S == 4 ? 2 : 1;
}

PmTuning fwd_adjust(const PmShape& s, PmTuning t, size_t elem, bool quad) {
    // a row group of G rows must fit the 32 KB LDS budget; wide rows run
    // without groups (S = 8 in FP32, S = 4 in FP64)
    const size_t nxp = (size_t)((s.nx + 3) & ~3);
    if (fwd_group(t.st_per_lane, elem, quad) > 1 &&
        nxp * elem * fwd_group(t.st_per_lane, elem, quad) > 32768)
        t.st_per_lane = elem == sizeof(float) ? 8 : 4;
    return t;
}

int fwd_rows_per_slice(const PmShape& s, const PmTuning& t0, size_t elem, bool quad) {
    const PmTuning t = fwd_adjust(s, t0, elem, quad);
    const int G = fwd_group(t.st_per_lane, elem, quad);
    const int nxp = (s.nx + 3) & ~3;
    const int lds_rows = std::max(1, (int)(32768 / ((size_t)nxp * elem)));
    int64_t rps;
    if (t.rows_per_slice > 0) {
        rps = std::min(t.rows_per_slice, lds_rows);
    } else {
        const int tiles = s.m_pad / (64 * t.st_per_lane);
        const int64_t want_slices = std::max<int64_t>(1, (kTargetWgs + tiles - 1) / tiles);
        rps = (s.n_rows + want_slices - 1) / want_slices;
        rps = std::max<int64_t>(rps, 4);
        rps = std::min<int64_t>(rps, lds_rows);
        rps = std::min<int64_t>(rps, s.n_rows);
        // the same number of row groups for each of the 4 waves: the nearest
        // multiple of 4 G rows that LDS holds (28 rows of 4-row groups were
        // 2, 2, 2 and 1 groups per wave)
        const int64_t q = 4 * G;
        if (rps > q && lds_rows >= q && s.n_rows > q)
            rps = std::min<int64_t>((rps + q / 2) / q * q, lds_rows / q * q);
    }
    if (rps > G) rps = rps / G * G;   // whole row groups per slice where possible
    return (int)std::max<int64_t>(rps, 1);
}

int adj_group(int C, size_t elem, bool quad) {
    if (quad) return 1;
    if (elem != sizeof(float)) return C >= 8 ? 2 : 4;
    return C >= 16 ? 2 : 4;
}

size_t adj_lds(int C, size_t elem, bool quad) {
    // station tile, then (FP32 point mass with row groups) the row-term
    // table; the cross-wave FP64 reduction reuses the same memory
    size_t b = std::max((size_t)256 * 4 * elem, (size_t)4 * C * 64 * sizeof(double));
    if (elem == sizeof(float) && !quad && adj_group(C, elem, quad) > 1)
        b = std::max(b, (size_t)256 * 4 * elem + (size_t)256 * kAdjRows * elem);
    return b;
}

int64_t adj_segments(const PmShape& s, int C, int G, int* segs_per_row) {
    const int spr = (s.nx + C - 1) / C;
    if (segs_per_row) *segs_per_row = spr;
    return (int64_t)s.nz * ((s.ny + G - 1) / G) * spr;
}

int adj_splits(const PmShape& s, const PmTuning& t, size_t elem, bool quad) {
    const int C = t.cells_per_lane;
    const int64_t segs = adj_segments(s, C, adj_group(C, elem, quad), nullptr);
    const int64_t wgs = (segs + 63) / 64;
    const int max_splits = s.m_pad / 256;
    int sp = t.adj_splits;
    if (sp <= 0) sp = (int)std::max<int64_t>(1, (kTargetWgs + wgs - 1) / wgs);
    return std::max(1, std::min(sp, max_splits));
}

namespace {

template <typename T, int S>
void launch_fwd(const PmShape& s, int rps, int n_slices, const void* st,
                const void* rho, T sy, T sz, T thr, double* part, bool quad,
                const QuadArgs& qa, hipStream_t stream) {
    using V4 = typename Vec4<T>::type;
    const int nxp = (s.nx + 3) & ~3;
    const int G = fwd_group(S, sizeof(T), quad);
    const size_t lds = std::max((size_t)((rps + G - 1) / G * G) * nxp * sizeof(T),
                                (size_t)4 * S * 64 * sizeof(double));
    const dim3 grid(s.m_pad / (64 * S), n_slices);
    if (quad)
        hipLaunchKernelGGL((pm_fwd_kernel<T, S, true>), grid, dim3(256), lds, stream,
                           (const V4*)st, (const T*)rho, s.nx, s.ny, (int)s.n_rows, sy, sz,
                           thr, rps, part, s.m_pad, qa);
    else
        hipLaunchKernelGGL((pm_fwd_kernel<T, S, false>), grid, dim3(256), lds, stream,
                           (const V4*)st, (const T*)rho, s.nx, s.ny, (int)s.n_rows, sy, sz,
                           thr, rps, part, s.m_pad, qa);
}

template <typename T, int C>
void launch_adj(const PmShape& s, int splits, int direct, const void* st, const void* r,
                T sy, T sz, double* part, void* out, double scale, bool quad,
                const QuadArgs& qa, hipStream_t stream) {
    using V4 = typename Vec4<T>::type;
    int segs_per_row = 0;
    const int64_t n_segs = adj_segments(s, C, adj_group(C, sizeof(T), quad), &segs_per_row);
    const int blocks_x = (int)((n_segs + 63) / 64);
    const int tiles = s.m_pad / 256;
    const int per_split = ((tiles + splits - 1) / splits) * 256;
    const size_t lds = adj_lds(C, sizeof(T), quad);
    if (quad)
        hipLaunchKernelGGL((pm_adj_kernel<T, C, true>), dim3(blocks_x, splits), dim3(256),
                           lds, stream, (const V4*)st, (const T*)r, s.m, s.nx, s.ny,
                           sy, sz, segs_per_row, n_segs, per_split, s.m_pad, part,
                           (T*)out, scale, s.n_cells, direct, qa);
    else
        hipLaunchKernelGGL((pm_adj_kernel<T, C, false>), dim3(blocks_x, splits), dim3(256),
                           lds, stream, (const V4*)st, (const T*)r, s.m, s.nx, s.ny,
                           sy, sz, segs_per_row, n_segs, per_split, s.m_pad, part,
                           (T*)out, scale, s.n_cells, direct, qa);
}

}  // namespace

size_t pm_fwd_workspace(const PmShape& s, const PmTuning& t, int precision, bool quad) {
    const size_t elem = precision == ZTM_FP64 ? sizeof(double) : sizeof(float);
    const int rps = fwd_rows_per_slice(s, t, elem, quad);
    const int64_t n_slices = (s.n_rows + rps - 1) / rps;
    return (size_t)n_slices * (size_t)s.m_pad;
}

size_t pm_adj_workspace(const PmShape& s, const PmTuning& t, int precision, bool near,
                        bool quad) {
    // near: a correction will run, so the main kernel writes partials
    const size_t elem = precision == ZTM_FP64 ? sizeof(double) : sizeof(float);
    const int sp = adj_splits(s, t, elem, quad);
    return (sp > 1 || near) ? (size_t)sp * (size_t)s.n_cells : 0;
}

int pm_fwd_launch(const PmShape& s, const PmTuning& t0, int precision,
                  const void* st, const void* rho, double sy, double sz,
                  double thr, double scale, double* part, double* g,
                  float* gref, const int* out_map, const NearArgs& near,
                  hipStream_t stream) {
    const size_t elem = precision == ZTM_FP64 ? sizeof(double) : sizeof(float);
    const int nxp = (s.nx + 3) & ~3;
    if ((size_t)nxp * elem * 1 > 65536)
        return fail(ZTM_EINVAL, "nx=%d too large for one LDS row (max %d)", s.nx,
                    (int)(65536 / elem) - 3);
    if (s.m == 0) return ZTM_OK;
    NearArgs na = near;
    // FP64 needs no near-field correction; with the FP32 path's prism zones
    // (kCorrPrism64) it adds the per-station FP64 corrections computed before
    if (precision == ZTM_FP64 && na.mode != kCorrPrism64) na.mode = kCorrNone;
    const bool quad = na.quad != 0 && (precision == ZTM_FP32 || na.mode == kCorrPrism64);
    const PmTuning t = fwd_adjust(s, t0, elem, quad);
    const int rps = fwd_rows_per_slice(s, t, elem, quad);
    const int n_slices = (int)((s.n_rows + rps - 1) / rps);
    if (n_slices > 65535) return fail(ZTM_EINVAL, "too many row slices (%d)", n_slices);
    {
        ProfScope prof("pm_fwd", stream);
        if (precision == ZTM_FP64) {
            const double a = sy, b = sz, c = thr;
            switch (t.st_per_lane) {
                case 1: launch_fwd<double, 1>(s, rps, n_slices, st, rho, a, b, c, part, quad, na.q, stream); break;
                case 2: launch_fwd<double, 2>(s, rps, n_slices, st, rho, a, b, c, part, quad, na.q, stream); break;
                default: launch_fwd<double, 4>(s, rps, n_slices, st, rho, a, b, c, part, quad, na.q, stream); break;
            }
        } else {
            const float a = (float)sy, b = (float)sz, c = (float)thr;
            switch (t.st_per_lane) {
                case 1: launch_fwd<float, 1>(s, rps, n_slices, st, rho, a, b, c, part, quad, na.q, stream); break;
                case 2: launch_fwd<float, 2>(s, rps, n_slices, st, rho, a, b, c, part, quad, na.q, stream); break;
                case 8: launch_fwd<float, 8>(s, rps, n_slices, st, rho, a, b, c, part, quad, na.q, stream); break;
                default: launch_fwd<float, 4>(s, rps, n_slices, st, rho, a, b, c, part, quad, na.q, stream); break;
            }
        }
        ZTM_HIP(hipGetLastError());
    }
    {
        ProfScope prof("pm_fwd_reduce", stream);
        hipLaunchKernelGGL(pm_fwd_reduce_kernel, dim3((s.m + 3) / 4), dim3(256), 0,
                           stream, part, n_slices, s.m_pad, s.m, scale, g, gref, out_map,
                           (const float*)rho, na);
        ZTM_HIP(hipGetLastError());
    }
    return ZTM_OK;
}

namespace {
constexpr int kFieldsMax = 8;
// Row group of the F-field forward (kFieldsGroup on the device).
int fields_group(int F) { return F > 4 ? 2 : fwd_group(2, sizeof(float), false); }
// Rows per slice for the F-field forward: the single-field choice for S = 2,
// held to F density tiles within the 32 KB LDS budget, in whole row groups.
int fields_rows_per_slice(const PmShape& s, int F) {
    PmTuning t;
    t.st_per_lane = 2;
    t.rows_per_slice = 0;
    const int nxp = (s.nx + 3) & ~3;
    const int G = fields_group(F);
    int rps = fwd_rows_per_slice(s, t, sizeof(float), false);
    const int cap = 32768 / (nxp * (int)sizeof(float) * F);   // rows of F tiles in LDS
    const int q = cap >= 4 * G ? 4 * G : G;   // as many row groups for each wave where LDS allows
    rps = std::min(rps, std::max(G, cap / q * q));
    return std::max(rps, 1);
}
}  // namespace

bool pm_fwd_fields_supported(const PmShape& s, int F) {
    if (F < 1 || F > kFieldsMax) return false;
    const int nxp = (s.nx + 3) & ~3;
    return (size_t)nxp * sizeof(float) * fields_group(F) * F <= 32768;
}

size_t pm_fwd_fields_workspace(const PmShape& s, int F) {
    const int rps = fields_rows_per_slice(s, F);
    return (size_t)((s.n_rows + rps - 1) / rps) * F * (size_t)s.m_pad;
}

int pm_fwd_fields_launch(const PmShape& s, int F, const float4* st, const float* rho,
                         int64_t field_stride, double sy, double sz, double thr, double scale,
                         double* part, double* g, const NearArgs& near, hipStream_t stream) {
    if (!pm_fwd_fields_supported(s, F))
        return fail(ZTM_EINVAL, "%d fields of rows of %d cells exceed the LDS budget", F, s.nx);
    if (near.quad) return fail(ZTM_EINVAL, "the multi-field forward has no quadrupole path");
    if (s.m == 0) return ZTM_OK;
    const int rps = fields_rows_per_slice(s, F);
    const int n_slices = (int)((s.n_rows + rps - 1) / rps);
    if (n_slices > 65535) return fail(ZTM_EINVAL, "too many row slices (%d)", n_slices);
    const int nxp = (s.nx + 3) & ~3;
    const int G = fields_group(F);
    const size_t lds = std::max((size_t)((rps + G - 1) / G * G) * nxp * sizeof(float) * F,
                                (size_t)4 * 2 * 64 * sizeof(double) * F);
    const dim3 grid(s.m_pad / 128, n_slices);
    {
        ProfScope prof("pm_fwd_fields", stream);
#define ZTM_FIELDS(FF)                                                                          \
    hipLaunchKernelGGL(pm_fwd_fields_kernel<FF>, grid, dim3(256), lds, stream, st, rho,         \
                       field_stride, s.nx, s.ny, (int)s.n_rows, (float)sy, (float)sz,            \
                       (float)thr, rps, part, s.m_pad)
        switch (F) {
            case 1: ZTM_FIELDS(1); break;
            case 2: ZTM_FIELDS(2); break;
            case 3: ZTM_FIELDS(3); break;
            case 4: ZTM_FIELDS(4); break;
            case 5: ZTM_FIELDS(5); break;
            case 6: ZTM_FIELDS(6); break;
            case 7: ZTM_FIELDS(7); break;
            default: ZTM_FIELDS(8); break;
        }
#undef ZTM_FIELDS
        ZTM_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(pm_fwd_fields_reduce_kernel, dim3((s.m + 3) / 4, F), dim3(256), 0, stream,
                       part, n_slices, F * s.m_pad, s.m, s.m_pad, scale, g, rho, field_stride,
                       near);
    ZTM_HIP(hipGetLastError());
    return ZTM_OK;
}

int pm_adj_launch(const PmShape& s, const PmTuning& t, int precision,
                  const void* st, const void* r, double sy, double sz,
                  double scale, double* part, void* out, const NearArgs& near,
                  hipStream_t stream) {
    if (s.n_cells == 0) return ZTM_OK;
    NearArgs na = near;
    if (precision == ZTM_FP64 && na.mode != kCorrPrism64) na.mode = kCorrNone;
    const bool quad = na.quad != 0 && (precision == ZTM_FP32 || na.mode == kCorrPrism64);
    const size_t elem = precision == ZTM_FP64 ? sizeof(double) : sizeof(float);
    const int splits = adj_splits(s, t, elem, quad);
    const int direct = (splits == 1 && na.mode == kCorrNone) ? 1 : 0;
    {
        ProfScope prof("pm_adj", stream);
        if (precision == ZTM_FP64) {
            switch (t.cells_per_lane) {
                case 4: launch_adj<double, 4>(s, splits, direct, st, r, sy, sz, part, out, scale, quad, na.q, stream); break;
                default: launch_adj<double, 8>(s, splits, direct, st, r, sy, sz, part, out, scale, quad, na.q, stream); break;
            }
        } else {
            const float a = (float)sy, b = (float)sz;
            switch (t.cells_per_lane) {
                case 4: launch_adj<float, 4>(s, splits, direct, st, r, a, b, part, out, scale, quad, na.q, stream); break;
                case 8: launch_adj<float, 8>(s, splits, direct, st, r, a, b, part, out, scale, quad, na.q, stream); break;
                default: launch_adj<float, 16>(s, splits, direct, st, r, a, b, part, out, scale, quad, na.q, stream); break;
            }
        }
        ZTM_HIP(hipGetLastError());
    }
    if (!direct) {
        ProfScope prof("pm_adj_reduce", stream);
        int klo = 0, khi = -1;
        near_layers(na, s.nz, &klo, &khi);
        const int64_t plane = (int64_t)s.nx * s.ny;
        const unsigned blocks = (unsigned)((s.n_cells + 255) / 256);
        if (precision == ZTM_FP64)
            hipLaunchKernelGGL((pm_adj_sum_kernel<double>), dim3(blocks), dim3(256), 0, stream,
                               part, splits, s.n_cells, plane, klo, khi, scale, (double*)out);
        else
            hipLaunchKernelGGL((pm_adj_sum_kernel<float>), dim3(blocks), dim3(256), 0, stream,
                               part, splits, s.n_cells, plane, klo, khi, scale, (float*)out);
        ZTM_HIP(hipGetLastError());
        if (klo <= khi && na.mode == kCorrPrism64) {
            const int64_t n0 = klo * plane, count = (int64_t)(khi - klo + 1) * plane;
            const unsigned nb = (unsigned)std::min<int64_t>((count + 7) / 8, 1 << 20);
            if (quad)
                hipLaunchKernelGGL((pm_adj_reduce64_kernel<true>), dim3(nb), dim3(256), 0, stream,
                                   part, splits, s.n_cells, n0, count, scale, (double*)out,
                                   (const double*)r, na);
            else
                hipLaunchKernelGGL((pm_adj_reduce64_kernel<false>), dim3(nb), dim3(256), 0, stream,
                                   part, splits, s.n_cells, n0, count, scale, (double*)out,
                                   (const double*)r, na);
            ZTM_HIP(hipGetLastError());
        } else if (klo <= khi && precision == ZTM_FP32) {
            const int64_t n0 = klo * plane, count = (int64_t)(khi - klo + 1) * plane;
            const unsigned nb = (unsigned)std::min<int64_t>((count + 7) / 8, 1 << 20);
#define ZTM_ADJ_NEAR(Q, M)                                                                 \
    hipLaunchKernelGGL((pm_adj_reduce_kernel<Q, M>), dim3(nb), dim3(256), 0, stream, part,  \
                       splits, s.n_cells, n0, count, scale, (float*)out, (const float*)r, na)
            if (na.mode == kCorrPrism) {
                if (quad) ZTM_ADJ_NEAR(true, kCorrPrism); else ZTM_ADJ_NEAR(false, kCorrPrism);
            } else {
                if (quad) ZTM_ADJ_NEAR(true, kCorrPmNear); else ZTM_ADJ_NEAR(false, kCorrPmNear);
            }
#undef ZTM_ADJ_NEAR
            ZTM_HIP(hipGetLastError());
        }
    }
    return ZTM_OK;
}


// ------------------------------------------------------ grouped launchers
int pm_grouped_workspace(ztm_plan p) {
    const PmTuning t = effective_tuning(p, ZTM_FP32);
    return ensure_workspace(p, std::max(pm_fwd_workspace(p->shape, t, ZTM_FP32, false),
                                        pm_adj_workspace(p->shape, t, ZTM_FP32, true, false)));
}

int pm_fwd_desc(ztm_plan p, int64_t cell_off, int64_t st_off, GFwdDesc* d, size_t* lds,
                int64_t* blocks) {
    const PmShape& s = p->shape;
    const PmTuning t = fwd_adjust(s, effective_tuning(p, ZTM_FP32), sizeof(float), false);
    const int nxp = (s.nx + 3) & ~3;
    if ((size_t)nxp * sizeof(float) > 65536) return fail(ZTM_EINVAL, "nx=%d too large", s.nx);
    if (t.st_per_lane != 2 && t.st_per_lane != 4 && t.st_per_lane != 8)
        return fail(ZTM_EINVAL, "grouped forward needs st_per_lane 2, 4 or 8 (got %d)",
                    t.st_per_lane);
    const int rps = fwd_rows_per_slice(s, t, sizeof(float), false);
    const int G = fwd_group(t.st_per_lane, sizeof(float), false);
    d->st = p->st_f;
    d->part = p->part;
    d->cell_off = cell_off;
    d->st_off = st_off;
    d->nx = s.nx;
    d->ny = s.ny;
    d->n_rows = (int)s.n_rows;
    d->rps = rps;
    d->m_pad = s.m_pad;
    d->tiles_x = s.m_pad / (64 * t.st_per_lane);
    d->n_slices = (int)((s.n_rows + rps - 1) / rps);
    d->m = s.m;
    d->S = t.st_per_lane;
    d->sy = (float)p->sy;
    d->sz = (float)p->sz;
    d->thr = (float)p->prob.threshold;
    d->scale = pm_scale(p);
    d->na = near_args(p);
    *lds = std::max((size_t)((rps + G - 1) / G * G) * nxp * sizeof(float),
                    (size_t)4 * t.st_per_lane * 64 * sizeof(double));
    *blocks = s.m == 0 ? 0 : (int64_t)d->tiles_x * d->n_slices;
    return ZTM_OK;
}

int pm_adj_desc(ztm_plan p, int64_t cell_off, int64_t st_off, GAdjDesc* d, size_t* lds,
                int64_t* blocks) {
    const PmTuning t = effective_tuning(p, ZTM_FP32);
    const PmShape& s = p->shape;
    const int C = t.cells_per_lane;
    if (C != 8 && C != 16)
        return fail(ZTM_EINVAL, "grouped adjoint needs cells_per_lane 8 or 16 (got %d)", C);
    const int splits = adj_splits(s, t, sizeof(float), false);
    int segs_per_row = 0;
    const int64_t n_segs = adj_segments(s, C, adj_group(C, sizeof(float), false), &segs_per_row);
    const int tiles = s.m_pad / 256;
    d->st = p->st_f;
    d->part = p->part;
    d->cell_off = cell_off;
    d->st_off = st_off;
    d->m = s.m;
    d->nx = s.nx;
    d->ny = s.ny;
    d->segs_per_row = segs_per_row;
    d->st_per_split = ((tiles + splits - 1) / splits) * 256;
    d->m_pad = s.m_pad;
    d->splits = splits;
    d->blocks_x = (int)((n_segs + 63) / 64);
    d->C = C;
    d->n_segs = n_segs;
    d->n_cells = s.n_cells;
    d->sy = (float)p->sy;
    d->sz = (float)p->sz;
    d->scale = pm_scale(p);
    d->na = near_args(p);
    near_layers(d->na, s.nz, &d->klo, &d->khi);
    const int64_t plane = (int64_t)s.nx * s.ny;
    d->near_off = d->klo <= d->khi ? d->klo * plane : 0;
    d->n_near = d->klo <= d->khi ? (int64_t)(d->khi - d->klo + 1) * plane : 0;
    *lds = adj_lds(C, sizeof(float), false);
    *blocks = (int64_t)d->blocks_x * splits;
    return ZTM_OK;
}

namespace {
// Persistent grid: a few resident workgroups per CU (256 CUs).
unsigned persistent_grid(int64_t total_blocks, int per_cu) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(total_blocks, 256 * per_cu));
}
}  // namespace

int pm_fwd_grouped_launch(const GroupedPass& gp, const float* rho, double* g, float* gf,
                          hipStream_t stream) {
    if (gp.np == 0) return ZTM_OK;
    if (gp.total_blocks > INT32_MAX) return fail(ZTM_EINVAL, "too many forward blocks");
    if (gp.total_blocks > 0) {
        ProfScope prof("pm_fwd_grouped", stream);
        const dim3 grid(persistent_grid(gp.total_blocks, gp.per_cu));
        const int total = (int)gp.total_blocks;
        const auto* dd = (const GFwdDesc*)gp.desc;
        if (gp.max_variant > 4)
            hipLaunchKernelGGL(pm_fwd_grouped_kernel<8>, grid, dim3(256), gp.lds, stream, dd,
                               gp.blk_start, gp.np, total, rho, gp.ticket);
        else
            hipLaunchKernelGGL(pm_fwd_grouped_kernel<4>, grid, dim3(256), gp.lds, stream, dd,
                               gp.blk_start, gp.np, total, rho, gp.ticket);
        ZTM_HIP(hipGetLastError());
    }
    const int64_t items = std::max<int64_t>(gp.total_items, 1);
    hipLaunchKernelGGL(pm_fwd_reduce_grouped_kernel, dim3((unsigned)((items + 3) / 4)),
                       dim3(256), 0, stream, (const GFwdDesc*)gp.desc, gp.item_start, gp.np,
                       gp.total_items, rho, g, gf, gp.ticket);
    ZTM_HIP(hipGetLastError());
    return ZTM_OK;
}

int pm_adj_grouped_launch(const GroupedPass& gp, const float* r, float* out,
                          hipStream_t stream) {
    if (gp.np == 0) return ZTM_OK;
    if (gp.total_blocks > INT32_MAX) return fail(ZTM_EINVAL, "too many adjoint blocks");
    if (gp.total_blocks > 0) {
        ProfScope prof("pm_adj_grouped", stream);
        const dim3 grid(persistent_grid(gp.total_blocks, gp.per_cu));
        const int total = (int)gp.total_blocks;
        const auto* dd = (const GAdjDesc*)gp.desc;
        if (gp.max_variant > 8)
            hipLaunchKernelGGL(pm_adj_grouped_kernel<16>, grid, dim3(256), gp.lds, stream, dd,
                               gp.blk_start, gp.np, total, r, gp.ticket);
        else
            hipLaunchKernelGGL(pm_adj_grouped_kernel<8>, grid, dim3(256), gp.lds, stream, dd,
                               gp.blk_start, gp.np, total, r, gp.ticket);
        ZTM_HIP(hipGetLastError());
    }
    const int64_t items = std::max<int64_t>(gp.total_items, 1);
    hipLaunchKernelGGL(pm_adj_sum_grouped_kernel, dim3((unsigned)((items + 255) / 256)),
                       dim3(256), 0, stream, (const GAdjDesc*)gp.desc, gp.item_start, gp.np,
                       gp.total_items, out, gp.ticket);
    ZTM_HIP(hipGetLastError());
    if (gp.total_near > 0) {
        hipLaunchKernelGGL(pm_adj_reduce_grouped_kernel,
                           dim3((unsigned)std::min<int64_t>((gp.total_near + 7) / 8, 1 << 20)),
                           dim3(256), 0, stream,
                           (const GAdjDesc*)gp.desc, gp.near_start, gp.np, gp.total_near, r,
                           out);
        ZTM_HIP(hipGetLastError());
    }
    return ZTM_OK;
}

}  // namespace ztm
