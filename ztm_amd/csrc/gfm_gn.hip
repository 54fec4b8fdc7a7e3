// Dense LA of the Gauss-Newton step (LMA::fit, create, annihilate;
// TFM.cpp:150,159-161,173), FP64 throughout:
//   normal equations  C = [J | r]^T [J | r]: J^T J (MTM, zmath.cpp:260-262)
//                     and J^T r in one symmetric product on FP64 MFMA;
//   damped steps      B[k] = b + (J^T J + d0 diag(J^T J) + d1 diag(1/sig^2))^-1
//                     J^T r for the damping grid, all conditions at once;
//   inverse diagonal  diag((J^T J + I d)^-1).
// The reference inverts with Gauss-Jordan (zmath.cpp:129-140) and multiplies;
// here the systems, symmetric positive definite whenever a damping term is
// positive, are factored by a blocked Cholesky, batched over the damping
// conditions, and solved by substitution. A system whose factorisation meets
// a pivot that is not positive is reported singular and, as the reference's
// zero inverse does, gives no step (and a zero inverse diagonal).
//
// Layouts (device, row-major): J [rows][n], r [rows], J^T J [n][n] (full,
// symmetric), J^T r [n]; factors [nd][n][n] (lower triangle used) in the
// caller's workspace, flags [nd] after them.
#include "gfm_internal.h"

#include <algorithm>
#include <cmath>

namespace ztm {

typedef double f64x4 __attribute__((ext_vector_type(4)));

namespace {

__device__ __forceinline__ f64x4 mfma64(double a, double b, f64x4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// ------------------------------------------------------- normal equations
// C = A^T A for A = [J | r] (rows x n1, n1 = n + 1 with r). A workgroup (4
// waves) forms a 128 x 128 tile (bi, bj), bi <= bj, over one K-slice of rows;
// wave w the 64 x 64 quadrant (w & 1, w >> 1): 4 x 4 MFMA tiles of 16 x 16.
// The slice walks 16 rows at a time: both 16 x 128 row strips go to LDS
// (stride 129: the lanes' 16 consecutive columns sit on distinct banks), the
// next strip's loads issued unconditionally into registers right after the
// barrier. MFMA f64 16x16x4: A[i][k] at lane i + 16k, B[k][j] at lane
// j + 16k, D[i][j] in register i / 4 of lane j + 16 (i % 4).
constexpr int kSyrkTile = 128;
constexpr int kSyrkLd = kSyrkTile + 1;
constexpr int kSyrkRows = 16;
constexpr int kSyrkLoads = kSyrkRows * kSyrkTile / 256;   // per thread per strip

struct SyrkStrip {
    double v[kSyrkLoads];
    unsigned ok;
    // strip rows k0 .. k0 + 15 (before k_end) of columns c0 .. c0 + 127 of A
    __device__ __forceinline__ void load(const double* __restrict__ J,
                                         const double* __restrict__ r, int n, int n1,
                                         int64_t k0, int64_t k_end, int c0) {
        ok = 0u;
#pragma unroll
        for (int u = 0; u < kSyrkLoads; ++u) {
            const int e = threadIdx.x + 256 * u;
            const int64_t k = k0 + e / kSyrkTile;
            const int c = c0 + e % kSyrkTile;
            const bool in = k < k_end && c < n1;
            ok |= in ? 1u << u : 0u;
            const double* p = !in ? J : c < n ? J + k * n + c : r + k;
            v[u] = *p;
        }
    }
    __device__ __forceinline__ void store(double* lds) const {
#pragma unroll
        for (int u = 0; u < kSyrkLoads; ++u) {
            const int e = threadIdx.x + 256 * u;
            lds[(e / kSyrkTile) * kSyrkLd + e % kSyrkTile] = (ok >> u) & 1u ? v[u] : 0.0;
        }
    }
};

__device__ __forceinline__ void upper_tile(int t, int nb, int& bi, int& bj) {
    // t-th tile (bi, bj), bi <= bj, row by row: row bi holds nb - bi tiles
    bi = 0;
    while (t >= nb - bi) {
        t -= nb - bi;
        ++bi;
    }
    bj = bi + t;
}

__global__ __launch_bounds__(256) void gn_syrk_kernel(const double* __restrict__ J,
                                                      const double* __restrict__ r, int n,
                                                      int n1, int64_t rows,
                                                      int64_t rows_per_slice, int nb,
                                                      double* __restrict__ part) {
    __shared__ double li[kSyrkRows * kSyrkLd];
    __shared__ double lj[kSyrkRows * kSyrkLd];
    int bi, bj;
    upper_tile(blockIdx.x, nb, bi, bj);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, lk = lane >> 4;
    const int wi = (wave & 1) * 64, wj = (wave >> 1) * 64;
    const int64_t k_begin = (int64_t)blockIdx.y * rows_per_slice;
    const int64_t k_end = min(rows, k_begin + rows_per_slice);
    f64x4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = f64x4{0.0, 0.0, 0.0, 0.0};
    SyrkStrip si, sj;
    si.load(J, r, n, n1, k_begin, k_end, bi * kSyrkTile);
    sj.load(J, r, n, n1, k_begin, k_end, bj * kSyrkTile);
    for (int64_t k0 = k_begin; k0 < k_end; k0 += kSyrkRows) {
        __syncthreads();   // every wave is done with the previous strip
        si.store(li);
        sj.store(lj);
        __syncthreads();
        if (k0 + kSyrkRows < k_end) {
            si.load(J, r, n, n1, k0 + kSyrkRows, k_end, bi * kSyrkTile);
            sj.load(J, r, n, n1, k0 + kSyrkRows, k_end, bj * kSyrkTile);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int kq = 4 * q + lk;
            double a[4], b[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                a[t] = li[kq * kSyrkLd + wi + 16 * t + l16];
                b[t] = lj[kq * kSyrkLd + wj + 16 * t + l16];
            }
#pragma unroll
            for (int ta = 0; ta < 4; ++ta)
#pragma unroll
                for (int tb = 0; tb < 4; ++tb) acc[ta][tb] = mfma64(a[ta], b[tb], acc[ta][tb]);
        }
    }
    // partial tile of this slice, full-matrix layout [n1][n1], upper tiles only
    double* out = part + (size_t)blockIdx.y * n1 * n1;
#pragma unroll
    for (int ta = 0; ta < 4; ++ta)
#pragma unroll
        for (int tb = 0; tb < 4; ++tb)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int i = bi * kSyrkTile + wi + 16 * ta + 4 * e + lk;
                const int j = bj * kSyrkTile + wj + 16 * tb + l16;
                if (i < n1 && j < n1) out[(size_t)i * n1 + j] = acc[ta][tb][e];
            }
}

// JJ[i][j] = JJ[j][i] = sum over slices of C[min(i,j)][max(i,j)] (an upper
// tile's entry, so the result is exactly symmetric), Jtr[i] = C[i][n].
__global__ __launch_bounds__(256) void gn_syrk_reduce_kernel(const double* __restrict__ part,
                                                             int ks, int n, int n1,
                                                             double* __restrict__ JJ,
                                                             double* __restrict__ jtr) {
    const int64_t total = (int64_t)n * n1;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * 256) {
        const int i = (int)(e / n1), j = (int)(e % n1);
        const size_t off = i <= j ? (size_t)i * n1 + j : (size_t)j * n1 + i;
        double s = 0.0;
        for (int k = 0; k < ks; ++k) s += part[(size_t)k * n1 * n1 + off];
        if (j < n)
            JJ[(size_t)i * n + j] = s;
        else if (jtr)
            jtr[i] = s;
    }
}

// ------------------------------------------------------ blocked Cholesky
constexpr int kNb = 32;            // block size of the factorisation
constexpr int kNbLd = kNb + 1;

// K_b = JJ with its diagonal raised: mode 0 (damped steps) adds
// d0 JJ_ii + d1 / sig_i^2 in the reference's order (TFM.cpp:160,
// (JJ + d0 M) + d1 A); mode 1 (inverse diagonal) adds d. Flags cleared.
__global__ __launch_bounds__(256) void gn_form_kernel(const double* __restrict__ JJ, int n,
                                                      int nd, int mode,
                                                      const double* __restrict__ sig,
                                                      const double* __restrict__ dm, double d,
                                                      double* __restrict__ K,
                                                      int* __restrict__ flags) {
    const int64_t nn = (int64_t)n * n;
    const int64_t total = nn * nd;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * 256) {
        const int b = (int)(e / nn);
        const int64_t o = e - (int64_t)b * nn;
        const int i = (int)(o / n), j = (int)(o % n);
        double v = JJ[o];
        if (i == j) {
            if (mode == 0) {
                const double s = sig[i];
                v = (v + dm[2 * b] * v) + dm[2 * b + 1] * (1.0 / (s * s));
            } else {
                v = v + d;
            }
        }
        K[e] = v;
        if (o == 0) flags[b] = 0;
    }
}

// (a) the diagonal block at p of every matrix, in LDS. grid = nd.
__global__ __launch_bounds__(256) void gn_potrf_diag_kernel(double* __restrict__ K, int n, int p,
                                                            int* __restrict__ flags) {
    __shared__ double a[kNb * kNbLd];
    double* Kb = K + (size_t)blockIdx.x * n * n;
    const int w = min(kNb, n - p);
    for (int t = threadIdx.x; t < w * w; t += 256) {
        const int i = t / w, j = t % w;
        a[i * kNbLd + j] = Kb[(size_t)(p + i) * n + p + j];
    }
    __syncthreads();
    for (int j = 0; j < w; ++j) {
        if (threadIdx.x == 0) {
            double d = a[j * kNbLd + j];
            if (!(d > 0.0)) {   // not positive definite (or NaN): singular
                flags[blockIdx.x] = 1;
                d = 1.0;
            }
            a[j * kNbLd + j] = sqrt(d);
        }
        __syncthreads();
        const double djj = a[j * kNbLd + j];
        for (int i = j + 1 + threadIdx.x; i < w; i += 256) a[i * kNbLd + j] /= djj;
        __syncthreads();
        const int m = w - j - 1;   // trailing (i, k), j < k <= i < w
        for (int t = threadIdx.x; t < m * m; t += 256) {
            const int i = j + 1 + t / m, k = j + 1 + t % m;
            if (k <= i) a[i * kNbLd + k] -= a[i * kNbLd + j] * a[k * kNbLd + j];
        }
        __syncthreads();
    }
    for (int t = threadIdx.x; t < w * w; t += 256) {
        const int i = t / w, j = t % w;
        if (j <= i) Kb[(size_t)(p + i) * n + p + j] = a[i * kNbLd + j];
    }
}

// (b) the panel below the diagonal block: rows r >= p + w, columns p .. p + w,
// X = A L^-T, one row per thread. grid = (ceil((n - p - w) / 64), nd); block 64.
__global__ __launch_bounds__(64) void gn_trsm_kernel(double* __restrict__ K, int n, int p) {
    __shared__ double l[kNb * kNbLd];
    __shared__ double x[64 * kNbLd];
    double* Kb = K + (size_t)blockIdx.y * n * n;
    const int w = min(kNb, n - p);
    const int r0 = p + w + blockIdx.x * 64;
    const int rows = min(64, n - r0);
    for (int t = threadIdx.x; t < w * w; t += 64) {
        const int i = t / w, j = t % w;
        l[i * kNbLd + j] = j <= i ? Kb[(size_t)(p + i) * n + p + j] : 0.0;
    }
    for (int t = threadIdx.x; t < rows * w; t += 64) {
        const int i = t / w, j = t % w;
        x[i * kNbLd + j] = Kb[(size_t)(r0 + i) * n + p + j];
    }
    __syncthreads();
    if (threadIdx.x < rows) {
        double* xr = x + threadIdx.x * kNbLd;
        for (int j = 0; j < w; ++j) {
            double s = xr[j];
            for (int k = 0; k < j; ++k) s -= xr[k] * l[j * kNbLd + k];
            xr[j] = s / l[j * kNbLd + j];
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < rows * w; t += 64) {
        const int i = t / w, j = t % w;
        Kb[(size_t)(r0 + i) * n + p + j] = x[i * kNbLd + j];
    }
}

// (c) the trailing lower triangle: K[i][j] -= sum_k L[i][p + k] L[j][p + k],
// i >= j >= p + kNb, 64 x 64 tiles (bi >= bj) on FP64 MFMA (each wave a
// 32 x 32 quadrant, 2 x 2 tiles of 16, K = kNb in steps of 4).
// grid = (tiles, nd). Only called with a full block (w = kNb).
__global__ __launch_bounds__(256) void gn_update_kernel(double* __restrict__ K, int n, int p,
                                                        int nt) {
    __shared__ double li[64 * kNbLd];
    __shared__ double lj[64 * kNbLd];
    double* Kb = K + (size_t)blockIdx.y * n * n;
    // blockIdx.x -> (bi, bj), bi >= bj, row by row: row bi holds bi + 1 tiles
    int t = blockIdx.x, bi = 0;
    while (t > bi) {
        t -= bi + 1;
        ++bi;
    }
    const int bj = t;
    const int q0 = p + kNb;
    const int i0 = q0 + 64 * bi, j0 = q0 + 64 * bj;
    for (int e = threadIdx.x; e < 64 * kNb; e += 256) {
        const int rr = e / kNb, k = e % kNb;
        li[rr * kNbLd + k] = i0 + rr < n ? Kb[(size_t)(i0 + rr) * n + p + k] : 0.0;
        lj[rr * kNbLd + k] = j0 + rr < n ? Kb[(size_t)(j0 + rr) * n + p + k] : 0.0;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, lk = lane >> 4;
    const int wi = (wave & 1) * 32, wj = (wave >> 1) * 32;
    f64x4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < kNb / 4; ++q) {
        const int k = 4 * q + lk;
        double a[2], b[2];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            a[s] = li[(wi + 16 * s + l16) * kNbLd + k];
            b[s] = lj[(wj + 16 * s + l16) * kNbLd + k];
        }
#pragma unroll
        for (int ta = 0; ta < 2; ++ta)
#pragma unroll
            for (int tb = 0; tb < 2; ++tb) acc[ta][tb] = mfma64(a[ta], b[tb], acc[ta][tb]);
    }
#pragma unroll
    for (int ta = 0; ta < 2; ++ta)
#pragma unroll
        for (int tb = 0; tb < 2; ++tb)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int i = i0 + wi + 16 * ta + 4 * e + lk;
                const int j = j0 + wj + 16 * tb + l16;
                if (i < n && j <= i) Kb[(size_t)i * n + j] -= acc[ta][tb][e];
            }
}

// ------------------------------------------------------------ substitution
// Forward substitution L y = rhs over rows q_begin .. n - 1 (earlier y are 0),
// y kept in LDS; blocks of kNb rows: 8 threads per row form the dot product
// with the solved part, then wave 0 solves the block's triangle.
__device__ void forward_subst(const double* __restrict__ L, int n, int q_begin, double* y,
                              double* s_blk) {
    const int tid = threadIdx.x;
    for (int q = q_begin; q < n; q += kNb) {
        const int w = min(kNb, n - q);
        {
            const int rho = tid >> 3, part = tid & 7;
            double s = 0.0;
            if (rho < w) {
                const double* row = L + (size_t)(q + rho) * n;
                for (int k = q_begin + part; k < q; k += 8) s += row[k] * y[k];
            }
            s += __shfl_xor(s, 1);
            s += __shfl_xor(s, 2);
            s += __shfl_xor(s, 4);
            if (part == 0 && rho < w) s_blk[rho] = y[q + rho] - s;
        }
        __syncthreads();
        if (tid < 64) {
            double v = tid < w ? s_blk[tid] : 0.0;
            for (int j = 0; j < w; ++j) {
                const double yj = __shfl(v, j) / L[(size_t)(q + j) * n + q + j];
                if (tid == j) v = yj;
                if (tid > j && tid < w) v -= L[(size_t)(q + tid) * n + q + j] * yj;
            }
            if (tid < w) y[q + tid] = v;
        }
        __syncthreads();
    }
}

// Backward substitution L^T x = y in place (x overwrites y), blocks from the
// last: the dot product with the solved part reads rows k of L (32 threads
// per row segment), then wave 0 solves the block's triangle.
__device__ void backward_subst(const double* __restrict__ L, int n, double* y, double* red) {
    const int tid = threadIdx.x;
    const int nblk = (n + kNb - 1) / kNb;
    for (int bq = nblk - 1; bq >= 0; --bq) {
        const int q = bq * kNb;
        const int w = min(kNb, n - q);
        {
            const int rho = tid & 31, part = tid >> 5;   // 8 parts over k
            double s = 0.0;
            if (rho < w)
                for (int k = q + w + part; k < n; k += 8) s += L[(size_t)k * n + q + rho] * y[k];
            red[part * kNb + rho] = s;
        }
        __syncthreads();
        if (tid < 64) {
            double v = 0.0;
            if (tid < w) {
                double s = 0.0;
                for (int pp = 0; pp < 8; ++pp) s += red[pp * kNb + tid];
                v = y[q + tid] - s;
            }
            for (int j = w - 1; j >= 0; --j) {
                const double xj = __shfl(v, j) / L[(size_t)(q + j) * n + q + j];
                if (tid == j) v = xj;
                if (tid < j) v -= L[(size_t)(q + j) * n + q + tid] * xj;
            }
            if (tid < w) y[q + tid] = v;
        }
        __syncthreads();
    }
}

// the substitutions keep the solution vector and 8 kNb partial sums in LDS,
// within the 64 KB a workgroup may take
constexpr int kSolveMaxN = 7680;

// B[b] = bvec + x_b with K_b x_b = rhs, or bvec when K_b is singular. grid = nd.
__global__ __launch_bounds__(256) void gn_solve_kernel(const double* __restrict__ K, int n,
                                                       const int* __restrict__ flags,
                                                       const double* __restrict__ rhs,
                                                       const double* __restrict__ bvec,
                                                       double* __restrict__ B) {
    extern __shared__ double sm[];
    double* y = sm;                  // [n]
    double* red = sm + n;            // [8 * kNb]
    const double* L = K + (size_t)blockIdx.x * n * n;
    for (int i = threadIdx.x; i < n; i += 256) y[i] = rhs[i];
    __syncthreads();
    forward_subst(L, n, 0, y, red);
    backward_subst(L, n, y, red);
    const bool sing = flags[blockIdx.x] != 0;
    for (int i = threadIdx.x; i < n; i += 256)
        B[(size_t)blockIdx.x * n + i] = bvec[i] + (sing ? 0.0 : y[i]);
}

// out[i] = (K^-1)_ii = |L^-1 e_i|^2, or 0 when K is singular. grid = n.
__global__ __launch_bounds__(256) void gn_invdiag_kernel(const double* __restrict__ L, int n,
                                                         const int* __restrict__ flags,
                                                         double* __restrict__ out) {
    extern __shared__ double sm[];
    double* y = sm;
    double* red = sm + n;
    const int i = blockIdx.x;
    const int q0 = i / kNb * kNb;
    for (int k = q0 + threadIdx.x; k < n; k += 256) y[k] = k == i ? 1.0 : 0.0;
    __syncthreads();
    forward_subst(L, n, q0, y, red);
    double s = 0.0;
    for (int k = i + threadIdx.x; k < n; k += 256) s += y[k] * y[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    __shared__ double wsum[4];
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0)
        out[i] = flags[0] ? 0.0 : (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]);
}

// K-slices of the normal-equation product: whole rounds of 512 resident
// workgroups to within 5%, at most 64, at least one 16-row strip each.
int syrk_slices(int tiles, int64_t rows) {
    const int64_t strips = (rows + kSyrkRows - 1) / kSyrkRows;
    const int max_ks = (int)std::max<int64_t>(1, std::min<int64_t>(64, strips));
    int best = 1;
    double best_fill = 0.0;
    for (int ks = 1; ks <= max_ks; ++ks) {
        const double rounds = (double)tiles * ks / 512.0;
        const double fill = rounds / std::ceil(rounds);
        if (fill >= 0.95) return ks;
        if (fill > best_fill + 1e-9) {
            best_fill = fill;
            best = ks;
        }
    }
    return best;
}

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

int syrk_tiles(int n1) {
    const int nb = (n1 + kSyrkTile - 1) / kSyrkTile;
    return nb * (nb + 1) / 2;
}

// Factor nd matrices K [nd][n][n] in place (lower triangle).
int cholesky_batched(double* K, int n, int nd, int* flags, hipStream_t stream) {
    for (int p = 0; p < n; p += kNb) {
        const int w = std::min(kNb, n - p);
        hipLaunchKernelGGL(gn_potrf_diag_kernel, dim3(nd), dim3(256), 0, stream, K, n, p, flags);
        ZTM_HIP(hipGetLastError());
        const int below = n - p - w;
        if (below <= 0) break;
        hipLaunchKernelGGL(gn_trsm_kernel, dim3((below + 63) / 64, nd), dim3(64), 0, stream, K,
                           n, p);
        ZTM_HIP(hipGetLastError());
        const int nt = (below + 63) / 64;
        hipLaunchKernelGGL(gn_update_kernel, dim3(nt * (nt + 1) / 2, nd), dim3(256), 0, stream,
                           K, n, p, nt);
        ZTM_HIP(hipGetLastError());
    }
    return ZTM_OK;
}

}  // namespace

size_t gn_workspace_bytes(int64_t rows, int n, int nd) {
    const int n1 = n + 1;
    const size_t syrk =
        sizeof(double) * (size_t)syrk_slices(syrk_tiles(n1), rows) * n1 * n1;
    const size_t chol = align256(sizeof(double) * (size_t)std::max(nd, 1) * n * n) +
                        align256(sizeof(int) * (size_t)std::max(nd, 1));
    return std::max(syrk, chol);
}

int gn_normal_launch(int64_t rows, int n, const double* J, const double* r, double* JJ,
                     double* jtr, void* work, hipStream_t stream) {
    const int n1 = r ? n + 1 : n;
    const int nb = (n1 + kSyrkTile - 1) / kSyrkTile;
    const int tiles = nb * (nb + 1) / 2;
    // the slice count the workspace was sized for (with the r column)
    const int ks = syrk_slices(syrk_tiles(n + 1), rows);
    const int64_t strips = (rows + kSyrkRows - 1) / kSyrkRows;
    const int64_t rps = (strips + ks - 1) / ks * kSyrkRows;   // whole strips per slice
    const int ks_used = (int)((rows + rps - 1) / rps);
    double* part = (double*)work;
    {
        ProfScope prof("gn_syrk", stream);
        hipLaunchKernelGGL(gn_syrk_kernel, dim3(tiles, ks_used), dim3(256), 0, stream, J, r, n,
                           n1, rows, rps, nb, part);
        ZTM_HIP(hipGetLastError());
    }
    const int64_t total = (int64_t)n * n1;
    const unsigned blocks = (unsigned)std::min<int64_t>((total + 255) / 256, 1 << 16);
    hipLaunchKernelGGL(gn_syrk_reduce_kernel, dim3(blocks), dim3(256), 0, stream, part, ks_used,
                       n, n1, JJ, r ? jtr : nullptr);
    ZTM_HIP(hipGetLastError());
    return ZTM_OK;
}

int gn_steps_launch(int n, int nd, const double* JJ, const double* jtr, const double* sig,
                    const double* dm, const double* b, double* B, void* work,
                    hipStream_t stream) {
    if (n > kSolveMaxN)
        return fail(ZTM_EINVAL, "the damped steps support up to %d parameters (got %d)",
                    kSolveMaxN, n);
    double* K = (double*)work;
    int* flags = (int*)((char*)work + align256(sizeof(double) * (size_t)nd * n * n));
    const int64_t total = (int64_t)nd * n * n;
    hipLaunchKernelGGL(gn_form_kernel, dim3((unsigned)std::min<int64_t>((total + 255) / 256,
                                                                         1 << 16)),
                       dim3(256), 0, stream, JJ, n, nd, 0, sig, dm, 0.0, K, flags);
    ZTM_HIP(hipGetLastError());
    {
        ProfScope prof("gn_cholesky", stream);
        if (int rc = cholesky_batched(K, n, nd, flags, stream)) return rc;
    }
    const size_t lds = sizeof(double) * ((size_t)n + 8 * kNb);
    ProfScope prof("gn_solve", stream);
    hipLaunchKernelGGL(gn_solve_kernel, dim3(nd), dim3(256), lds, stream, K, n, flags, jtr, b, B);
    ZTM_HIP(hipGetLastError());
    return ZTM_OK;
}

int gn_inverse_diag_launch(int n, const double* JJ, double d, double* out, void* work,
                           hipStream_t stream) {
    if (n > kSolveMaxN)
        return fail(ZTM_EINVAL, "the inverse diagonal supports up to %d parameters (got %d)",
                    kSolveMaxN, n);
    double* K = (double*)work;
    int* flags = (int*)((char*)work + align256(sizeof(double) * (size_t)n * n));
    const int64_t total = (int64_t)n * n;
    hipLaunchKernelGGL(gn_form_kernel, dim3((unsigned)std::min<int64_t>((total + 255) / 256,
                                                                         1 << 16)),
                       dim3(256), 0, stream, JJ, n, 1, 1, nullptr, nullptr, d, K, flags);
    ZTM_HIP(hipGetLastError());
    if (int rc = cholesky_batched(K, n, 1, flags, stream)) return rc;
    const size_t lds = sizeof(double) * ((size_t)n + 8 * kNb);
    hipLaunchKernelGGL(gn_invdiag_kernel, dim3(n), dim3(256), lds, stream, K, n, flags, out);
    ZTM_HIP(hipGetLastError());
    return ZTM_OK;
}

}  // namespace ztm
