// Dense LA of the Gauss-Newton step (LMA::fit, create, annihilate;
// TFM.cpp:150,159-161,173), FP64 throughout:
//   normal equations  J^T J (MTM, zmath.cpp:260-262) on FP64 MFMA and J^T r
//                     in the same pass over J;
//   damped steps      B[k] = b + (J^T J + d0 diag(J^T J) + d1 diag(1/sig^2))^-1
//                     J^T r for the damping grid, all conditions at once;
//   inverse diagonal  diag((J^T J + I d)^-1).
// The reference inverts with Gauss-Jordan (zmath.cpp:129-140) and multiplies;
// here the systems, symmetric positive definite whenever a damping term is
// positive, are factored by a blocked Cholesky, batched over the damping
// conditions, and solved by substitution with the inverses of the 32 x 32
// diagonal factor blocks. A system whose factorisation meets a pivot that is
// not positive is reported singular and, as the reference's zero inverse
// does, gives no step (and a zero inverse diagonal).
//
// Layouts (device, row-major): J [rows][n], r [rows], J^T J [n][n] (full,
// symmetric), J^T r [n]. Workspace of the factorisations: factors K [nd][n][n]
// (lower triangle used), the inverses of their diagonal blocks
// Dinv [nd][n / 32][32][32] (lower triangle, zeros above), flags [nd].
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
// C = J^T J by 64 x 64 tiles (bi, bj), bi <= bj, each over one K-slice of
// rows: a workgroup (4 waves) forms one tile, wave w the 32 x 32 quadrant
// (w & 1, w >> 1) as 2 x 2 MFMA tiles of 16 x 16. The slice walks 16 rows at a
// time through two LDS buffers (one barrier per strip): while the MFMAs read
// strip s, the loads of strip s + 1 are in flight into registers. MFMA f64
// 16x16x4: A[i][k] at lane i + 16k, B[k][j] at lane j + 16k, D[i][j] in
// register i / 4 of lane j + 16 (i % 4). The diagonal tiles also form their
// 64 entries of J^T r on the VALU from the same strips.
//
// Workgroups go to the 8 XCDs round-robin. With the slice count a multiple
// of 8, workgroup w runs slice (w % 8) + 8 (w / 8 / tiles), so every tile of
// a slice runs on one XCD and the slice's rows are fetched into that XCD's L2
// once for all its tiles.
constexpr int kSyrkTile = 64;
constexpr int kSyrkLd = kSyrkTile + 1;
constexpr int kSyrkRows = 16;
constexpr int kSyrkLoads = kSyrkRows * kSyrkTile / 256;   // per thread per operand
constexpr int kXcds = 8;

struct SyrkStrip {
    double v[kSyrkLoads];
    unsigned ok;
    // strip rows k0 .. k0 + 15 (before k_end) of columns c0 .. c0 + 63 of J.
    // Entries outside J load J[0] and are zeroed at the store, so that no
    // instruction between the load and the store needs the loaded value.
    __device__ __forceinline__ void load(const double* __restrict__ J, int n, int64_t k0,
                                         int64_t k_end, int c0) {
        ok = 0u;
#pragma unroll
        for (int u = 0; u < kSyrkLoads; ++u) {
            const int e = threadIdx.x + 256 * u;
            const int64_t k = k0 + e / kSyrkTile;
            const int c = c0 + e % kSyrkTile;
            const bool in = k < k_end && c < n;
            ok |= in ? 1u << u : 0u;
            v[u] = *(in ? J + k * n + c : J);
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
                                                      int64_t rows, int64_t rows_per_slice,
                                                      int nb, int tiles, int xcd_map,
                                                      double* __restrict__ part,
                                                      double* __restrict__ jpart) {
    __shared__ double li[2][kSyrkRows * kSyrkLd];
    __shared__ double lj[2][kSyrkRows * kSyrkLd];
    __shared__ double lr[2][kSyrkRows];
    int tile, slice;
    if (xcd_map) {
        const int x = blockIdx.x % kXcds, idx = blockIdx.x / kXcds;
        slice = x + kXcds * (idx / tiles);
        tile = idx % tiles;
    } else {
        tile = blockIdx.x % tiles;
        slice = blockIdx.x / tiles;
    }
    int bi, bj;
    upper_tile(tile, nb, bi, bj);
    const bool diag = bi == bj;
    const int ci = bi * kSyrkTile, cj = bj * kSyrkTile;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, lk = lane >> 4;
    const int wi = (wave & 1) * 32, wj = (wave >> 1) * 32;
    const int64_t k_begin = (int64_t)slice * rows_per_slice;
    const int64_t k_end = min(rows, k_begin + rows_per_slice);   // empty when k_begin >= rows
    const bool with_r = diag && r != nullptr;
    f64x4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = f64x4{0.0, 0.0, 0.0, 0.0};
    double jr = 0.0;
    SyrkStrip si, sj;
    double rv = 0.0;
    si.load(J, n, k_begin, k_end, ci);
    if (!diag) sj.load(J, n, k_begin, k_end, cj);
    if (with_r && tid < kSyrkRows && k_begin + tid < k_end) rv = r[k_begin + tid];
    si.store(li[0]);
    if (!diag) sj.store(lj[0]);
    if (with_r && tid < kSyrkRows) lr[0][tid] = rv;
    __syncthreads();
    int buf = 0;
    for (int64_t k0 = k_begin; k0 < k_end; k0 += kSyrkRows) {
        const bool more = k0 + kSyrkRows < k_end;
        if (more) {
            si.load(J, n, k0 + kSyrkRows, k_end, ci);
            if (!diag) sj.load(J, n, k0 + kSyrkRows, k_end, cj);
            rv = 0.0;
            if (with_r && tid < kSyrkRows && k0 + kSyrkRows + tid < k_end)
                rv = r[k0 + kSyrkRows + tid];
        }
        const double* A = li[buf];
        const double* B = diag ? li[buf] : lj[buf];
#pragma unroll
        for (int q = 0; q < kSyrkRows / 4; ++q) {
            const int kq = 4 * q + lk;
            double a[2], b[2];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                a[t] = A[kq * kSyrkLd + wi + 16 * t + l16];
                b[t] = B[kq * kSyrkLd + wj + 16 * t + l16];
            }
#pragma unroll
            for (int ta = 0; ta < 2; ++ta)
#pragma unroll
                for (int tb = 0; tb < 2; ++tb) acc[ta][tb] = mfma64(a[ta], b[tb], acc[ta][tb]);
        }
        if (with_r && tid < kSyrkTile) {
#pragma unroll
            for (int k = 0; k < kSyrkRows; ++k) jr += A[k * kSyrkLd + tid] * lr[buf][k];
        }
        if (more) {
            si.store(li[buf ^ 1]);
            if (!diag) sj.store(lj[buf ^ 1]);
            if (with_r && tid < kSyrkRows) lr[buf ^ 1][tid] = rv;
        }
        __syncthreads();
        buf ^= 1;
    }
    // this slice's partial tile, full-matrix layout [n][n], upper tiles only
    double* out = part + (size_t)slice * n * n;
#pragma unroll
    for (int ta = 0; ta < 2; ++ta)
#pragma unroll
        for (int tb = 0; tb < 2; ++tb)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int i = ci + wi + 16 * ta + 4 * e + lk;
                const int j = cj + wj + 16 * tb + l16;
                if (i < n && j < n) out[(size_t)i * n + j] = acc[ta][tb][e];
            }
    if (with_r && tid < kSyrkTile && ci + tid < n) jpart[(size_t)slice * n + ci + tid] = jr;
}

// JJ from the slices' partial upper tiles, in 32 x 32 tiles (bi, bj), bi <= bj:
// each thread sums 4 entries over the slices in order (rows of 32 entries
// read across threads) and writes them, and, off the diagonal, the block
// writes the tile's transpose to (bj, bi) through LDS, so both halves are
// the same sums and JJ is exactly symmetric. On a diagonal tile every entry
// is summed from the upper one, C[min(i,j)][max(i,j)]. Blocks past the
// tiles sum J^T r from the diagonal tiles' entries. grid = tiles + ceil(n / 256).
constexpr int kRedTile = 32;

__global__ __launch_bounds__(256) void gn_syrk_reduce_kernel(const double* __restrict__ part,
                                                             const double* __restrict__ jpart,
                                                             int ks, int n, int nb, int tiles,
                                                             double* __restrict__ JJ,
                                                             double* __restrict__ jtr) {
    const size_t nn = (size_t)n * n;
    const int tid = threadIdx.x;
    if ((int)blockIdx.x >= tiles) {
        const int i = (blockIdx.x - tiles) * 256 + tid;
        if (jtr && i < n) {
            double s = 0.0;
            for (int k = 0; k < ks; ++k) s += jpart[(size_t)k * n + i];
            jtr[i] = s;
        }
        return;
    }
    __shared__ double t[kRedTile][kRedTile + 1];
    int bi, bj;
    upper_tile(blockIdx.x, nb, bi, bj);
    const int tx = tid & 31, ty = tid >> 5;
    // the thread's 4 entries summed side by side, slice by slice in order,
    // so that many loads are in flight (entries outside the matrix read
    // entry 0 and are discarded)
    size_t off[4];
    bool in[4];
    double s[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i = bi * kRedTile + ty + 8 * r, j = bj * kRedTile + tx;
        in[r] = i < n && j < n;
        off[r] = !in[r] ? 0 : i <= j ? (size_t)i * n + j : (size_t)j * n + i;
        s[r] = 0.0;
    }
#pragma unroll 8
    for (int k = 0; k < ks; ++k) {
        const double* pk = part + k * nn;
#pragma unroll
        for (int r = 0; r < 4; ++r) s[r] += pk[off[r]];
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int li = ty + 8 * r;
        const int i = bi * kRedTile + li, j = bj * kRedTile + tx;
        if (in[r]) JJ[(size_t)i * n + j] = s[r];
        t[li][tx] = in[r] ? s[r] : 0.0;
    }
    if (bi == bj) return;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int li = ty + 8 * r;
        const int i = bj * kRedTile + li, j = bi * kRedTile + tx;   // (j, i) of the tile above
        if (i < n && j < n) JJ[(size_t)i * n + j] = t[tx][li];
    }
}

// ------------------------------------------------------ blocked Cholesky
constexpr int kNb = 32;            // block size of the factorisation
constexpr int kNbLd = kNb + 1;
constexpr int kPanelRows = 64;     // panel rows per workgroup of the panel step

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

// One step of the blocked Cholesky, at column block q of every matrix
// (grid = nd x (panel workgroups + update tiles)), with a look-ahead of one
// block: the previous block's (p = q - 32) trailing update is split between
// this step's two roles, which touch disjoint columns.
//
// Panel role (workgroups x < pw). First the previous block's update of the
// columns q .. q + w of this workgroup's rows: the diagonal block and its 64
// panel rows, each minus L[row][p..p+32) L[q+c][p..p+32)^T, on FP64 MFMA with
// the operands read from K into registers. Then wave 0 factors the 32 x 32
// diagonal block A by symmetric elimination on [A | I] without pivoting,
// lanes l < 32 holding row l of A and lanes 32 + c column c of I in
// registers: step j publishes column j of A through LDS (a_jj is the pivot
// d_j) and every lane subtracts a_kj (reg[j] / d_j) from its reg[k], k > j,
// which updates the rows of A below j and the columns of the right half. No
// workgroup barrier sits in this chain. A = L' D L'^T, with L'^-1 left in the
// right half, each row j scaled by d_j^-1/2 once final, so that it holds
// L^-1 = D^-1/2 L'^-1. A pivot that is not positive (or NaN) flags the matrix
// singular and is taken as 1; the rows and columns past a last block's width
// are an identity. The workgroup then solves its 64 panel rows,
// X = A_panel L^-T, on FP64 MFMA; workgroup 0 also writes L^-1 (Dinv) and the
// flag. The factorisation is repeated in every panel workgroup: they run side
// by side, and the diagonal block of K keeps its values for all of them to
// read. L's diagonal blocks are never stored: the substitutions take them
// from Dinv.
//
// Update role (the other workgroups): the rest of the previous block's
// update, K[i][j] -= sum_k L[i][p + k] L[j][p + k] for i >= j >= q + 32, in
// 64 x 64 tiles (bi >= bj) on FP64 MFMA (each wave a 32 x 32 quadrant, 2 x 2
// tiles of 16, K = 32 in steps of 4), operands and the tile's entries loaded
// into registers at once.
//
// So a step costs one launch instead of two, and the next block's panel,
// which waits on only 32 columns of the update, no longer waits for all of it.
//
// MFMA f64 16x16x4: A[i][k] at lane i + 16k, B[k][j] at lane j + 16k, D[i][j]
// in register i / 4 of lane j + 16 (i % 4); a lane (l16, lk) = (lane % 16,
// lane / 16) supplies the operands' entries (l16, 4 s + lk) of k-step s.
constexpr int kStepLds = 4 * kNb * kNbLd;   // doubles: the diagonal block, L^-1, 64 panel rows

// Orders one wave's LDS accesses across its lanes: the wave's stores have
// completed (lgkmcnt(0)) before any later access issues, and the fences keep
// the compiler from moving accesses across this point.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0)
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ void chol_panel(double* __restrict__ Kb, int n, int q, int xw,
                           double* __restrict__ Db, int* __restrict__ flag, double* sm) {
    double* a = sm;                      // [kNb][kNbLd] the diagonal block (lower)
    double* e = sm + kNb * kNbLd;        // [kNb][kNbLd] L^-1 of the block
    double* x = sm + 2 * kNb * kNbLd;    // [kPanelRows][kNbLd] this workgroup's panel rows
    double* col = sm + kStepLds;         // [2][kNb] the elimination's published column
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, lk = lane >> 4;
    const int ct = wave & 1;             // column tile (16 of the block's 32) of both products
    const int w = min(kNb, n - q);
    const int r0 = q + w + xw * kPanelRows;
    const int prow = max(0, min(kPanelRows, n - r0));
    const bool prev = q > 0;
    const int p = q - kNb;
    // the previous block's update of columns q .. q + w: row tiles rt = 0, 1
    // (the diagonal block) and 2 .. 5 (the panel rows), this wave's three
    // (wave >> 1) + 2u, all against column tile ct; the loads go first
    f64x4 acc[3];
#pragma unroll
    for (int u = 0; u < 3; ++u) acc[u] = f64x4{0.0, 0.0, 0.0, 0.0};
    double av[3][kNb / 4], bv[kNb / 4];
    if (prev) {
        const int bj = 16 * ct + l16;
        const bool bin = bj < w;
        const double* bp = Kb + (size_t)(q + (bin ? bj : 0)) * n + p + lk;
#pragma unroll
        for (int s = 0; s < kNb / 4; ++s) {
            const double v = bp[4 * s];
            bv[s] = bin ? v : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 3; ++u) {
            const int rt = (wave >> 1) + 2 * u;
            const int ri = rt < 2 ? 16 * rt + l16 : 16 * (rt - 2) + l16;
            const bool in = rt < 2 ? ri < w : ri < prow;
            const int row = in ? (rt < 2 ? q + ri : r0 + ri) : q;
            const double* ap = Kb + (size_t)row * n + p + lk;
#pragma unroll
            for (int s = 0; s < kNb / 4; ++s) {
                const double v = ap[4 * s];
                av[u][s] = in ? v : 0.0;
            }
        }
    }
    for (int t = tid; t < kNb * kNb; t += 256) {
        const int i = t / kNb, j = t % kNb;
        const bool in = i < w && j <= i;
        const double v = *(in ? Kb + (size_t)(q + i) * n + q + j : Kb);
        a[i * kNbLd + j] = in ? v : 0.0;
    }
    for (int t = tid; t < kPanelRows * kNb; t += 256) {
        const int i = t / kNb, j = t % kNb;
        const bool in = i < prow && j < w;
        const double v = *(in ? Kb + (size_t)(r0 + i) * n + q + j : Kb);
        x[i * kNbLd + j] = in ? v : 0.0;
    }
    if (prev) {
#pragma unroll
        for (int u = 0; u < 3; ++u)
#pragma unroll
            for (int s = 0; s < kNb / 4; ++s) acc[u] = mfma64(av[u][s], bv[s], acc[u]);
    }
    __syncthreads();
    if (prev) {   // each entry belongs to one lane
#pragma unroll
        for (int u = 0; u < 3; ++u) {
            const int rt = (wave >> 1) + 2 * u;
            const int j = 16 * ct + l16;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = 16 * rt + 4 * r + lk;
                if (rt < 2) {
                    if (j <= i) a[i * kNbLd + j] -= acc[u][r];
                } else {
                    x[(i - kNb) * kNbLd + j] -= acc[u][r];
                }
            }
        }
        __syncthreads();
    }
    if (wave == 0) {
        const int l = lane & 31;
        const bool row_lane = lane < 32;
        double reg[kNb];
#pragma unroll
        for (int k = 0; k < kNb; ++k) {
            const double v = a[l * kNbLd + k];
            reg[k] = row_lane && l < w ? v : (k == l ? 1.0 : 0.0);
        }
        bool singular = false;
        double piv = 1.0;   // row lanes: the pivot of their row
        if (row_lane) col[l] = reg[0];
#pragma unroll
        for (int j = 0; j < kNb; ++j) {
            const double* cj = col + (j & 1) * kNb;   // column j, published
            double* cn = col + ((j + 1) & 1) * kNb;   // column j + 1, published here
            wave_lds_sync();
            double d = cj[j];
            if (!(d > 0.0)) {
                singular = true;
                d = 1.0;
            }
            double cv[kNb];   // the column below the pivot, all loads issued at once
#pragma unroll
            for (int k = j + 1; k < kNb; ++k) cv[k] = cj[k];
            piv = l == j ? d : piv;
            const double y = reg[j] * (1.0 / d);
#pragma unroll
            for (int k = j + 1; k < kNb; ++k) {
                reg[k] -= cv[k] * y;
                // done here: left free, the compiler defers each update to
                // the step that reads it and keeps every loaded column live
                asm volatile("" : "+v"(reg[k]));
                // the next column goes out first, the rest of the step behind it
                if (k == j + 1 && row_lane) cn[l] = reg[k];
            }
        }
        // L^-1 = D^-1/2 L'^-1: row k scaled by d_k^-1/2, published by row lane k
        if (row_lane) col[l] = 1.0 / sqrt(piv);
        wave_lds_sync();
        if (!row_lane) {
#pragma unroll
            for (int k = 0; k < kNb; ++k)
                e[k * kNbLd + l] = k < w && l <= k ? reg[k] * col[k] : 0.0;
        }
        if (xw == 0 && lane == 0 && singular) *flag = 1;
    }
    __syncthreads();
    if (xw == 0) {
        for (int t = tid; t < kNb * kNb; t += 256) Db[t] = e[(t / kNb) * kNbLd + t % kNb];
    }
    // X[i][j] = sum_k A[i][k] L^-1[j][k]: row tiles (wave >> 1) + 2u of the 4
    if (prow > 0) {
        double lb[kNb / 4];
#pragma unroll
        for (int s = 0; s < kNb / 4; ++s) lb[s] = e[(16 * ct + l16) * kNbLd + 4 * s + lk];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int rt = (wave >> 1) + 2 * u;
            f64x4 c = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int s = 0; s < kNb / 4; ++s)
                c = mfma64(x[(16 * rt + l16) * kNbLd + 4 * s + lk], lb[s], c);
            const int j = 16 * ct + l16;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = 16 * rt + 4 * r + lk;
                if (i < prow && j < w) Kb[(size_t)(r0 + i) * n + q + j] = c[r];
            }
        }
    }
}

// Update tile t of the region i >= j >= u0 from the panel at columns p.
__device__ void chol_update(double* __restrict__ Kb, int n, int p, int u0, int t) {
    // t -> (bi, bj), bi >= bj, row by row: row bi holds bi + 1 tiles
    int bi = 0;
    while (t > bi) {
        t -= bi + 1;
        ++bi;
    }
    const int bj = t;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, lk = lane >> 4;
    const int i0 = u0 + 64 * bi + (wave & 1) * 32, j0 = u0 + 64 * bj + (wave >> 1) * 32;
    // this lane's 16 entries of the tile start the accumulators, and the
    // products come in negated (an exact sign flip of one operand), so the
    // MFMAs leave K[i][j] - sum_k directly; all loads are issued at once
    f64x4 acc[2][2];
#pragma unroll
    for (int ta = 0; ta < 2; ++ta)
#pragma unroll
        for (int tb = 0; tb < 2; ++tb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + 16 * ta + 4 * r + lk, j = j0 + 16 * tb + l16;
                const bool in = i < n && j <= i;
                const double v = *(in ? Kb + (size_t)i * n + j : Kb);
                acc[ta][tb][r] = in ? v : 0.0;
            }
    double av[2][kNb / 4], bv[2][kNb / 4];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int ra = i0 + 16 * h + l16, rb = j0 + 16 * h + l16;
        const double* ap = Kb + (size_t)(ra < n ? ra : 0) * n + p + lk;
        const double* bp = Kb + (size_t)(rb < n ? rb : 0) * n + p + lk;
#pragma unroll
        for (int s = 0; s < kNb / 4; ++s) {
            const double va = ap[4 * s], vb = bp[4 * s];
            av[h][s] = ra < n ? va : 0.0;
            bv[h][s] = rb < n ? -vb : 0.0;
        }
    }
#pragma unroll
    for (int s = 0; s < kNb / 4; ++s)
#pragma unroll
        for (int ta = 0; ta < 2; ++ta)
#pragma unroll
            for (int tb = 0; tb < 2; ++tb) acc[ta][tb] = mfma64(av[ta][s], bv[tb][s], acc[ta][tb]);
#pragma unroll
    for (int ta = 0; ta < 2; ++ta)
#pragma unroll
        for (int tb = 0; tb < 2; ++tb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + 16 * ta + 4 * r + lk, j = j0 + 16 * tb + l16;
                if (i < n && j <= i) Kb[(size_t)i * n + j] = acc[ta][tb][r];
            }
}

// grid = nd (pw + ut): the panel workgroups of every matrix first (they are
// dispatched first, and the next step waits on them), then the update tiles.
__global__ __launch_bounds__(256) void gn_chol_step_kernel(double* __restrict__ K, int n, int q,
                                                           int pw, int nd,
                                                           double* __restrict__ Dinv,
                                                           int* __restrict__ flags) {
    __shared__ double sm[kStepLds + 2 * kNb];
    const int x = blockIdx.x;
    if (x < pw * nd) {
        const int b = x / pw;
        double* Db = Dinv + ((size_t)b * ((n + kNb - 1) / kNb) + q / kNb) * kNb * kNb;
        chol_panel(K + (size_t)b * n * n, n, q, x - b * pw, Db, flags + b, sm);
    } else {
        const int ut = (gridDim.x - pw * nd) / nd;
        const int b = (x - pw * nd) / ut;
        chol_update(K + (size_t)b * n * n, n, q - kNb, q + kNb, x - pw * nd - b * ut);
    }
}

// ------------------------------------------------------------ substitution
// One workgroup of 1,024 threads per system: the substitutions are a chain
// of n / 32 dependent block steps, so each step issues all its global loads
// at once (32 per thread at most, per 1,024 columns) and pays one memory
// round trip and two barriers.
constexpr int kSolveThreads = 1024;

__device__ __forceinline__ double half_wave_sum(double v) {   // over 32 lanes
#pragma unroll
    for (int off = 16; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Workgroup barrier for the substitutions: orders the workgroup's LDS
// accesses (lgkmcnt(0), then s_barrier) and leaves global loads in flight,
// where __syncthreads()' release fence would wait for them too and so stall
// on the next block's prefetched rows. The memory clobber keeps the compiler
// from moving memory accesses across it.
constexpr int kWaitLgkm0 = 0xC07F;   // s_waitcnt lgkmcnt(0), vmcnt and expcnt at their maxima
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_s_waitcnt(kWaitLgkm0);
    asm volatile("s_barrier" ::: "memory");
}

// Forward substitution L y = rhs over rows q_begin .. n - 1 (earlier y are 0),
// y in LDS, rhs in y on entry, by blocks of kNb rows. For block q, the 32
// threads of row q + i (i = tid / 32) form its product with the solved part
// over consecutive columns; then y_q = Dinv_q (rhs_q minus those products),
// the 32 threads of row i over the block's columns. A block's rows on the
// first 1,024 columns and its Dinv entry are loaded while the block before
// it is solved, so the step waits on no memory round trip below 1,024
// columns.
__device__ void forward_subst(const double* __restrict__ L, const double* __restrict__ Dv, int n,
                              int q_begin, double* y, double* s_blk) {
    const int tid = threadIdx.x, i = tid >> 5, c = tid & 31;
    constexpr int kChunk = 32 * kNb;   // columns a row's 32 threads cover per pass
    double lv[kNb], dv;
    int lim = 0;   // the prefetched entries m with 32 m < lim are in the solved part
    auto prefetch = [&](int qq) {   // block qq's rows on columns q_begin .. q_begin + kChunk
        const int ww = min(kNb, n - qq);
        const bool rin = i < ww;
        const double* ptr = L + (size_t)(qq + (rin ? i : 0)) * n + q_begin + c;
        lim = rin ? qq - q_begin - c : 0;
#pragma unroll
        for (int m = 0; m < kNb; ++m) {
            lv[m] = 0.0;
            if (32 * m < lim) lv[m] = ptr[32 * m];
        }
        const double d = Dv[(size_t)(qq / kNb) * kNb * kNb + (rin ? i : 0) * kNb + c];
        dv = rin ? d : 0.0;
    };
    prefetch(q_begin);
    for (int q = q_begin; q < n; q += kNb) {
        const int w = min(kNb, n - q);
        double s = 0.0;
        const double* yc = y + q_begin + c;
#pragma unroll
        for (int m = 0; m < kNb; ++m)
            if (32 * m < lim) s += lv[m] * yc[32 * m];
        if (i < w) {   // columns past the first kChunk (n > kChunk)
            const double* row = L + (size_t)(q + i) * n;
#pragma unroll 8
            for (int k = q_begin + kChunk + c; k < q; k += 32) s += row[k] * y[k];
        }
        const double dq = dv;
        if (q + kNb < n) prefetch(q + kNb);
        s = half_wave_sum(s);
        if (c == 0 && i < w) s_blk[i] = y[q + i] - s;
        lds_barrier();
        double v = i < w ? dq * s_blk[c] : 0.0;   // Dinv is zero beyond the block
        v = half_wave_sum(v);
        if (c == 0 && i < w) y[q + i] = v;
        lds_barrier();
    }
}

// Backward substitution L^T x = y in place (x overwrites y), blocks from the
// last, right-looking: x_q = Dinv_q^T y_q (the 32 threads of row j over the
// block), then every earlier entry takes its share,
// y[j] -= sum_i L[q + i][j] x_q[i] for j < q (one column per thread, the
// block's rows of L read across threads). A block's rows on the thread's
// first column and its Dinv entry are loaded at the end of the block after
// it, so they arrive across the barrier and that block's Dinv step.
__device__ void backward_subst(const double* __restrict__ L, const double* __restrict__ Dv, int n,
                               double* y, double* s_blk) {
    const int tid = threadIdx.x, j = tid >> 5, c = tid & 31;
    const int nblk = (n + kNb - 1) / kNb;
    double lb[kNb], dv;
    auto prefetch = [&](int bq) {   // block bq's rows on column tid, and its Dinv entry
        const int q = bq * kNb, w = min(kNb, n - q);
        const double* ptr = L + (size_t)q * n + tid;
        const int lim = tid < q ? w : 0;   // rows ii < lim
#pragma unroll
        for (int ii = 0; ii < kNb; ++ii) {
            lb[ii] = 0.0;
            if (ii < lim) lb[ii] = ptr[(size_t)ii * n];
        }
        const bool din = j < w && c < w;
        const double d = Dv[(size_t)bq * kNb * kNb + (din ? c * kNb + j : 0)];
        dv = din ? d : 0.0;
    };
    prefetch(nblk - 1);
    for (int bq = nblk - 1; bq >= 0; --bq) {
        const int q = bq * kNb;
        const int w = min(kNb, n - q);
        double v = dv * y[q + (j < w && c < w ? c : 0)];   // dv is 0 outside the block
        v = half_wave_sum(v);
        if (c == 0 && j < w) s_blk[j] = v;
        lds_barrier();
        if (tid < w) y[q + tid] = s_blk[tid];
        if (tid < q) {
            double acc = 0.0;
#pragma unroll
            for (int ii = 0; ii < kNb; ++ii) acc += lb[ii] * s_blk[ii < w ? ii : 0];
            y[tid] -= acc;
        }
        for (int jj = tid + kSolveThreads; jj < q; jj += kSolveThreads) {
            double acc = 0.0;
#pragma unroll 8
            for (int ii = 0; ii < w; ++ii) acc += L[(size_t)(q + ii) * n + jj] * s_blk[ii];
            y[jj] -= acc;
        }
        if (bq > 0) prefetch(bq - 1);   // in flight across the barrier and the next Dinv step
        lds_barrier();
    }
}

// the substitutions keep the solution vector in LDS, within the 64 KB a
// workgroup may take (with a kNb-entry scratch vector)
constexpr int kSolveMaxN = 7680;

// B[b] = bvec + x_b with K_b x_b = rhs, or bvec when K_b is singular. grid = nd.
__global__ __launch_bounds__(kSolveThreads) void gn_solve_kernel(const double* __restrict__ K,
                                                       const double* __restrict__ Dinv, int n,
                                                       const int* __restrict__ flags,
                                                       const double* __restrict__ rhs,
                                                       const double* __restrict__ bvec,
                                                       double* __restrict__ B) {
    extern __shared__ double sm[];
    double* y = sm;                    // [n]
    double* s_blk = sm + n;            // [kNb]
    const int nblk = (n + kNb - 1) / kNb;
    const double* L = K + (size_t)blockIdx.x * n * n;
    const double* Dv = Dinv + (size_t)blockIdx.x * nblk * kNb * kNb;
    for (int i = threadIdx.x; i < n; i += kSolveThreads) y[i] = rhs[i];
    __syncthreads();
    forward_subst(L, Dv, n, 0, y, s_blk);
    backward_subst(L, Dv, n, y, s_blk);
    const bool sing = flags[blockIdx.x] != 0;
    for (int i = threadIdx.x; i < n; i += kSolveThreads)
        B[(size_t)blockIdx.x * n + i] = bvec[i] + (sing ? 0.0 : y[i]);
}

// out[i] = (K^-1)_ii = |L^-1 e_i|^2, or 0 when K is singular. grid = n.
__global__ __launch_bounds__(kSolveThreads) void gn_invdiag_kernel(const double* __restrict__ L,
                                                         const double* __restrict__ Dinv, int n,
                                                         const int* __restrict__ flags,
                                                         double* __restrict__ out) {
    extern __shared__ double sm[];
    double* y = sm;
    double* s_blk = sm + n;
    const int i = blockIdx.x;
    const int q0 = i / kNb * kNb;
    for (int k = q0 + threadIdx.x; k < n; k += kSolveThreads) y[k] = k == i ? 1.0 : 0.0;
    __syncthreads();
    forward_subst(L, Dinv, n, q0, y, s_blk);
    double s = 0.0;
    for (int k = i + threadIdx.x; k < n; k += kSolveThreads) s += y[k] * y[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    __shared__ double wsum[kSolveThreads / 64];
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int v = 0; v < kSolveThreads / 64; ++v) t += wsum[v];
        out[i] = flags[0] ? 0.0 : t;
    }
}

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

int syrk_tiles(int n) {
    const int nb = (n + kSyrkTile - 1) / kSyrkTile;
    return nb * (nb + 1) / 2;
}

// K-slices of the normal-equation product: a multiple of the 8 XCDs, enough
// for about 4,096 workgroups and at most 64; fewer (no XCD mapping) when J
// has fewer 16-row strips than that. 4 workgroups fit a CU, so a slice's
// tiles run in rounds of 128 per XCD; more, shorter slices shrink the last,
// partly filled round (at 1,024 columns: 136 tiles per slice, 0.98 ms with
// 8 slices, 0.84 ms with 32).
int syrk_slices(int tiles, int64_t rows) {
    const int64_t strips = (rows + kSyrkRows - 1) / kSyrkRows;
    const int per_xcd = std::max(1, std::min(8, (4096 + kXcds * tiles - 1) / (kXcds * tiles)));
    const int ks = kXcds * per_xcd;
    return (int)std::max<int64_t>(1, std::min<int64_t>(ks, strips));
}

size_t factor_bytes(int n, int nd) {
    const size_t nblk = (size_t)(n + kNb - 1) / kNb;
    return align256(sizeof(double) * (size_t)nd * n * n) +
           align256(sizeof(double) * (size_t)nd * nblk * kNb * kNb) +
           align256(sizeof(int) * (size_t)nd);
}

// Factor nd matrices K [nd][n][n] in place (lower triangle), the inverses of
// the diagonal blocks into Dinv: one launch per 32-column block (the panel,
// with the previous block's update of its columns, beside the rest of that
// update).
int cholesky_batched(double* K, double* Dinv, int n, int nd, int* flags, hipStream_t stream) {
    for (int q = 0; q < n; q += kNb) {
        const int w = std::min(kNb, n - q);
        const int below = n - q - w;
        const int pw = std::max(1, (below + kPanelRows - 1) / kPanelRows);
        int ut = 0;   // the previous block's update tiles, rows and columns >= q + 32
        if (q > 0 && n > q + kNb) {
            const int nt = (n - q - kNb + 63) / 64;
            ut = nt * (nt + 1) / 2;
        }
        const int64_t blocks = (int64_t)(pw + ut) * nd;
        if (blocks >= ((int64_t)1 << 31))
            return fail(ZTM_EINVAL, "Cholesky: %d matrices of %d need too many workgroups", nd, n);
        hipLaunchKernelGGL(gn_chol_step_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, K, n,
                           q, pw, nd, Dinv, flags);
        ZTM_HIP(hipGetLastError());
    }
    return ZTM_OK;
}

}  // namespace

size_t gn_workspace_bytes(int64_t rows, int n, int nd) {
    const size_t ks = (size_t)syrk_slices(syrk_tiles(n), rows);
    const size_t syrk = align256(sizeof(double) * ks * n * n) + sizeof(double) * ks * n;
    return std::max(syrk, factor_bytes(n, std::max(nd, 1)));
}

int gn_normal_launch(int64_t rows, int n, const double* J, const double* r, double* JJ,
                     double* jtr, void* work, hipStream_t stream) {
    const int nb = (n + kSyrkTile - 1) / kSyrkTile;
    const int tiles = nb * (nb + 1) / 2;
    const int ks = syrk_slices(tiles, rows);
    const int64_t strips = (rows + kSyrkRows - 1) / kSyrkRows;
    const int64_t rps = (strips + ks - 1) / ks * kSyrkRows;   // whole strips per slice
    double* part = (double*)work;
    double* jpart = (double*)((char*)work + align256(sizeof(double) * (size_t)ks * n * n));
    if ((int64_t)tiles * ks >= ((int64_t)1 << 31))
        return fail(ZTM_EINVAL, "normal equations: %d parameters need too many tiles", n);
    {
        ProfScope prof("gn_syrk", stream);
        hipLaunchKernelGGL(gn_syrk_kernel, dim3((unsigned)(tiles * ks)), dim3(256), 0, stream, J,
                           r, n, rows, rps, nb, tiles, ks % kXcds == 0 ? 1 : 0, part, jpart);
        ZTM_HIP(hipGetLastError());
    }
    const int nr = (n + kRedTile - 1) / kRedTile;
    const int rtiles = nr * (nr + 1) / 2;
    const int jblocks = r ? (n + 255) / 256 : 0;
    hipLaunchKernelGGL(gn_syrk_reduce_kernel, dim3(rtiles + jblocks), dim3(256), 0, stream, part,
                       jpart, ks, n, nr, rtiles, JJ, r ? jtr : nullptr);
    ZTM_HIP(hipGetLastError());
    return ZTM_OK;
}

int gn_steps_launch(int n, int nd, const double* JJ, const double* jtr, const double* sig,
                    const double* dm, const double* b, double* B, void* work,
                    hipStream_t stream) {
    if (n > kSolveMaxN)
        return fail(ZTM_EINVAL, "the damped steps support up to %d parameters (got %d)",
                    kSolveMaxN, n);
    const size_t nblk = (size_t)(n + kNb - 1) / kNb;
    double* K = (double*)work;
    double* Dinv = (double*)((char*)work + align256(sizeof(double) * (size_t)nd * n * n));
    int* flags = (int*)((char*)Dinv + align256(sizeof(double) * (size_t)nd * nblk * kNb * kNb));
    const int64_t total = (int64_t)nd * n * n;
    hipLaunchKernelGGL(gn_form_kernel, dim3((unsigned)std::min<int64_t>((total + 255) / 256,
                                                                         1 << 16)),
                       dim3(256), 0, stream, JJ, n, nd, 0, sig, dm, 0.0, K, flags);
    ZTM_HIP(hipGetLastError());
    {
        ProfScope prof("gn_cholesky", stream);
        if (int rc = cholesky_batched(K, Dinv, n, nd, flags, stream)) return rc;
    }
    const size_t lds = sizeof(double) * ((size_t)n + kNb);
    ProfScope prof("gn_solve", stream);
    hipLaunchKernelGGL(gn_solve_kernel, dim3(nd), dim3(kSolveThreads), lds, stream, K, Dinv, n, flags, jtr,
                       b, B);
    ZTM_HIP(hipGetLastError());
    return ZTM_OK;
}

int gn_inverse_diag_launch(int n, const double* JJ, double d, double* out, void* work,
                           hipStream_t stream) {
    if (n > kSolveMaxN)
        return fail(ZTM_EINVAL, "the inverse diagonal supports up to %d parameters (got %d)",
                    kSolveMaxN, n);
    const size_t nblk = (size_t)(n + kNb - 1) / kNb;
    double* K = (double*)work;
    double* Dinv = (double*)((char*)work + align256(sizeof(double) * (size_t)n * n));
    int* flags = (int*)((char*)Dinv + align256(sizeof(double) * nblk * kNb * kNb));
    const int64_t total = (int64_t)n * n;
    hipLaunchKernelGGL(gn_form_kernel, dim3((unsigned)std::min<int64_t>((total + 255) / 256,
                                                                         1 << 16)),
                       dim3(256), 0, stream, JJ, n, 1, 1, nullptr, nullptr, d, K, flags);
    ZTM_HIP(hipGetLastError());
    if (int rc = cholesky_batched(K, Dinv, n, 1, flags, stream)) return rc;
    const size_t lds = sizeof(double) * ((size_t)n + kNb);
    ProfScope prof("gn_invdiag", stream);
    hipLaunchKernelGGL(gn_invdiag_kernel, dim3(n), dim3(kSolveThreads), lds, stream, K, Dinv, n, flags,
                       out);
    ZTM_HIP(hipGetLastError());
    return ZTM_OK;
}

}  // namespace ztm
