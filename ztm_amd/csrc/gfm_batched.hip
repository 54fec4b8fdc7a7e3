// Batched (ensemble) forward and adjoint on MFMA: Y = G P for P = [B models x
// N cells] and X = G^T R for R = [B residual vectors x M stations], with G
// (stations x cells) generated on the fly -- never stored. (Up to 4 models
// or vectors take the FP32 point-mass kernels instead; see gfm_abi.hip.)
//
// Split-BF16 scheme: G and P are each split into BF16 parts, x = x0 + x1 (+ x2)
// with x0 = bf16(x), x1 = bf16(x - x0), x2 = bf16(x - x0 - x1), and the product
// keeps the terms down to the split's precision on v_mfma_f32_32x32x16_bf16
// with FP32 accumulation:
//   2-way (hi/lo, ~2^-16):  x0 y0 + x0 y1 + x1 y0                 (3 MFMAs)
//   3-way (~2^-24):         + x0 y2 + x1 y1 + x2 y0               (6 MFMAs)
// Split-FP16 (split code 4): hi/lo FP16 parts (~2^-22) on
// v_mfma_f32_32x32x16_f16, 3 MFMAs. FP16 stops at 65504, so the operands
// carry powers of two that bring their largest magnitudes to 2^14: each
// model or residual vector its own (from a device absmax before the pack),
// G in the forward one per station (the plan's per-station bound on |G|)
// and in the adjoint one per plan. The reduce takes them out in FP64.
// Each K-slice's FP32 accumulators are written out and the slices are summed
// in FP64, which bounds the FP32 accumulation run to one slice (a few hundred
// MFMA steps by default).
//
// Work decomposition: a workgroup = 4 waves = 128 stations (32 per wave, the
// MFMA's M); each wave keeps all B models' accumulators (B/32 tiles of 32x32,
// 16 floats per lane each). The K loop walks the slice's cells 16 at a time:
// the workgroup stages the chunk of packed P (hi and lo, B x 16 cells) in LDS
// once, every wave generates its own 32 x 16 block of G in registers directly
// in the MFMA A-operand layout (lane: station lane%32, cells 8*(lane/32)+0..7),
// converts it to hi/lo and issues 3 MFMAs per model tile. Workgroups of one
// K-slice differ only in stations; with a slice count that is a multiple of
// 8 all the station tiles of a slice run on one XCD (fwd_tile_slice), so
// that its L2 fetches the slice's P chunks once for all of them.
//
// Layouts (HBM):  P parts [N16/16][2][Bp][8] bf16 each (chunk, cell half,
//                 model, cell), zero padded (N16 = N up to a multiple of 16,
//                 Bp = B up to 32 x a power of two), so that the B operands
//                 of one half-wave are consecutive 16-byte slots in LDS;
//                 part [KS][m_pad][Bp] FP32;  Y [B][M] FP64 (model-major).
//                 The adjoint packs R the same way over stations, and writes
//                 X [B][N] FP32 directly (one K-slice, BF16) or through part
//                 [KS][n_pad][Bp] and the reduce (one K-slice with FP16:
//                 the transpose). Split 4 keeps the models' absmax bits
//                 [256] after the partials.
#include "gfm_internal.h"
#include "gfm_pair.h"

#include <algorithm>
#include <cmath>
#include <type_traits>

namespace ztm {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// Element type of the split parts: BF16 (8 significant bits, the FP32 range)
// or FP16 (11 significant bits below 65504; used with power-of-two scales
// that bring every operand's largest magnitude to at most 2^kHalfTop).
template <typename H> struct Half;
template <> struct Half<__bf16> {
    using v8 = bf16x8;
    static constexpr bool scaled = false;
    static __device__ __forceinline__ f32x16 mfma(v8 a, v8 b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
};
template <> struct Half<_Float16> {
    using v8 = f16x8;
    static constexpr bool scaled = true;
    static __device__ __forceinline__ f32x16 mfma(v8 a, v8 b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
};
constexpr int kHalfTop = 14;

// Power-of-two factor that brings a largest magnitude a (given by its float
// bits) to at most 2^kHalfTop; 1 for a = 0.
// The exponent is clamped to +-120, as half_g_scale clamps it, so tiny
// maxima (below 2^-106) give a finite scale rather than 2^(14 - e) = inf.
__device__ __forceinline__ float half_scale(unsigned abits) {
    const float a = __uint_as_float(abits);
    if (!(a > 0.f)) return 1.f;
    int e;
    frexpf(a, &e);   // a <= 2^e
    return ldexpf(1.f, max(-120, min(120, kHalfTop - e)));
}

// amax[b] |= bits of max over n of |X[b][n]| among entries with |X| > thr
// (non-negative floats order like their bits); amax must start at 0.
// grid = (blocks, B); block = 256.
__global__ __launch_bounds__(256) void batched_absmax_kernel(const float* __restrict__ X,
                                                             int64_t n, float thr,
                                                             unsigned* __restrict__ amax) {
    const int b = blockIdx.y;
    const float* x = X + (size_t)b * n;
    float m = 0.f;
    auto take = [&](float v) {
        v = fabsf(v);
        if (v > thr) m = fmaxf(m, v);
    };
    int64_t i0 = 0;
    if (((uintptr_t)x & 15) == 0) {   // 16-byte loads over the whole float4s of the row
        const float4* x4 = reinterpret_cast<const float4*>(x);
        const int64_t n4 = n >> 2;
        for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n4;
             i += (int64_t)gridDim.x * 256) {
            const float4 v = x4[i];
            take(v.x); take(v.y); take(v.z); take(v.w);
        }
        i0 = n4 << 2;
    }
    for (int64_t i = i0 + (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * 256)
        take(x[i]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    __shared__ float wm[4];
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)   // one atomic per block
        atomicMax(&amax[b], __float_as_uint(fmaxf(fmaxf(wm[0], wm[1]), fmaxf(wm[2], wm[3]))));
}

__global__ void batched_clear_kernel(unsigned* __restrict__ x, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) x[i] = 0u;
}

constexpr int kBatchChunk = 16;   // cells per MFMA K step
// s_waitcnt immediate (gfx9 encoding): vmcnt(0), expcnt and lgkmcnt left at
// their maxima, i.e. wait only for this wave's vector memory operations.
constexpr int kWaitVm0 = 0x0F70;
constexpr int kBatchTile = 128;   // stations per workgroup

// Station tile and K-slice of a forward workgroup, grid = (tiles, slices).
// Workgroups are dispatched in grid order (x fastest) and go to the 8 XCDs
// round-robin. With a slice count that is a multiple of 8, the workgroup at
// dispatch position w takes slice (w % 8) + 8 (w / 8 / tiles) and tile
// (w / 8) % tiles: every tile of a slice runs on one XCD, whose L2 then
// fetches the slice's P parts once for all of them (otherwise each XCD
// fetches every slice). Same work per workgroup either way.
constexpr int kXcds = 8;
__device__ __forceinline__ void fwd_tile_slice(int& tile, int& slice) {
    const int tiles = gridDim.x;
    if (gridDim.y % kXcds == 0) {
        const int64_t w = blockIdx.x + (int64_t)tiles * blockIdx.y;
        const int64_t idx = w / kXcds;
        slice = (int)(w % kXcds + kXcds * (idx / tiles));
        tile = (int)(idx % tiles);
    } else {
        tile = blockIdx.x;
        slice = blockIdx.y;
    }
}

// Workgroup barrier without __syncthreads()'s release fence: that fence makes
// the compiler wait for every outstanding vector memory operation
// (vmcnt(0)) before the barrier, which drains the loads a prefetch ring keeps
// in flight for later chunks. The caller first waits (s_waitcnt) for the
// loads the other waves need next; the memory clobber keeps the compiler from
// moving memory accesses across the barrier.
__device__ __forceinline__ void ring_barrier() { asm volatile("s_barrier" ::: "memory"); }
// the adjoint's station chunks end inside the padded station records
static_assert(kStationPad % 16 == 0, "station padding must hold whole chunks");

// ---------------------------------------------------------------- packing
// A block packs 32 models x 64 cells (4 chunks): it reads the tile with
// 256-byte row reads into LDS (16-byte loads when vec4: n_cells a multiple of
// 4 and P 16-byte aligned, so no load crosses a row's end), then each thread
// converts the 8 cells of one (chunk, half, model) slot and writes them as one
// 16-byte store per part, contiguous across the models of a slot.
// grid = (ceil(n_chunks / 4), Bp / 32). Element v (times mul, and times the
// model's half_scale(amax[model]) when amax is given, formed once per model
// and block) is split as h_1 = H(v), h_2 = H(v - h_1), ...; entries outside
// the model or cell range, and |v| <= thr, are 0.
constexpr int kPackCells = 64;

template <int SPLIT, typename H>
__global__ __launch_bounds__(256) void batched_pack_kernel(const float* __restrict__ P, int B,
                                                           int64_t n_cells, int Bp,
                                                           int64_t n_chunks, int64_t n_packed,
                                                           float thr, float mul,
                                                           const unsigned* __restrict__ amax,
                                                           int vec4, H* __restrict__ parts) {
    using v8 = typename Half<H>::v8;
    __shared__ float tile[32][kPackCells + 1];
    __shared__ float fm[32];   // each model's factor
    const int t = threadIdx.x;
    const int64_t c0 = (int64_t)blockIdx.x * kPackCells;
    const int m0 = blockIdx.y * 32;
    if (t < 32) fm[t] = amax && m0 + t < B ? mul * half_scale(amax[m0 + t]) : mul;
    __syncthreads();
    if (vec4) {
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int idx = i * 256 + t;
            const int row = idx >> 4, col = 4 * (idx & 15);
            const int model = m0 + row;
            const int64_t cell = c0 + col;
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
            if (model < B && cell < n_cells) {
                v = *reinterpret_cast<const float4*>(P + (size_t)model * n_cells + cell);
                const float f = fm[row];
                v.x = fabsf(v.x) > thr ? v.x * f : 0.f;
                v.y = fabsf(v.y) > thr ? v.y * f : 0.f;
                v.z = fabsf(v.z) > thr ? v.z * f : 0.f;
                v.w = fabsf(v.w) > thr ? v.w * f : 0.f;
            }
            tile[row][col] = v.x;
            tile[row][col + 1] = v.y;
            tile[row][col + 2] = v.z;
            tile[row][col + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int idx = i * 256 + t;
            const int row = idx >> 6, col = idx & 63;
            const int model = m0 + row;
            const int64_t cell = c0 + col;
            float v = 0.f;
            if (model < B && cell < n_cells) {
                v = P[(size_t)model * n_cells + cell];
                v = fabsf(v) > thr ? v * fm[row] : 0.f;
            }
            tile[row][col] = v;
        }
    }
    __syncthreads();
    const int slot = t >> 5;          // 8 slots: chunk (c0/16 + slot/2), half slot%2
    const int mrow = t & 31;
    const int64_t chunk = c0 / kBatchChunk + (slot >> 1);
    if (chunk >= n_chunks) return;
    v8 h[SPLIT];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        float v = tile[mrow][slot * 8 + e];
#pragma unroll
        for (int q = 0; q < SPLIT; ++q) {
            const H b = (H)v;
            h[q][e] = b;
            v -= (float)b;
        }
    }
    const int64_t off = ((chunk * 2 + (slot & 1)) * Bp + m0 + mrow) * 8;
#pragma unroll
    for (int q = 0; q < SPLIT; ++q)
        *reinterpret_cast<v8*>(parts + q * n_packed + off) = h[q];
}

// --------------------------------------------------------------- forward
// q = n / d, r = n % d for 0 <= n < 2^31 through a double reciprocal and one
// correction step (the product's error is far below one unit for n < 2^31).
__device__ __forceinline__ void divmod_pos(int n, int d, double inv, int& q, int& r) {
    q = (int)((double)n * inv);
    r = n - q * d;
    const bool lo = r < 0, hi = r >= d;
    q += hi ? 1 : (lo ? -1 : 0);
    r += hi ? -d : (lo ? d : 0);
}

// One G value into element e of the SPLIT parts. FP16 parts take the value
// times the power-of-two G scale gs (already applied when PRE), held to
// +-2^15 when CLAMP: the adjoint's padding station records, whose |G| the
// scale does not bound, then stay finite, so that their products with the
// zero-padded residuals stay 0. The forward needs no clamp: its per-station
// bounds cover the cells that pad the grid, and a padding station's row
// only reaches its own (discarded) outputs. The empty asm makes the parts
// come from the rounded FP32 value through v_cvt_f16_f32: left to itself the
// compiler folds the product into v_fma_mix*_f16 conversions, which measured
// ~2e-5 errors where these give ~1e-6.
template <int SPLIT, typename H, bool PRE = false, bool CLAMP = true>
__device__ __forceinline__ void store_split(float gv, float gs,
                                            typename Half<H>::v8 (&a)[SPLIT], int e) {
    if constexpr (Half<H>::scaled) {
        if constexpr (!PRE) gv *= gs;
        if constexpr (CLAMP) gv = __builtin_amdgcn_fmed3f(gv, -32768.f, 32768.f);
        asm volatile("" : "+v"(gv));
    }
#pragma unroll
    for (int q = 0; q < SPLIT; ++q) {
        const H h = (H)gv;
        a[q][e] = h;
        gv -= (float)h;
    }
}

// This lane's 8 G values of a chunk (cells cell0 .. cell0 + 7, raster order),
// split into SPLIT BF16 parts in the MFMA A-operand layout: begin() forms
// the chunk's shared terms, elem<E>() the value of cell cell0 + E, so that
// the kernel can spread the 8 elements between its MFMAs. MODE 1 (nx >= 8):
// the 8 cells span at most two grid rows, so both row terms are formed once
// and each cell selects one. MODE 2 (nx a multiple of 8): the 8 cells lie in
// one row (cell0 is a multiple of 8), so there is nothing to select; for such
// grids it gives the same values as MODE 1. MODE 0: cell-by-cell stepping
// (elements in increasing E). Cells past the grid give finite values
// (r^2 >= 1e-20, so |dz r^-3| <= 1e36), which the zero-padded P cancels.
template <int SPLIT, int MODE, typename H = __bf16>
struct BatchedGen {
    using v8 = typename Half<H>::v8;
    float gs = 1.f;                    // G scale (FP16 parts)
    float xa, R0, dz0, R1, dz1, nxf;   // MODE 1, 2
    int ew;                            // MODE 1: cells e >= ew lie on the next row
    float px, py, pz, sy, sz;          // MODE 0
    int i, j, k, nx, ny;               // MODE 0

    __device__ __forceinline__ void begin(const float4 p, int cell0, int nx_, int ny_,
                                          double inv_nx, double inv_ny, float sy_, float sz_) {
        int rw, i0, k0, j0;
        divmod_pos(cell0, nx_, inv_nx, rw, i0);
        divmod_pos(rw, ny_, inv_ny, k0, j0);
        begin_at(p, i0, j0, k0, nx_, ny_, sy_, sz_);
    }

    // The same from the grid indices (i0, j0, k0) of the chunk's first cell.
    __device__ __forceinline__ void begin_at(const float4 p, int i0, int j0, int k0, int nx_,
                                             int ny_, float sy_, float sz_) {
        if constexpr (MODE == 0) {
            px = p.x; py = p.y; pz = p.z; sy = sy_; sz = sz_;
            i = i0; j = j0; k = k0; nx = nx_; ny = ny_;
        } else {
            dz0 = p.z - (float)k0 * sz_;
            const float dy0 = p.y - (float)j0 * sy_;
            // no near-field pass follows, so the exact row term (floored only
            // against 0/0 at a coincident station), not row_R32
            R0 = fmaf(dy0, dy0, fmaf(dz0, dz0, 1e-20f));
            xa = p.x - (float)i0;
            if constexpr (MODE == 1) {
                const bool wrap = j0 + 1 == ny_;
                const int j1 = wrap ? 0 : j0 + 1, k1 = wrap ? k0 + 1 : k0;
                const float dy1 = p.y - (float)j1 * sy_;
                dz1 = p.z - (float)k1 * sz_;
                R1 = fmaf(dy1, dy1, fmaf(dz1, dz1, 1e-20f));
                ew = nx_ - i0;
                nxf = (float)nx_;
                if constexpr (Half<H>::scaled) dz1 *= gs;
            }
            if constexpr (Half<H>::scaled) dz0 *= gs;   // G scale folded into the row terms
        }
    }

    // Empty asm operands that tie the chunk's shared terms to this point of
    // the instruction stream, so that an element computed from them cannot
    // be hoisted above it.
    __device__ __forceinline__ void pin() {
        if constexpr (MODE == 2) {
            asm volatile("" : "+v"(xa), "+v"(R0), "+v"(dz0));
        } else if constexpr (MODE == 1) {
            asm volatile("" : "+v"(xa), "+v"(R0), "+v"(dz0), "+v"(R1), "+v"(dz1));
        } else {
            asm volatile("" : "+v"(px), "+v"(py), "+v"(pz));
        }
    }

    template <int E>
    __device__ __forceinline__ void elem(v8 (&a)[SPLIT]) {
        float gv;
        if constexpr (MODE == 2) {
            const float x = xa - (float)E;
            const float v = rsq_f(fmaf(x, x, R0));
            gv = dz0 * (v * v * v);
        } else if constexpr (MODE == 1) {
            const bool w = E >= ew;
            const float xe = xa - (float)E;
            const float x = w ? xe + nxf : xe;
            const float v = rsq_f(fmaf(x, x, w ? R1 : R0));
            gv = (w ? dz1 : dz0) * (v * v * v);
        } else {
            const float dy = py - (float)j * sy;
            const float dz = pz - (float)k * sz;
            const float x = px - (float)i;
            const float v = rsq_f(fmaf(x, x, fmaf(dy, dy, fmaf(dz, dz, 1e-20f))));
            gv = dz * (v * v * v);
            const bool wi = ++i == nx;
            i = wi ? 0 : i;
            j += wi ? 1 : 0;
            const bool wj = j == ny;
            j = wj ? 0 : j;
            k += wj ? 1 : 0;
        }
        store_split<SPLIT, H, MODE != 0, false>(gv, gs, a, E);
    }
};

template <int E, int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (E < N) {
        f(std::integral_constant<int, E>{});
        static_for<E + 1, N>(f);
    }
}

template <int SPLIT, int MODE, typename H>
__device__ __forceinline__ void batched_gen(const float4 p, int cell0, int nx, int ny,
                                            double inv_nx, double inv_ny, float sy, float sz,
                                            float gs, typename Half<H>::v8 (&a)[SPLIT]) {
    BatchedGen<SPLIT, MODE, H> g;
    g.gs = gs;
    g.begin(p, cell0, nx, ny, inv_nx, inv_ny, sy, sz);
    static_for<0, 8>([&](auto ec) { g.template elem<decltype(ec)::value>(a); });
}

// Issue order within a chunk: after MFMA m (of the chunk's TOTAL) come the G
// elements e with e * TOTAL / 8 == m (integer division: the 8 elements of
// the next chunk's G block spread evenly over the MFMAs), fenced by
// scheduling barriers, so that the wave issues independent VALU work while
// each MFMA runs.
template <int TOTAL, typename F>
__device__ __forceinline__ void gen_after_mfma(int m, F&& elem) {
    static_for<0, 8>([&](auto ec) {
        constexpr int e = decltype(ec)::value;
        if ((e * TOTAL) / 8 == m) elem(ec);
    });
}

// grid = (station tiles of 128, K-slices); block = 256 (4 waves x 32 stations).
// Software pipeline over the slice's chunks, one barrier per chunk: while the
// MFMAs of chunk c run on its G block (registers) and P parts (LDS buffer
// c&1), the next chunk's P parts load from global memory straight into the
// other LDS buffer and its G block is generated. The last iteration repeats
// its own chunk's loads into the buffer that is not read again.
// H: part type (BF16, or FP16 with G's rows scaled by the stations' power-of-
// two scales gsc[m_pad]).
template <int NT, int SPLIT, int MODE, typename H>   // model tiles of 32 (B <= 32 * NT); 2- or 3-way split; batched_gen mode
__global__ __launch_bounds__(256) void batched_fwd_kernel(
    const float4* __restrict__ st, const typename Half<H>::v8* __restrict__ parts,
    int64_t part_stride, int nx, int ny, double inv_nx, double inv_ny, float sy, float sz,
    const float* __restrict__ gsc, int n_chunks, int chunks_per_slice, int m_pad,
    float* __restrict__ part) {
    using v8 = typename Half<H>::v8;
    constexpr int Bp = 32 * NT;
    constexpr int V = Bp * 2;                          // 16-byte slots per part per chunk
    constexpr int LOADS = (V * SPLIT + 255) / 256;     // per thread per chunk
    __shared__ v8 lp[2][SPLIT][V];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int row = lane & 31, kh = lane >> 5;
    int tile, slice;
    fwd_tile_slice(tile, slice);
    const int mbase = tile * kBatchTile + wave * 32;
    const float4 p = st[mbase + row];
    const float gs = Half<H>::scaled ? gsc[mbase + row] : 1.f;

    f32x16 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int i = 0; i < 16; ++i) acc[t][i] = 0.f;

    const int c0 = slice * chunks_per_slice;
    const int c1 = min(c0 + chunks_per_slice, n_chunks);
    // P parts of a chunk go from global memory straight into LDS buffer buf
    // (global_load_lds_dwordx4: lane l of a wave writes 16 bytes at the wave's
    // LDS base + 16 l). Element idx = q V + v of the flattened [SPLIT][V]
    // buffer is part q, slot v; a wave's 64 lanes cover 64 consecutive idx,
    // and V * SPLIT is a multiple of 64, so the guard is wave-uniform.
    auto fetch = [&](int chunk, int buf) {
#pragma unroll
        for (int u = 0; u < LOADS; ++u) {
            const int idx = threadIdx.x + 256 * u;
            if ((V * SPLIT) % 256 == 0 || idx < V * SPLIT)
                __builtin_amdgcn_global_load_lds(
                    (const void*)(parts + (idx / V) * part_stride + (int64_t)chunk * V + (idx % V)),
                    (__attribute__((address_space(3))) void*)(&lp[buf][0][0] + (idx - lane)),
                    16, 0, 0);
        }
    };
    v8 a[SPLIT];
    if (c0 < c1) {
        fetch(c0, 0);
        batched_gen<SPLIT, MODE, H>(p, c0 * kBatchChunk + 8 * kh, nx, ny, inv_nx, inv_ny, sy, sz,
                                    gs, a);
    }
    __builtin_amdgcn_s_waitcnt(kWaitVm0);   // this wave's LDS loads have landed
    __syncthreads();
    for (int c = c0; c < c1; ++c) {
        const int buf = (c - c0) & 1;
        const int cn = min(c + 1, c1 - 1);
        fetch(cn, buf ^ 1);   // buffer buf ^ 1 was last read before the previous barrier
        v8 cur[SPLIT];
#pragma unroll
        for (int q = 0; q < SPLIT; ++q) cur[q] = a[q];
        BatchedGen<SPLIT, MODE, H> g;
        g.gs = gs;
        g.begin(p, cn * kBatchChunk + 8 * kh, nx, ny, inv_nx, inv_ny, sy, sz);
        // Tiles in pairs so that consecutive MFMAs update different
        // accumulators (no back-to-back dependence on one accumulator).
        // The B operands of the next tile pair are read from LDS while the
        // current pair's MFMAs run (two register sets); the next chunk's G
        // elements go between the MFMAs.
        constexpr int TP = NT >= 2 ? 2 : 1;
        constexpr int kPerPair = TP * SPLIT * (SPLIT + 1) / 2;   // MFMAs per tile pair
        constexpr int TOTAL = NT / TP * kPerPair;
        v8 b[2][TP][SPLIT];
        auto load_b = [&](int t, v8 (&bb)[TP][SPLIT]) {
#pragma unroll
            for (int u = 0; u < TP; ++u)
#pragma unroll
                for (int q = 0; q < SPLIT; ++q) bb[u][q] = lp[buf][q][kh * Bp + 32 * (t + u) + row];
        };
        load_b(0, b[0]);
#pragma unroll
        for (int t = 0; t < NT; t += TP) {
            const int cb = (t / TP) & 1;
            if (t + TP < NT) load_b(t + TP, b[cb ^ 1]);
            // products x_qa * y_qb with qa + qb < SPLIT
#pragma unroll
            for (int s2 = 0; s2 < SPLIT; ++s2)
#pragma unroll
                for (int qa = 0; qa <= s2; ++qa)
#pragma unroll
                    for (int u = 0; u < TP; ++u) {
                        acc[t + u] = Half<H>::mfma(cur[qa], b[cb][u][s2 - qa], acc[t + u]);
                        const int m = t / TP * kPerPair + (s2 * (s2 + 1) / 2 + qa) * TP + u;
                        gen_after_mfma<TOTAL>(m, [&](auto ec) {
                            g.pin();
                            g.template elem<decltype(ec)::value>(a);
#pragma unroll
                            for (int q = 0; q < SPLIT; ++q) asm volatile("" : "+v"(a[q]));
                        });
                    }
        }
        __builtin_amdgcn_s_waitcnt(kWaitVm0);
        __syncthreads();
    }
    // D layout: lane holds model 32t + lane%32, stations 8*(i/4) + 4*(lane/32) + i%4
    float* out = part + ((size_t)slice * m_pad + mbase) * Bp;
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const int srow = 8 * (e >> 2) + 4 * kh + (e & 3);
            out[(size_t)srow * Bp + 32 * t + row] = acc[t][e];
        }
}

// Variant 1 of the forward for the two-part splits (FP16 hi/lo, BF16 2-way):
// the same work and results as batched_fwd_kernel, another issue order. All
// 2 NT B operands of a chunk are read from LDS at the top of the iteration
// (the kernel may take up to 256 registers: 2 waves per SIMD), the MFMAs go
// term by term over the model tiles (hi x hi on every tile, then hi x lo, then
// lo x hi), so that consecutive MFMAs update different accumulators and none
// waits on an LDS read, and scheduling groups place the next chunk's G
// elements (VALU and the rsq, a transcendental) between the MFMAs.
// PF: chunks of P in flight ahead of the one the MFMAs read (1: two LDS
// buffers, the next chunk loads during this one; 2: three buffers, the loads
// run two chunks ahead and the end-of-chunk wait covers only the older ones).
template <int NT, int MODE, typename H, int WV, int PF = 1>   // WV waves (32 WV stations) per workgroup
__global__ __launch_bounds__(64 * WV) __attribute__((amdgpu_waves_per_eu(2, 2)))
void batched_fwd2_kernel(const float4* __restrict__ st, const typename Half<H>::v8* __restrict__ parts,
                         int64_t part_stride, int nx, int ny, double inv_nx, double inv_ny,
                         float sy, float sz, const float* __restrict__ gsc, int n_chunks,
                         int chunks_per_slice, int m_pad, float* __restrict__ part) {
    using v8 = typename Half<H>::v8;
    constexpr int SPLIT = 2;
    constexpr int Bp = 32 * NT;
    constexpr int V = Bp * 2;
    constexpr int LOADS = (V * SPLIT + 64 * WV - 1) / (64 * WV);
    constexpr int NBUF = PF + 1;
    __shared__ v8 lp[NBUF][SPLIT][V];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // uniform: SALU math
    const int lane = threadIdx.x & 63;
    const int row = lane & 31, kh = lane >> 5;
    int tile, slice;
    fwd_tile_slice(tile, slice);
    const int mbase = tile * (32 * WV) + wave * 32;
    const float4 p = st[mbase + row];
    const float gs = Half<H>::scaled ? gsc[mbase + row] : 1.f;

    f32x16 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int i = 0; i < 16; ++i) acc[t][i] = 0.f;

    const int c0 = slice * chunks_per_slice;
    const int c1 = min(c0 + chunks_per_slice, n_chunks);
    // Per-thread source offsets of the chunk's P slots (fixed), and the
    // wave's LDS destinations (uniform).
    int64_t src[LOADS];
#pragma unroll
    for (int u = 0; u < LOADS; ++u) {
        const int idx = 64 * wave + lane + 64 * WV * u;
        src[u] = (idx / V) * part_stride + (idx % V);
    }
    auto fetch = [&](int chunk, int buf) {
        const typename Half<H>::v8* base = parts + (int64_t)chunk * V;
#pragma unroll
        for (int u = 0; u < LOADS; ++u) {
            const int widx = 64 * wave + 64 * WV * u;   // the wave's first slot
            if ((V * SPLIT) % (64 * WV) == 0 || widx < V * SPLIT)
                __builtin_amdgcn_global_load_lds(
                    (const void*)(base + src[u]),
                    (__attribute__((address_space(3))) void*)(&lp[buf][0][0] + widx), 16, 0, 0);
        }
    };
    // Grid indices of this lane's first cell of the next chunk to generate,
    // advanced by one chunk (16 cells) per iteration instead of a divide.
    int gi = 0, gj = 0, gk = 0;
    auto locate = [&](int cell) {
        int rw;
        divmod_pos(cell, nx, inv_nx, rw, gi);
        divmod_pos(rw, ny, inv_ny, gk, gj);
    };
    auto advance = [&]() {   // nx >= 8: at most two row wraps per 16 cells
        gi += kBatchChunk;
#pragma unroll
        for (int w = 0; w < 2; ++w) {
            const bool wrap = gi >= nx;
            gi = wrap ? gi - nx : gi;
            gj += wrap ? 1 : 0;
            const bool wk = gj == ny;
            gj = wk ? 0 : gj;
            gk += wk ? 1 : 0;
        }
    };
    v8 a[SPLIT];
    if (c0 < c1) {
        fetch(c0, 0);
        if constexpr (PF == 2) fetch(min(c0 + 1, c1 - 1), 1);
        batched_gen<SPLIT, MODE, H>(p, c0 * kBatchChunk + 8 * kh, nx, ny, inv_nx, inv_ny, sy, sz,
                                    gs, a);
        if constexpr (MODE != 0) locate(min(c0 + 1, c1 - 1) * kBatchChunk + 8 * kh);
    }
    if constexpr (PF == 2) {
        __builtin_amdgcn_s_waitcnt(kWaitVm0 | LOADS);   // chunk c0's loads have landed
        ring_barrier();
    } else {
        __builtin_amdgcn_s_waitcnt(kWaitVm0);
        __syncthreads();
    }
    int buf = 0;   // LDS buffer of chunk c: (c - c0) % NBUF
    for (int c = c0; c < c1; ++c) {
        const int cn = min(c + 1, c1 - 1);
        // buffer of chunk c + PF, last read in iteration c - 1 (before the barrier)
        const int fb = buf + PF >= NBUF ? buf + PF - NBUF : buf + PF;
        fetch(min(c + PF, c1 - 1), fb);
        v8 cur[SPLIT];
#pragma unroll
        for (int q = 0; q < SPLIT; ++q) cur[q] = a[q];
        v8 b[SPLIT][NT];
#pragma unroll
        for (int q = 0; q < SPLIT; ++q)
#pragma unroll
            for (int t = 0; t < NT; ++t) b[q][t] = lp[buf][q][kh * Bp + 32 * t + row];
        BatchedGen<SPLIT, MODE, H> g;
        g.gs = gs;
        if constexpr (MODE == 0) {
            g.begin(p, cn * kBatchChunk + 8 * kh, nx, ny, inv_nx, inv_ny, sy, sz);
        } else {
            g.begin_at(p, gi, gj, gk, nx, ny, sy, sz);
            if (cn + 1 < c1) advance();   // the chunk after cn (the last chunk repeats)
        }
        constexpr int TOTAL = 3 * NT;
        static_for<0, 3>([&](auto pc) {
            constexpr int pr = decltype(pc)::value;
            constexpr int qa = pr == 2 ? 1 : 0, qb = pr == 1 ? 1 : 0;
            static_for<0, NT>([&](auto tc) {
                constexpr int t = decltype(tc)::value;
                acc[t] = Half<H>::mfma(cur[qa], b[qb][t], acc[t]);
                gen_after_mfma<TOTAL>(pr * NT + t,
                                      [&](auto ec) { g.template elem<decltype(ec)::value>(a); });
                __builtin_amdgcn_sched_barrier(0);   // keep this order
            });
        });
        if constexpr (PF == 2) {
            __builtin_amdgcn_s_waitcnt(kWaitVm0 | LOADS);   // all but chunk c + 2's loads
            ring_barrier();
        } else {
            __builtin_amdgcn_s_waitcnt(kWaitVm0);
            __syncthreads();
        }
        buf = buf + 1 == NBUF ? 0 : buf + 1;
    }
    // no load into this workgroup's LDS may still be in flight when it ends
    if constexpr (PF == 2) __builtin_amdgcn_s_waitcnt(kWaitVm0);
    float* out = part + ((size_t)slice * m_pad + mbase) * Bp;
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const int srow = 8 * (e >> 2) + 4 * kh + (e & 3);
            out[(size_t)srow * Bp + 32 * t + row] = acc[t][e];
        }
}

// --------------------------------------------------------------- adjoint
// X = G^T R for R = [B residual vectors x M stations]: the same MFMA scheme
// with the roles of stations and cells exchanged. A wave owns 32 cells (the
// MFMA's M; the cell's grid coordinates stay in registers) and the K loop
// walks the slice's stations 16 at a time. The B operand is the packed
// residual chunk (the forward's packing over stations instead of cells), and
// the lane's 8 G values of a chunk are those of its cell with the stations
// 8 (lane / 32) + 0..7 of the chunk, evaluated with exactly the forward's
// arithmetic (MODE 0), so both directions use bit-identical G entries.
//
// The stations of a chunk never touch LDS: lane l holds the record of
// station 16 c + 8 (l / 32) + l % 8 of the chunk in registers, and a lane's
// 8 stations 8 (l / 32) + e are broadcast to it by DPP row_newbcast:e (lane
// e of each 16-lane row to the whole row), a register move that needs no
// memory wait. G of chunk c + 1 is generated while chunk c's MFMAs run, so
// the record of chunk c + 2 is loaded (plain global load) during iteration c
// and used in iteration c + 1. The only LDS traffic is the residual parts,
// which load straight from global memory into the buffer that is not being
// read, as in the forward.
//   ROW 1 (nx % 32 == 0: a wave's 32 cells lie in one grid row and share y
//          and z): each lane first turns its record into (x, dy^2 + dz^2 +
//          1e-20, dz) for the wave's row, so a G entry costs 1 sub + 1 FMA +
//          rsq + 3 mul instead of 3 sub + 3 FMA + rsq + 3 mul, with the same
//          roundings.
//
// With one K-slice (X != nullptr) the kernel stores X[b][n] itself: lane
// (row, kh) holds vector 32 t + row at cells 8 g + 4 kh + 0..3 of its wave's
// 32, i.e. 4 consecutive cells per (t, g) -- one 16-byte store when vec4
// (N % 4 == 0 and X 16-byte aligned). The scale is then folded into the
// packed residuals, so the accumulators are stored as they are (no register
// work that would cost occupancy).
template <int E>
__device__ __forceinline__ float row_bcast(float v) {   // lane E of this lane's 16-lane row
    return __builtin_bit_cast(
        float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x150 + E, 0xF, 0xF, false));
}

// VAR 0: per tile pair, the pair's B operands read while the previous pair's
// MFMAs run, scheduling-group hints for the G elements. VAR 1 (two-part
// splits; batched_adj2_kernel, up to 256 registers): as batched_fwd2_kernel,
// every B operand of the chunk read at the top, the MFMAs term by term over
// the tiles, the G elements fixed between them. Same products in the same
// order per accumulator: equal results.
template <int NT, int SPLIT, int ROW, typename H, int VAR, int WV = 4>   // WV waves per workgroup
__device__ __forceinline__ void batched_adj_body(
    const float4* __restrict__ st, const typename Half<H>::v8* __restrict__ parts,
    int64_t part_stride, int nx, int ny, double inv_nx, double inv_ny, float sy, float sz,
    float gs, int n_chunks, int chunks_per_slice, int n_pad, float* __restrict__ part,
    float* __restrict__ X, int n_cells, int B, int vec4, const unsigned* __restrict__ amax,
    double sdiv) {
    using v8 = typename Half<H>::v8;
    constexpr int Bp = 32 * NT;
    constexpr int V = Bp * 2;                          // 16-byte slots per part per chunk
    constexpr int LOADS = (V * SPLIT + 64 * WV - 1) / (64 * WV);   // per thread per chunk
    __shared__ v8 lp[2][SPLIT][V];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int row = lane & 31, kh = lane >> 5;
    const int nbase = blockIdx.x * (32 * WV) + wave * 32;
    float cx, cy, cz;
    {
        int rw, i, j, k;
        divmod_pos(nbase + row, nx, inv_nx, rw, i);
        divmod_pos(rw, ny, inv_ny, k, j);
        cx = (float)i;
        cy = (float)j * sy;
        cz = (float)k * sz;
        // rounded here: without the pin the compiler may fuse j * sy into a
        // later subtraction (an FMA) in one variant's code and not in
        // another's, and the variants' G entries would differ in the last bit
        asm volatile("" : "+v"(cx), "+v"(cy), "+v"(cz));
    }

    f32x16 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int i = 0; i < 16; ++i) acc[t][i] = 0.f;

    const int c0 = blockIdx.y * chunks_per_slice;
    const int c1 = min(c0 + chunks_per_slice, n_chunks);
    auto fetch = [&](int chunk, int buf) {
#pragma unroll
        for (int u = 0; u < LOADS; ++u) {
            const int idx = threadIdx.x + 64 * WV * u;
            if ((V * SPLIT) % (64 * WV) == 0 || idx < V * SPLIT)
                __builtin_amdgcn_global_load_lds(
                    (const void*)(parts + (idx / V) * part_stride + (int64_t)chunk * V + (idx % V)),
                    (__attribute__((address_space(3))) void*)(&lp[buf][0][0] + (idx - lane)),
                    16, 0, 0);
        }
    };
    // x, y, z only: a load whose unused fourth register is reallocated while
    // the load is in flight would need a wait for it
    auto record = [&](int chunk) {
        const float* r = &st[(int64_t)chunk * kBatchChunk + 8 * kh + (lane & 7)].x;
        return make_float3(r[0], r[1], r[2]);
    };
    auto prep = [&](float3 p) {
        if constexpr (ROW) {   // (x, R0, dz) of this lane's station for the wave's row
            const float dy = p.y - cy;
            const float dz = p.z - cz;
            p.y = fmaf(dy, dy, fmaf(dz, dz, 1e-20f));
            p.z = Half<H>::scaled ? dz * gs : dz;   // G scale folded into dz
        }
        return p;
    };
    auto elem = [&](auto ec, const float3& p, v8 (&a)[SPLIT]) {
        constexpr int e = decltype(ec)::value;
        float r2, dz;
        const float x = row_bcast<e>(p.x) - cx;
        if constexpr (ROW) {
            r2 = fmaf(x, x, row_bcast<e>(p.y));
            dz = row_bcast<e>(p.z);
        } else {
            const float dy = row_bcast<e>(p.y) - cy;
            dz = row_bcast<e>(p.z) - cz;
            r2 = fmaf(x, x, fmaf(dy, dy, fmaf(dz, dz, 1e-20f)));
        }
        const float v = rsq_f(r2);
        store_split<SPLIT, H, ROW != 0>(dz * (v * v * v), gs, a, e);
    };
    auto gen = [&](float3 p, v8 (&a)[SPLIT]) {
        p = prep(p);
        static_for<0, 8>([&](auto ec) { elem(ec, p, a); });
    };
    v8 a[SPLIT];
    float3 rec = make_float3(0.f, 0.f, 0.f);
    if (c0 < c1) {
        fetch(c0, 0);
        gen(record(c0), a);
        rec = record(min(c0 + 1, c1 - 1));
    }
    __builtin_amdgcn_s_waitcnt(kWaitVm0);
    __syncthreads();
    for (int c = c0; c < c1; ++c) {
        const int buf = (c - c0) & 1;
        fetch(min(c + 1, c1 - 1), buf ^ 1);
        const float3 now = rec;                 // chunk c + 1
        rec = record(min(c + 2, c1 - 1));
        v8 cur[SPLIT];
#pragma unroll
        for (int q = 0; q < SPLIT; ++q) cur[q] = a[q];
        if constexpr (VAR == 1) {
            static_assert(SPLIT == 2, "variant 1 is for the two-part splits");
            v8 b[SPLIT][NT];
#pragma unroll
            for (int q = 0; q < SPLIT; ++q)
#pragma unroll
                for (int t = 0; t < NT; ++t) b[q][t] = lp[buf][q][kh * Bp + 32 * t + row];
            const float3 pn = prep(now);
            constexpr int TOTAL = 3 * NT;
            static_for<0, 3>([&](auto pc) {
                constexpr int pr = decltype(pc)::value;
                constexpr int qa = pr == 2 ? 1 : 0, qb = pr == 1 ? 1 : 0;
                static_for<0, NT>([&](auto tc) {
                    constexpr int t = decltype(tc)::value;
                    acc[t] = Half<H>::mfma(cur[qa], b[qb][t], acc[t]);
                    gen_after_mfma<TOTAL>(pr * NT + t, [&](auto ec) { elem(ec, pn, a); });
                    __builtin_amdgcn_sched_barrier(0);   // keep this order
                });
            });
            __builtin_amdgcn_s_waitcnt(kWaitVm0);
            __syncthreads();
            continue;
        }
        gen(now, a);
        constexpr int TP = NT >= 2 ? 2 : 1;
        v8 b[2][TP][SPLIT];
        auto load_b = [&](int t, v8 (&bb)[TP][SPLIT]) {
#pragma unroll
            for (int u = 0; u < TP; ++u)
#pragma unroll
                for (int q = 0; q < SPLIT; ++q) bb[u][q] = lp[buf][q][kh * Bp + 32 * (t + u) + row];
        };
        load_b(0, b[0]);
#pragma unroll
        for (int t = 0; t < NT; t += TP) {
            const int cb = (t / TP) & 1;
            if (t + TP < NT) load_b(t + TP, b[cb ^ 1]);
#pragma unroll
            for (int s2 = 0; s2 < SPLIT; ++s2)
#pragma unroll
                for (int qa = 0; qa <= s2; ++qa)
#pragma unroll
                    for (int u = 0; u < TP; ++u)
                        acc[t + u] = Half<H>::mfma(cur[qa], b[cb][u][s2 - qa], acc[t + u]);
        }
        // Issue-order hints: global loads, the first tile pair's LDS reads,
        // then per tile pair the next pair's reads and its MFMAs, each
        // followed by a few VALU instructions of the next chunk's G block.
        // (The forward's source-order interleave measured the same on this
        // kernel's shapes, and here it moved the copy of the prefetched
        // record into the iteration, with a wait for every load.)
        constexpr int kPerPair = TP * SPLIT * (SPLIT + 1) / 2;
        __builtin_amdgcn_sched_group_barrier(0x020, LOADS + 1, 0);   // VMEM read
        __builtin_amdgcn_sched_group_barrier(0x100, TP * SPLIT, 0);
#pragma unroll
        for (int t = 0; t < NT; t += TP) {
            if (t + TP < NT) __builtin_amdgcn_sched_group_barrier(0x100, TP * SPLIT, 0);
#pragma unroll
            for (int i = 0; i < kPerPair; ++i) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x002, SPLIT == 3 ? 4 : 6, 0);
            }
        }
        __builtin_amdgcn_s_waitcnt(kWaitVm0);
        __syncthreads();
    }
    // D layout: lane holds model 32t + lane%32, cells 8*(i/4) + 4*(lane/32) + i%4
    if (X) {
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int b = 32 * t + row;
            if (b >= B) continue;
            // BF16: the scale was folded into the packed residuals. FP16 (here
            // only in variant 1, whose register budget leaves room): the
            // vector's factor sdiv / half_scale(amax[b]), formed in FP64.
            float f = 1.f;
            if constexpr (Half<H>::scaled) f = (float)(sdiv / (double)half_scale(amax[b]));
            float* xb = X + (size_t)b * n_cells;
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int n = nbase + 8 * g + 4 * kh;
                float v[4];
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    v[i] = Half<H>::scaled ? acc[t][4 * g + i] * f : acc[t][4 * g + i];
                if (vec4 && n + 3 < n_cells) {
                    *reinterpret_cast<float4*>(xb + n) = make_float4(v[0], v[1], v[2], v[3]);
                } else {
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        if (n + i < n_cells) xb[n + i] = v[i];
                }
            }
        }
        return;
    }
    float* out = part + ((size_t)blockIdx.y * n_pad + nbase) * Bp;
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const int crow = 8 * (e >> 2) + 4 * kh + (e & 3);
            out[(size_t)crow * Bp + 32 * t + row] = acc[t][e];
        }
}

template <int NT, int SPLIT, int ROW, typename H>   // model tiles of 32 (B <= 32 * NT); 2- or 3-way split; part type
__global__ __launch_bounds__(256) void batched_adj_kernel(
    const float4* __restrict__ st, const typename Half<H>::v8* __restrict__ parts,
    int64_t part_stride, int nx, int ny, double inv_nx, double inv_ny, float sy, float sz,
    float gs, int n_chunks, int chunks_per_slice, int n_pad, float* __restrict__ part,
    float* __restrict__ X, int n_cells, int B, int vec4) {
    batched_adj_body<NT, SPLIT, ROW, H, 0>(st, parts, part_stride, nx, ny, inv_nx, inv_ny, sy, sz,
                                           gs, n_chunks, chunks_per_slice, n_pad, part, X,
                                           n_cells, B, vec4, nullptr, 1.0);
}

template <int NT, int ROW, typename H, int WV>
__global__ __launch_bounds__(64 * WV) __attribute__((amdgpu_waves_per_eu(2, 2)))
void batched_adj2_kernel(const float4* __restrict__ st, const typename Half<H>::v8* __restrict__ parts,
                         int64_t part_stride, int nx, int ny, double inv_nx, double inv_ny,
                         float sy, float sz, float gs, int n_chunks, int chunks_per_slice,
                         int n_pad, float* __restrict__ part, float* __restrict__ X, int n_cells,
                         int B, int vec4, const unsigned* __restrict__ amax, double sdiv) {
    batched_adj_body<NT, 2, ROW, H, 1, WV>(st, parts, part_stride, nx, ny, inv_nx, inv_ny, sy, sz, gs,
                                       n_chunks, chunks_per_slice, n_pad, part, X, n_cells, B,
                                       vec4, amax, sdiv);
}

// Y[b][m] = scale * sum_ks part[ks][m][b]. A block sums a 16 (stations) x 64
// (models) tile: thread (station t / 16, models 4 (t % 16) .. + 3) reads one
// 16-byte run per slice, 256 contiguous bytes per station, and sums the four
// entries in FP64, slice by slice in order; the sums leave transposed through
// LDS, 4 consecutive stations of one model per thread. The adjoint uses it
// with cells in place of stations and a float output.
// With FP16 parts (amax given) the sums carry model b's operand scale
// half_scale(amax[b]) and, in the forward, station m's G scale gsc[m] (gsc
// given), which the FP64 scale takes out.
// grid = (ceil(m / 16), ceil(B / 64)); Bp a multiple of 32, part 16-byte
// aligned.
template <typename TO>
__global__ __launch_bounds__(256) void batched_reduce_kernel(const float* __restrict__ part,
                                                             int ks, int m, int m_pad, int B,
                                                             int Bp, double scale,
                                                             const unsigned* __restrict__ amax,
                                                             const float* __restrict__ gsc,
                                                             TO* __restrict__ Y) {
    __shared__ double tile[16][65];
    const int mb = blockIdx.x * 16, bb = blockIdx.y * 64;
    const int t = threadIdx.x;
    {
        const int i = t >> 4, j = 4 * (t & 15);
        const int mm = mb + i, b0 = bb + j;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        if (mm < m && b0 < Bp) {   // b0 .. b0 + 3 < Bp: Bp is a multiple of 4
            const float4* p = reinterpret_cast<const float4*>(part + (size_t)mm * Bp + b0);
            const size_t step = (size_t)m_pad * Bp / 4;
#pragma unroll 8
            for (int q = 0; q < ks; ++q) {
                const float4 v = p[(size_t)q * step];
                s0 += (double)v.x;
                s1 += (double)v.y;
                s2 += (double)v.z;
                s3 += (double)v.w;
            }
        }
        tile[i][j] = s0;
        tile[i][j + 1] = s1;
        tile[i][j + 2] = s2;
        tile[i][j + 3] = s3;
    }
    __syncthreads();
    // 64 models x 16 stations out: thread (model t / 4, stations 4 (t % 4) + 0..3)
    const int bl = t >> 2, b = bb + bl;
    if (b >= B) return;
    const double f = amax ? scale / (double)half_scale(amax[b]) : scale;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int i = 4 * (t & 3) + u, mm = mb + i;
        if (mm < m) {
            double g = f;
            if (gsc) g /= (double)gsc[mm];
            Y[(size_t)b * m + mm] = (TO)(g * tile[i][bl]);
        }
    }
}

// One K-slice: X[b][c] = f_b * part[c][b], f_b = scale, or with FP16 parts
// scale / half_scale(amax[b]): a transpose through LDS in tiles of 64
// cells x 32 models, so that each model's 64 cells leave as one 256-byte run
// (the reduce's 8-entry runs halve the write rate at this size).
// grid = (ceil(n / 64), ceil(B / 32)); block = 256.
__global__ __launch_bounds__(256) void batched_transpose_kernel(const float* __restrict__ part,
                                                                int n, int B, int Bp, double scale,
                                                                const unsigned* __restrict__ amax,
                                                                float* __restrict__ X) {
    __shared__ float tile[32][65];
    __shared__ double f[32];
    const int nb = blockIdx.x * 64, bb = blockIdx.y * 32;
    const int t = threadIdx.x;
    if (t < 32) {
        const int b = bb + t;
        f[t] = amax && b < B ? scale / (double)half_scale(amax[b]) : scale;
    }
    {
        const int tx = t & 31, ty = t >> 5;   // model, cell row
        const int b = bb + tx;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int c = nb + ty + 8 * i;
            tile[tx][ty + 8 * i] = c < n && b < B ? part[(size_t)c * Bp + b] : 0.f;
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int idx = i * 256 + t;
        const int bl = idx >> 6, c = idx & 63;
        const int b = bb + bl;
        if (b < B && nb + c < n)
            X[(size_t)b * n + nb + c] = (float)(f[bl] * (double)tile[bl][c]);
    }
}

// --------------------------------------------------------------- launchers
namespace {
// model tiles of 32, padded to a power of two (1, 2, 4 or 8 for B <= 256)
int batched_nt(int B) {
    const int t = (B + 31) / 32;
    int p = 1;
    while (p < t) p <<= 1;
    return p;
}
}  // namespace

namespace {
// Split codes: 2 and 3 = BF16 parts, 4 = two FP16 parts.
int split_parts(int split) { return split == 4 ? 2 : split; }
constexpr size_t kAmaxBytes = 256 * sizeof(unsigned);   // per-model absmax (FP16)

static_assert(kHalfTop == 14, "half_g_scale brings G to 2^14");

int check_split(int split, bool bound) {
    if (split != 2 && split != 3 && split != 4)
        return fail(ZTM_EINVAL, "split must be 2, 3 or 4 (got %d)", split);
    if (split == 4 && !bound)
        return fail(ZTM_EINVAL, "the FP16 split needs the plan's bound on |G| (none for this plan)");
    return ZTM_OK;
}
}  // namespace

size_t batched_workspace(const PmShape& s, int B, int ks, int split) {
    const int Bp = 32 * batched_nt(B);
    const int64_t n16 = (s.n_cells + kBatchChunk - 1) / kBatchChunk * kBatchChunk;
    const size_t packed = (size_t)n16 * Bp * split_parts(split) * sizeof(__bf16);
    const size_t partial = (size_t)ks * s.m_pad * Bp * sizeof(float);
    return packed + partial + kAmaxBytes;
}

namespace {
// K-slices for a grid of `tiles` row tiles (workgroups per slice) over
// n_chunks K chunks.
// Resident workgroups on 256 CUs for the kernel instance B selects
// (registers: 2 per CU at 256 models, more for fewer model tiles).
int64_t resident_groups(int B) {
    const int nt = batched_nt(std::max(B, 1));
    return 256 * (nt >= 8 ? 2 : nt == 4 ? 3 : nt == 2 ? 5 : 8);
}

int default_slices(int64_t tiles, int64_t n_chunks, int B) {
    const int64_t cap = resident_groups(B);
    // At least one full round of workgroups and at most ~384 chunks (MFMA
    // accumulation steps) per slice, to bound the FP32 accumulation error;
    // then as many slices as fit the same number of rounds, so the last round
    // is not left mostly empty (time ~ rounds x chunks per slice).
    int64_t ks = std::max<int64_t>((cap + tiles - 1) / tiles, (n_chunks + 383) / 384);
    const int64_t rounds = (tiles * ks + cap - 1) / cap;
    ks = std::max<int64_t>(ks, rounds * cap / tiles);
    ks = std::min<int64_t>(ks, std::max<int64_t>(1, n_chunks));
    return (int)std::min<int64_t>(ks, 65535);
}
}  // namespace

int batched_default_slices(const PmShape& s, int B) {
    return default_slices((s.m + kBatchTile - 1) / kBatchTile,
                          (s.n_cells + kBatchChunk - 1) / kBatchChunk, B);
}

int batched_adj_default_slices(const PmShape& s, int B) {
    const int64_t tiles = (s.n_cells + kBatchTile - 1) / kBatchTile;
    const int64_t n_chunks = ((int64_t)s.m + kBatchChunk - 1) / kBatchChunk;
    const int ks = default_slices(tiles, n_chunks, B);
    // One slice stores G^T R from the accumulators (no partials, no reduce
    // pass). Take it when it costs no more rounds x chunks per slice than the
    // split and its FP32 accumulation run stays within 768 chunks (12,288
    // stations).
    if (ks > 1 && n_chunks <= 768) {
        const int64_t cap = resident_groups(B);
        const int64_t one = (tiles + cap - 1) / cap * n_chunks;
        const int64_t split = (tiles * ks + cap - 1) / cap * ((n_chunks + ks - 1) / ks);
        if (one <= split) return 1;
    }
    return ks;
}

int batched_fwd_launch(const PmShape& s, int B, int ks, int split, const float* st,
                       const float* P, double thr, double sy, double sz, double scale,
                       const float* gsc, void* ws, double* Y, hipStream_t stream,
                       int variant, int waves) {
    if (waves != 4 && waves != 8) return fail(ZTM_EINVAL, "waves per workgroup must be 4 or 8");
    if (variant < 0 || variant > 2) return fail(ZTM_EINVAL, "kernel variant must be 0, 1 or 2");
    if (waves == 8 && (variant == 0 || split == 3))
        return fail(ZTM_EINVAL, "8 waves per workgroup need variant 1 and a two-part split");
    if (B <= 0 || s.m == 0) return ZTM_OK;
    if (int rc = check_split(split, gsc != nullptr)) return rc;
    const int nt = batched_nt(B);
    if (nt > 8) return fail(ZTM_EINVAL, "batched forward supports at most 256 models per call (got %d)", B);
    const int Bp = 32 * nt;
    if (s.n_cells > (int64_t)INT32_MAX - 2 * kBatchChunk)
        return fail(ZTM_EINVAL, "batched forward supports up to 2^31 - 32 cells");
    const int64_t n_chunks = (s.n_cells + kBatchChunk - 1) / kBatchChunk;
    ks = std::max(1, std::min<int>(ks, (int)std::min<int64_t>(n_chunks, 65535)));
    const int per = (int)((n_chunks + ks - 1) / ks);
    ks = (int)((n_chunks + per - 1) / per);
    const int64_t n_packed = n_chunks * Bp * kBatchChunk;
    const int np = split_parts(split);
    __bf16* parts = (__bf16*)ws;
    float* part = (float*)(parts + np * n_packed);
    unsigned* amax = nullptr;
    const dim3 pgrid((unsigned)((s.n_cells + kPackCells - 1) / kPackCells), (unsigned)(Bp / 32));
    const int pvec4 = s.n_cells % 4 == 0 && (uintptr_t)P % 16 == 0;
    if (split == 4) {
        amax = (unsigned*)(part + (size_t)ks * s.m_pad * Bp);
        hipLaunchKernelGGL(batched_clear_kernel, dim3(1), dim3(256), 0, stream, amax, Bp);
        const int64_t blocks = std::min<int64_t>(16, (s.n_cells + 4095) / 4096);
        hipLaunchKernelGGL(batched_absmax_kernel, dim3((unsigned)blocks, (unsigned)B), dim3(256),
                           0, stream, P, s.n_cells, (float)thr, amax);
        hipLaunchKernelGGL((batched_pack_kernel<2, _Float16>), pgrid, dim3(256), 0, stream, P, B,
                           s.n_cells, Bp, n_chunks, n_packed, (float)thr, 1.f,
                           (const unsigned*)amax, pvec4, (_Float16*)parts);
    } else if (split == 3) {
        hipLaunchKernelGGL((batched_pack_kernel<3, __bf16>), pgrid, dim3(256), 0, stream, P, B,
                           s.n_cells, Bp, n_chunks, n_packed, (float)thr, 1.f, nullptr, pvec4, parts);
    } else {
        hipLaunchKernelGGL((batched_pack_kernel<2, __bf16>), pgrid, dim3(256), 0, stream, P, B,
                           s.n_cells, Bp, n_chunks, n_packed, (float)thr, 1.f, nullptr, pvec4, parts);
    }
    ZTM_HIP(hipGetLastError());
    const int tile = 32 * waves;   // stations per workgroup (m_pad is a multiple of 512)
    const dim3 grid((unsigned)((s.m + tile - 1) / tile), (unsigned)ks);
    const int64_t stride = n_packed / 8;   // in 16-byte units
    {
        ProfScope prof("batched_fwd", stream);
#define ZTM_BATCHED_FWD2(NT, MODE, H, WV, PF)                                                   \
    hipLaunchKernelGGL((batched_fwd2_kernel<NT, MODE, H, WV, PF>), grid, dim3(64 * WV), 0,      \
                       stream, (const float4*)st, (const typename Half<H>::v8*)parts, stride,   \
                       s.nx, s.ny, 1.0 / s.nx, 1.0 / s.ny, (float)sy, (float)sz, gsc,           \
                       (int)n_chunks, per, s.m_pad, part)
#define ZTM_BATCHED_LAUNCH(NT, SP, MODE, H)                                                     \
    if (SP == 2 && variant == 2 && waves == 8)                                                 \
        ZTM_BATCHED_FWD2(NT, MODE, H, 8, 2);                                                   \
    else if (SP == 2 && variant == 2)                                                          \
        ZTM_BATCHED_FWD2(NT, MODE, H, 4, 2);                                                   \
    else if (SP == 2 && variant == 1 && waves == 8)                                            \
        ZTM_BATCHED_FWD2(NT, MODE, H, 8, 1);                                                   \
    else if (SP == 2 && variant == 1)                                                          \
        ZTM_BATCHED_FWD2(NT, MODE, H, 4, 1);                                                   \
    else                                                                                       \
        hipLaunchKernelGGL((batched_fwd_kernel<NT, SP, MODE, H>), grid, dim3(256), 0, stream,  \
                           (const float4*)st, (const typename Half<H>::v8*)parts, stride, s.nx, \
                           s.ny, 1.0 / s.nx, 1.0 / s.ny, (float)sy, (float)sz, gsc,             \
                           (int)n_chunks, per, s.m_pad, part)
#define ZTM_BATCHED_CASE(NT, SP, H)                                                            \
    if (s.nx % 8 == 0)                                                                         \
        ZTM_BATCHED_LAUNCH(NT, SP, 2, H);                                                      \
    else if (s.nx >= 8)                                                                        \
        ZTM_BATCHED_LAUNCH(NT, SP, 1, H);                                                      \
    else                                                                                       \
        ZTM_BATCHED_LAUNCH(NT, SP, 0, H)
#define ZTM_BATCHED_NT(SP, H)                                                                  \
    switch (nt) {                                                                              \
        case 1: ZTM_BATCHED_CASE(1, SP, H); break;                                             \
        case 2: ZTM_BATCHED_CASE(2, SP, H); break;                                             \
        case 4: ZTM_BATCHED_CASE(4, SP, H); break;                                             \
        default: ZTM_BATCHED_CASE(8, SP, H); break;                                            \
    }
        if (split == 4) {
            ZTM_BATCHED_NT(2, _Float16)
        } else if (split == 3) {
            ZTM_BATCHED_NT(3, __bf16)
        } else {
            ZTM_BATCHED_NT(2, __bf16)
        }
#undef ZTM_BATCHED_NT
#undef ZTM_BATCHED_CASE
#undef ZTM_BATCHED_LAUNCH
#undef ZTM_BATCHED_FWD2
        ZTM_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(batched_reduce_kernel<double>, dim3((s.m + 15) / 16, (B + 63) / 64), dim3(256), 0,
                       stream, part, ks, s.m, s.m_pad, B, Bp, scale, (const unsigned*)amax,
                       split == 4 ? gsc : nullptr, Y);
    ZTM_HIP(hipGetLastError());
    return ZTM_OK;
}


namespace {
__global__ void batched_zero_kernel(float* __restrict__ x, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        x[i] = 0.f;
}
}  // namespace

size_t batched_adj_workspace(const PmShape& s, int B, int ks, int split) {
    const int Bp = 32 * batched_nt(B);
    const int64_t m16 = ((int64_t)s.m + kBatchChunk - 1) / kBatchChunk * kBatchChunk;
    const int64_t n_pad = (s.n_cells + 255) / 256 * 256;   // both workgroup sizes
    const size_t packed = (size_t)m16 * Bp * split_parts(split) * sizeof(__bf16);
    // one BF16 slice stores X directly; FP16 always goes through the reduce,
    // which takes out the per-vector scales
    const bool direct = ks == 1 && split != 4;
    const size_t partial = direct ? 0 : (size_t)ks * n_pad * Bp * sizeof(float);
    return packed + partial + kAmaxBytes;
}

int batched_adj_launch(const PmShape& s, int B, int ks, int split, const float* st,
                       const float* R, double sy, double sz, double scale, double gmax, void* ws,
                       float* X, hipStream_t stream, int variant, int waves) {
    if (waves != 4 && waves != 8) return fail(ZTM_EINVAL, "waves per workgroup must be 4 or 8");
    if (variant < 0 || variant > 2) return fail(ZTM_EINVAL, "kernel variant must be 0, 1 or 2");
    if (waves == 8 && (variant == 0 || split == 3))
        return fail(ZTM_EINVAL, "8 waves per workgroup need variant 1 and a two-part split");
    if (B <= 0 || s.n_cells == 0) return ZTM_OK;
    if (split != 4 || s.m > 0)
        if (int rc = check_split(split, gmax >= 0.0 && std::isfinite(gmax))) return rc;
    const int nt = batched_nt(B);
    if (nt > 8) return fail(ZTM_EINVAL, "batched adjoint supports at most 256 vectors per call (got %d)", B);
    if (s.n_cells > (int64_t)INT32_MAX - 2 * kBatchTile)
        return fail(ZTM_EINVAL, "batched adjoint supports up to 2^31 - 256 cells");
    if (s.m == 0) {   // no stations: G^T R = 0
        const int64_t n = (int64_t)B * s.n_cells;
        const int64_t blocks = std::min<int64_t>((n + 255) / 256, 1 << 20);
        hipLaunchKernelGGL(batched_zero_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, X,
                           n);
        ZTM_HIP(hipGetLastError());
        return ZTM_OK;
    }
    const int Bp = 32 * nt;
    // station chunks read records up to 16 n_chunks - 1 < m_pad (a multiple of 16)
    const int64_t n_chunks = ((int64_t)s.m + kBatchChunk - 1) / kBatchChunk;
    ks = std::max(1, std::min<int>(ks, (int)std::min<int64_t>(n_chunks, 65535)));
    const int per = (int)((n_chunks + ks - 1) / ks);
    ks = (int)((n_chunks + per - 1) / per);
    const int64_t n_packed = n_chunks * Bp * kBatchChunk;
    const int tile = 32 * waves;   // cells per workgroup
    const int64_t tiles = (s.n_cells + tile - 1) / tile;
    const int n_pad = (int)(tiles * tile);
    const int np = split_parts(split);
    __bf16* parts = (__bf16*)ws;
    float* part = (float*)(parts + np * n_packed);
    // one slice stores X from the accumulators: BF16 with the scale folded
    // into the packed residuals; FP16 in variant 1 with each vector's factor
    // applied in the epilogue (variant 0 goes through the transpose)
    float* direct = ks == 1 && (split != 4 || variant >= 1) ? X : nullptr;
    // The residuals are packed like the forward's models, stations in place
    // of cells; with one slice they carry the scale (kernel stores as is).
    const float mul = direct && split != 4 ? (float)scale : 1.f;
    unsigned* amax = nullptr;
    float gs = 1.f;
    const dim3 pgrid((unsigned)((s.m + kPackCells - 1) / kPackCells), (unsigned)(Bp / 32));
    const int pvec4 = s.m % 4 == 0 && (uintptr_t)R % 16 == 0;
    if (split == 4) {
        amax = (unsigned*)(part + (size_t)ks * n_pad * Bp);
        gs = half_g_scale(gmax);
        hipLaunchKernelGGL(batched_clear_kernel, dim3(1), dim3(256), 0, stream, amax, Bp);
        const int64_t blocks = std::min<int64_t>(16, ((int64_t)s.m + 4095) / 4096);
        hipLaunchKernelGGL(batched_absmax_kernel, dim3((unsigned)blocks, (unsigned)B), dim3(256),
                           0, stream, R, (int64_t)s.m, 0.f, amax);
        hipLaunchKernelGGL((batched_pack_kernel<2, _Float16>), pgrid, dim3(256), 0, stream, R, B,
                           (int64_t)s.m, Bp, n_chunks, n_packed, 0.f, 1.f,
                           (const unsigned*)amax, pvec4, (_Float16*)parts);
    } else if (split == 3) {
        hipLaunchKernelGGL((batched_pack_kernel<3, __bf16>), pgrid, dim3(256), 0, stream, R, B,
                           (int64_t)s.m, Bp, n_chunks, n_packed, 0.f, mul, nullptr, pvec4, parts);
    } else {
        hipLaunchKernelGGL((batched_pack_kernel<2, __bf16>), pgrid, dim3(256), 0, stream, R, B,
                           (int64_t)s.m, Bp, n_chunks, n_packed, 0.f, mul, nullptr, pvec4, parts);
    }
    ZTM_HIP(hipGetLastError());
    const dim3 grid((unsigned)tiles, (unsigned)ks);
    const int64_t stride = n_packed / 8;   // in 16-byte units
    const int vec4 = s.n_cells % 4 == 0 && (uintptr_t)X % 16 == 0;
    {
        ProfScope prof("batched_adj", stream);
#define ZTM_BATCHED_ADJ2(NT, ROW, H, WV)                                                         \
    hipLaunchKernelGGL((batched_adj2_kernel<NT, ROW, H, WV>), grid, dim3(64 * WV), 0, stream,    \
                       (const float4*)st, (const typename Half<H>::v8*)parts, stride, s.nx,      \
                       s.ny, 1.0 / s.nx, 1.0 / s.ny, (float)sy, (float)sz, gs, (int)n_chunks,    \
                       per, n_pad, part, direct, (int)s.n_cells, B, vec4,                        \
                       (const unsigned*)amax, scale / gs)
#define ZTM_BATCHED_ADJ_ROW(NT, SP, ROW, H)                                                      \
    if (SP == 2 && variant >= 1 && waves == 8)                                                 \
        ZTM_BATCHED_ADJ2(NT, ROW, H, 8);                                                       \
    else if (SP == 2 && variant >= 1)                                                          \
        ZTM_BATCHED_ADJ2(NT, ROW, H, 4);                                                       \
    else                                                                                       \
        hipLaunchKernelGGL((batched_adj_kernel<NT, SP, ROW, H>), grid, dim3(256), 0, stream,     \
                           (const float4*)st, (const typename Half<H>::v8*)parts, stride, s.nx,  \
                           s.ny, 1.0 / s.nx, 1.0 / s.ny, (float)sy, (float)sz, gs,               \
                           (int)n_chunks, per, n_pad, part, direct, (int)s.n_cells, B, vec4)
#define ZTM_BATCHED_ADJ(NT, SP, H)                                                               \
    if (s.nx % 32 == 0)                                                                        \
        ZTM_BATCHED_ADJ_ROW(NT, SP, 1, H);                                                     \
    else                                                                                       \
        ZTM_BATCHED_ADJ_ROW(NT, SP, 0, H)
#define ZTM_BATCHED_ADJ_NT(SP, H)                                                              \
    switch (nt) {                                                                              \
        case 1: ZTM_BATCHED_ADJ(1, SP, H); break;                                              \
        case 2: ZTM_BATCHED_ADJ(2, SP, H); break;                                              \
        case 4: ZTM_BATCHED_ADJ(4, SP, H); break;                                              \
        default: ZTM_BATCHED_ADJ(8, SP, H); break;                                             \
    }
        if (split == 4) {
            ZTM_BATCHED_ADJ_NT(2, _Float16)
        } else if (split == 3) {
            ZTM_BATCHED_ADJ_NT(3, __bf16)
        } else {
            ZTM_BATCHED_ADJ_NT(2, __bf16)
        }
#undef ZTM_BATCHED_ADJ_NT
#undef ZTM_BATCHED_ADJ
#undef ZTM_BATCHED_ADJ_ROW
#undef ZTM_BATCHED_ADJ2
        ZTM_HIP(hipGetLastError());
    }
    if (direct) return ZTM_OK;
    if (ks == 1) {
        hipLaunchKernelGGL(batched_transpose_kernel,
                           dim3((unsigned)((s.n_cells + 63) / 64), (B + 31) / 32), dim3(256), 0,
                           stream, part, (int)s.n_cells, B, Bp, scale / gs,
                           (const unsigned*)amax, X);
        ZTM_HIP(hipGetLastError());
        return ZTM_OK;
    }
    hipLaunchKernelGGL(batched_reduce_kernel<float>,
                       dim3((unsigned)((s.n_cells + 15) / 16), (B + 63) / 64), dim3(256), 0, stream,
                       part, ks, (int)s.n_cells, n_pad, B, Bp, scale / gs,
                       (const unsigned*)amax, nullptr, X);
    ZTM_HIP(hipGetLastError());
    return ZTM_OK;
}

}  // namespace ztm
