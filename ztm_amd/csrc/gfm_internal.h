// Internal declarations shared by the HIP translation units of libztm_gfm.
#pragma once

#include <hip/hip_runtime.h>
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <string>

#include "../../include/ztm_gfm.h"

namespace ztm {

// Thread-local error detail; returns `code` so callers can `return fail(...)`.
int fail(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);

#define ZTM_HIP(call)                                             \
    do {                                                          \
        hipError_t e_ = (call);                                   \
        if (e_ != hipSuccess) return ::ztm::hip_fail(e_, #call);  \
    } while (0)

// Station records as the kernels read them: coordinates relative to the
// centre of cell (0,0,0) in units of L = |dx|, scaled so that cell (i,j,k)
// sits at (i, j*sy, k*sz) with sy = dy/L, sz = dz/L (signed). x is stored as
// (x - x0)/dx, so the sign of dx only flips the sign of a squared term.

// Small floor keeps r^2 > 0 in FP64 when a station coincides with a cell
// centre, so the (zero-length) vertical offset contributes 0 instead of
// 0*inf = NaN. In units of L^2 it is far below any real distance.
constexpr double kEpsD = 1e-20;
// FP32 kernels clamp a row's dy^2 + dz^2 (cell units) below at this value:
// half a cell width squared (see row_R32 in gfm_pair.h).
constexpr float  kRmin2F = 0.25f;

// Padding of prepared station arrays (largest forward tile = 64 lanes * 8).
constexpr int kStationPad = 512;

// Near field: pairs within kNear of the station's nearest cell (measured in
// units of the cell's largest edge, per axis) are re-evaluated in FP64 by the
// FP32 path. Prism mid zone: pairs within kMid are corrected from the
// far-field kernel to the prism integral (FP64 closed form in the near field,
// FP32 3x3x3 Gauss-Legendre beyond). Per-axis index radii: r_a =
// ceil(k * max_edge / edge_a), e.g. (2,2,2) and (7,7,7) for cubes.
constexpr int kNear = 2;
constexpr int kMid = 7;
// Least margin of the station-by-column index.
constexpr int kColMargin = 8;

// Quadrupole term of a non-cubic cell's far field, cell units (L = |dx|),
// divided by the cell volume: q2* = 2.5 * Q'_**, qz = Q'_zz with
// Q'_xx = (2 hx^2 - hy^2 - hz^2)/3 etc. for half-widths (hx, hy, hz).
struct QuadArgs {
    float q2x = 0.f, q2y = 0.f, q2z = 0.f, qz = 0.f;
};

// Correction applied by the reduce kernels of the FP32 path.
// kCorrPrism64: the FP64 path with the FP32 path's prism zones (precision
// ZTM_FP64_ZONED): mid and near zones corrected in FP64, no clamped rows.
enum CorrMode { kCorrNone = 0, kCorrPmNear = 1, kCorrPrism = 2, kCorrPrism64 = 3 };

// Everything the FP32 corrections need (mode kCorrNone disables them).
// st_c[m] = (ci, cj, ck, 0): the station's nearest cell index per axis.
// col_start/col_st: stations grouped by (ci, cj) column over the extended
// range [-cmx, nx+cmx) x [-cmy, ny+cmy) (adjoint gather); stations outside it
// are never near any cell. kmin/kmax: range of ck over them.
struct NearArgs {
    int mode = kCorrNone;
    int rx = kNear, ry = kNear, rz = kNear;      // correction box radii
    int nrx = kNear, nry = kNear, nrz = kNear;   // FP64 near-field radii
    int cmx = kColMargin, cmy = kColMargin;      // column-index margins
    int quad = 0;
    QuadArgs q;
    const float4* st_f = nullptr;
    const double4* st_d = nullptr;
    const int4* st_c = nullptr;
    const int* col_start = nullptr;
    const int* col_st = nullptr;
    const double* corr = nullptr;   // forward prism: per-station correction
    // Stations that may have clamped rows, grouped by (cj, ck) row key over
    // [-rmy, ny+rmy) x [-rmz, nz+rmz); jr/kr: the
    // row-index window around (cj, ck) a clamped row can lie in; rkmin/rkmax:
    // range of ck over them (adjoint gather).
    const int* row_start = nullptr;
    const int* row_st = nullptr;
    int jr = 1, kr = 1, rkmin = 1 << 29, rkmax = -(1 << 29);
    int rmy = kColMargin, rmz = kColMargin;   // row-key margins in j and k
    int nx = 0, ny = 0, nz = 0;
    float syf = 1.f, szf = 1.f, thr = 0.f;
    double syd = 1.0, szd = 1.0, vol = 1.0;   // vol = |sy*sz|, cell volume / L^3
    int kmin = 0, kmax = -1;
};

// Cell layers [*klo, *khi] of an nz-layer grid in which the adjoint gather
// may apply a near-field correction (empty, *klo > *khi, without one).
// Cells outside them take the plain split-K sum.
inline void near_layers(const NearArgs& na, int nz, int* klo, int* khi) {
    int lo = 1 << 30, hi = -(1 << 30);
    if (na.mode != kCorrNone) {
        if (na.kmin <= na.kmax) {
            lo = na.kmin - na.rz;
            hi = na.kmax + na.rz;
        }
        if (na.row_start && na.rkmin <= na.rkmax) {
            lo = lo < na.rkmin - na.kr ? lo : na.rkmin - na.kr;
            hi = hi > na.rkmax + na.kr ? hi : na.rkmax + na.kr;
        }
    }
    *klo = lo < 0 ? 0 : lo;
    *khi = hi > nz - 1 ? nz - 1 : hi;
}

// Launch shapes of the point-mass kernels.
// One batched kernel setting: issue-order variant (two-part splits: 0 or 1),
// waves per workgroup (variant 1: 4 or 8, i.e. 128 or 256 stations / cells),
// K-slices (0 = choose from the problem shape).
struct BatchedKernel {
    int variant = 1;
    int waves = 4;
    int slices = 0;
};
inline int batched_class(int B) { return B <= 32 ? 0 : B <= 64 ? 1 : B <= 128 ? 2 : 3; }

struct PmTuning {
    int st_per_lane = 2;      // forward: stations per lane (1,2,4,8)
    int rows_per_slice = 0;   // forward: 0 = choose from the problem shape
    int cells_per_lane = 8;   // adjoint: cells per lane (4,8,16)
    int adj_splits = 0;       // adjoint: 0 = choose from the problem shape
    // batched kernels, per direction (0 forward, 1 adjoint) and model-count
    // class (batched_class: up to 32, 64, 128, 256 models)
    BatchedKernel batched[2][4];
    int batched_split = 3;    // batched parts: 2 or 3 BF16, or 4 = FP16 hi/lo
                              // (plans made by ztm_plan_create start at 4)
};

struct PmShape {
    int nx, ny, nz;
    int64_t n_rows;           // ny*nz
    int64_t n_cells;          // nx*ny*nz
    int m, m_pad;             // stations, padded stations
};

// Kernel launchers (gfm_pointmass.hip). `part` must hold at least the number
// of doubles returned by the matching *_workspace() call.
size_t pm_fwd_workspace(const PmShape& s, const PmTuning& t, int precision, bool quad);
// Forward rows per LDS slice / adjoint station splits actually used for t.
// FP32 point-mass forward: rows per group sharing the per-column offset
// a = X - i (1 otherwise), and the tuning actually run (rows too wide for a
// row group fall back to S = 8).
int fwd_group(int S, size_t elem, bool quad);
PmTuning fwd_adjust(const PmShape& s, PmTuning t, size_t elem, bool quad);
int fwd_rows_per_slice(const PmShape& s, const PmTuning& t, size_t elem, bool quad = false);
// FP32 point-mass adjoint: rows a lane takes together (1 otherwise), the
// number of lane segments (C cells of G rows) and the station splits.
int adj_group(int C, size_t elem, bool quad);
size_t adj_lds(int C, size_t elem, bool quad);   // dynamic LDS of one adjoint workgroup
int64_t adj_segments(const PmShape& s, int C, int G, int* segs_per_row);
int adj_splits(const PmShape& s, const PmTuning& t, size_t elem, bool quad);
size_t pm_adj_workspace(const PmShape& s, const PmTuning& t, int precision,
                        bool near, bool quad);

// Forward: writes g[m] = scale * sum (FP64). When out_map != nullptr the
// result is instead added as float into gref[out_map[m]] (reference ABI).
int pm_fwd_launch(const PmShape& s, const PmTuning& t, int precision,
                  const void* st, const void* rho, double sy, double sz,
                  double thr, double scale, double* part, double* g,
                  float* gref, const int* out_map, const NearArgs& near,
                  hipStream_t stream);

// Adjoint: out[n] = scale * sum_m K(m,n) r[m] (float or double by precision).
// With a correction mode set (FP32 only) the result always goes through `part`.
int pm_adj_launch(const PmShape& s, const PmTuning& t, int precision,
                  const void* st, const void* r, double sy, double sz,
                  double scale, double* part, void* out, const NearArgs& near,
                  hipStream_t stream);

// Grouped (multi-problem) point-mass launches, FP32 path. One persistent
// grid takes tickets from a device counter; a ticket is one workgroup's block
// of one problem. blk_start is the prefix sum of the blocks per descriptor,
// item_start that of its stations (forward reduce) or cells (adjoint reduce).
// Inputs and outputs are concatenated vectors: each descriptor carries its
// problem's cell and station offsets. Descriptors live in device memory.
struct GFwdDesc {
    const float4* st;
    double* part;
    int64_t cell_off, st_off;
    int nx, ny, n_rows, rps, m_pad, tiles_x, n_slices, m, S;
    float sy, sz, thr;
    double scale;
    NearArgs na;
};
struct GAdjDesc {
    const float4* st;
    double* part;
    int64_t cell_off, st_off;
    int m, nx, ny, segs_per_row, st_per_split, m_pad, splits, blocks_x, C;
    int klo, khi;            // near-field layers (near_layers)
    int64_t n_segs, n_cells;
    int64_t near_off, n_near; // cells of layers klo..khi: first cell, count
    float sy, sz;
    double scale;
    NearArgs na;
};
struct GroupedPass {
    const void* desc = nullptr;          // GFwdDesc[np] or GAdjDesc[np] (device)
    const int64_t* blk_start = nullptr;  // device, np + 1
    const int64_t* item_start = nullptr; // device, np + 1
    const int64_t* near_start = nullptr; // device, np + 1 (adjoint: near-field cells)
    unsigned* ticket = nullptr;          // device counter, zero between passes
    int np = 0;
    int max_variant = 0;                 // largest S (forward) or C (adjoint) present
    int per_cu = 8;                      // resident workgroups per CU launched
    int64_t total_blocks = 0, total_items = 0, total_near = 0;
    size_t lds = 0;
};
// Sizes the plan's workspace for both grouped passes (call before building
// descriptors: they hold the workspace address).
int pm_grouped_workspace(ztm_plan p);
int pm_fwd_desc(ztm_plan p, int64_t cell_off, int64_t st_off, GFwdDesc* d, size_t* lds,
                int64_t* blocks);
int pm_adj_desc(ztm_plan p, int64_t cell_off, int64_t st_off, GAdjDesc* d, size_t* lds,
                int64_t* blocks);
int pm_fwd_grouped_launch(const GroupedPass& gp, const float* rho, double* g, float* gf,
                          hipStream_t stream);
int pm_adj_grouped_launch(const GroupedPass& gp, const float* r, float* out,
                          hipStream_t stream);

// Producer of the gravity path (gfm_tfm.hip): TFM.cpp:83-89 on the device.
// Frame of the reference's ENU grid: ENU -> ECEF rotation m (row-major),
// ECEF reference point, ENU offsets RC[3..5] and cell edge RC[6] (km).
struct TfmFrame {
    double m[9];
    double ref[3];
    double off[3];
    double h;
};
TfmFrame tfm_frame(const double* rc);
int tfm_drho_launch(int di, int dj, int dk, int fh, int fv, const double* rc, const double* sea,
                    const double* moho, const double* vel, double vm, double ro, double dr,
                    float* drho, hipStream_t stream);
// B variants at once: moho/vel advance by their strides (elements, 0 =
// shared) per variant, params[3b..3b+2] = (vm, ro, dr); drho [B][N].
int tfm_drho_batched_launch(int di, int dj, int dk, int fh, int fv, int B, const double* rc,
                            const double* sea, const double* moho, int64_t moho_stride,
                            const double* vel, int64_t vel_stride, const double* params,
                            float* drho, hipStream_t stream);
// g [B][M] = 0 except g[b][out_map[i]] = Y[b][i], i < m.
// d drho / d(vm, ro, dr) [3][N] at one parameter point (float).
int tfm_drho_jvp_launch(int di, int dj, int dk, int fh, int fv, const double* rc,
                        const double* sea, const double* moho, const double* vel, double vm,
                        double ro, double dr, float* out, hipStream_t stream);
// Tangent fields of nv directions on the Moho and velocity grids (dmoho
// [nv][di*dj], dvel [nv][di*dj*dk], either may be null): out [nv][N] =
// d drho / d(M, V) applied to them, at one parameter point (float).
int tfm_drho_jvp_grids_launch(int di, int dj, int dk, int fh, int fv, const double* rc,
                              const double* sea, const double* moho, const double* vel,
                              double vm, double ro, double dr, int nv, const double* dmoho,
                              const double* dvel, float* out, hipStream_t stream);
// Its transpose for a = G^T r on the fine raster: gmoho [di*dj] and gvel
// [di*dj*dk] (FP64; either may be null) = the gradients of r . G with
// respect to the grids; ws holds tfm_grid_vjp_workspace(...) bytes.
size_t tfm_grid_vjp_workspace(int di, int dj, int dk, int fh, int fv);
int tfm_grid_vjp_launch(int di, int dj, int dk, int fh, int fv, const double* rc,
                        const double* sea, const double* moho, const double* vel, double vm,
                        double ro, double dr, const float* a, double* ws, double* gmoho,
                        double* gvel, hipStream_t stream);
// Residual gather to the reference plan's station order, and FP64 dot
// products out[q] = sum_n a[n] f[q][n] (q < 3) in a fixed order;
// tfm_vjp_workspace: bytes for a[N], r_int[m] and the dot partials.
int tfm_gather_launch(const double* r, const int* out_map, int m, float* r_int,
                      hipStream_t stream);
int tfm_dot3_launch(const float* a, const float* f, int64_t N, double* part, double* out,
                    hipStream_t stream);
size_t tfm_vjp_workspace(int64_t N, int m);
int tfm_scatter_launch(const double* Y, const int* out_map, int m, int M, int B, double* g,
                       hipStream_t stream);

// Prism formula (gfm_prism.hip).
// FP32: per-station mid-zone correction (prism integral minus the far-field
// kernel) for cells within kMid of each station -> corr[m]; the forward then
// runs pm_fwd_launch with near.mode = kCorrPrism and near.corr = corr.
int prism_near_fwd_launch(const PmShape& s, const NearArgs& near, const float* rho,
                          double* corr, hipStream_t stream);
// The same in FP64 for ZTM_FP64_ZONED (rho double, FP64 far-field kernel).
int prism_near_fwd64_launch(const PmShape& s, const NearArgs& near, const double* rho,
                            double thr, double* corr, hipStream_t stream);
// FP64: exact closed form on every pair (the FP64 GPU path of the prism).
size_t prism64_workspace(const PmShape& s, const PmTuning& t, bool adjoint);
int prism64_fwd_launch(const PmShape& s, const PmTuning& t, const double4* st,
                       const double* rho, double sy, double sz, double thr, double scale,
                       double* part, double* g, hipStream_t stream);
int prism64_adj_launch(const PmShape& s, const PmTuning& t, const double4* st,
                       const double* r, double sy, double sz, double scale, double* part,
                       double* out, hipStream_t stream);

// Reference-mode station preparation on the device: interior stations of a
// di x dj raster with heights s[m] -> float4 records + output map.
int prep_ref_stations(const float* s, int bfr, int di, int dj, int fh, int fv,
                      float4* st, double4* st_d, int4* st_c, int* out_map,
                      int m_int, int m_pad, hipStream_t stream);

// Stored-matrix variant (gfm_stored.hip): G as FP32 [M][ld], ld = N rounded
// up to 4 (zero padding), entries scale * K(m,n) evaluated in FP64.
int64_t stored_ld(const PmShape& s);
int stored_build_launch(const PmShape& s, bool prism, const double4* st, double sy, double sz,
                        double scale, float* G, hipStream_t stream);
int stored_fwd_launch(const PmShape& s, const float* G, const float* rho, double thr,
                      double* g, hipStream_t stream);
int stored_adj_splits(const PmShape& s);
int stored_adj_launch(const PmShape& s, const float* G, const float* r, double* part, float* out,
                      hipStream_t stream);

// F <= 8 FP32 density fields (field f at rho + f * field_stride) through one
// point-mass forward that forms every pair's r^-3 once (gfm_pointmass.hip);
// g[f * m + i], each field with its FP64 near-field correction. |rho| <= thr
// is skipped by the main pass, |rho| <= near.thr by the correction. part must
// hold pm_fwd_fields_workspace(s, F) doubles.
bool pm_fwd_fields_supported(const PmShape& s, int F);
size_t pm_fwd_fields_workspace(const PmShape& s, int F);
int pm_fwd_fields_launch(const PmShape& s, int F, const float4* st, const float* rho,
                         int64_t field_stride, double sy, double sz, double thr, double scale,
                         double* part, double* g, const NearArgs& near, hipStream_t stream);

// Batched split-BF16 MFMA forward (gfm_batched.hip): P float [B][N] -> Y
// double [B][M]; ws must hold batched_workspace(s, B, ks) bytes.
size_t batched_workspace(const PmShape& s, int B, int ks, int split);
int batched_default_slices(const PmShape& s, int B);
// split: 2 or 3 BF16 parts, or 4 = two FP16 parts scaled by powers of two,
// which needs gsc [m_pad], the stations' power-of-two G scales
// (half_g_scale of each station's bound on |G|; else unused, null = none).
// variant: the kernel's issue order for the two-part splits (0, or 1 =
// batched_fwd2_kernel); same results.
int batched_fwd_launch(const PmShape& s, int B, int ks, int split, const float* st,
                       const float* P, double thr, double sy, double sz, double scale,
                       const float* gsc, void* ws, double* Y, hipStream_t stream,
                       int variant = 0, int waves = 4);
// Batched adjoint on the same scheme: R float [B][M] -> X float [B][N];
// ws must hold batched_adj_workspace(s, B, ks, split) bytes.
size_t batched_adj_workspace(const PmShape& s, int B, int ks, int split);
int batched_adj_default_slices(const PmShape& s, int B);
// The adjoint's split 4 takes one G scale, from gmax >= 0, the bound on |G|
// over all station-cell pairs (< 0 = none).
int batched_adj_launch(const PmShape& s, int B, int ks, int split, const float* st,
                       const float* R, double sy, double sz, double scale, double gmax, void* ws,
                       float* X, hipStream_t stream, int variant = 0, int waves = 4);

// Dense LA of the Gauss-Newton step (gfm_gn.hip).
size_t gn_workspace_bytes(int64_t rows, int n, int nd);
int gn_normal_launch(int64_t rows, int n, const double* J, const double* r, double* JJ,
                     double* jtr, void* work, hipStream_t stream);
int gn_steps_launch(int n, int nd, const double* JJ, const double* jtr, const double* sig,
                    const double* dm, const double* b, double* B, void* work,
                    hipStream_t stream);
int gn_inverse_diag_launch(int n, const double* JJ, double d, double* out, void* work,
                           hipStream_t stream);

// FP64 ensemble on FP64 MFMA (gfm_batched64.hip): workspace bytes and the
// K-slices it allows; the launchers (B in [1, 256], n_cells < 2^31).
size_t batched64_fwd_workspace(const PmShape& s, int B, int* ks_out = nullptr);
size_t batched64_adj_workspace(const PmShape& s, int B, int* ks_out = nullptr);
int batched64_fwd_launch(const PmShape& s, int B, const double4* st, const double* P,
                         double thr, double sy, double sz, double scale, double* ws, double* Y,
                         hipStream_t stream);
int batched64_adj_launch(const PmShape& s, int B, const double4* st, const double* R, double sy,
                         double sz, double scale, double* ws, double* X, hipStream_t stream);

// Power-of-two G scale of the FP16 parts for a bound gmax on |G|: brings it
// to at most 2^14; 1 when gmax is 0.
inline float half_g_scale(double gmax) {
    if (!(gmax > 0.0)) return 1.f;
    int e;
    std::frexp(gmax, &e);   // gmax < 2^e
    return (float)std::ldexp(1.0, std::max(-120, std::min(120, 14 - e)));
}

// out = L^T L rho for forward-difference gradients on an nx x ny x nz grid
// (gfm_reg.hip).
int ltl_launch(int precision, const void* rho, void* out, int nx, int ny, int nz,
               hipStream_t stream);

// Dominant-kernel timing (gfm_abi.hip).
struct ProfScope {
    ProfScope(const char* name, hipStream_t s);
    ~ProfScope();
    const char* name_;
    hipStream_t stream_;
    hipEvent_t a_ = nullptr, b_ = nullptr;
};

}  // namespace ztm

// A plan: one grid + one station set prepared on one device (ztm_plan_*).
struct ztm_plan_s {
    using PmShape = ztm::PmShape;
    using PmTuning = ztm::PmTuning;
    static constexpr int kColMargin = ztm::kColMargin;
    static constexpr int kNear = ztm::kNear;
    int device = 0;
    ztm_problem prob{};
    PmShape shape{};
    PmTuning tuning{};
    double L = 1.0, sy = 1.0, sz = 1.0;
    float4* st_f = nullptr;      // [m_pad] prepared stations, FP32
    double4* st_d = nullptr;     // [m_pad] prepared stations, FP64
    int4* st_c = nullptr;        // [m_pad] nearest cell index per station
    int* col_start = nullptr;    // near-field gather index (general plans)
    int* col_st = nullptr;
    int kmin = 0, kmax = -1;
    int* row_start = nullptr;    // clamped-row gather index (general plans)
    int* row_st = nullptr;
    int jr = 1, kr = 1, rmy = kColMargin, rmz = kColMargin;
    int rx = kNear, ry = kNear, rz = kNear;      // correction box radii
    int nrx = kNear, nry = kNear, nrz = kNear;   // FP64 near-field radii
    int cmx = kColMargin, cmy = kColMargin;      // column-index margins
    int clamp_margin = kColMargin;               // nearest-index clamp
    int rkmin = 1 << 29, rkmax = -(1 << 29);
    int* out_map = nullptr;      // reference mode: slot -> g index
    double* part = nullptr;      // split-K workspace
    size_t part_cap = 0;         // doubles
    float* G = nullptr;          // stored matrix [m][stored_ld], or null
    double gmax = -1.0;          // bound on |G| (kernel units) over station-cell pairs; < 0: none
    bool fp16_adjoint = false;   // the FP16 split is accurate for the batched adjoint here
    float* st_gs = nullptr;      // [m_pad] per-station FP16 G scales, or null
    void* bws = nullptr;         // batched forward workspace
    size_t bws_cap = 0;          // bytes
    std::atomic<int> refs{1};    // ztm_plan_retain / ztm_plan_destroy
};

namespace ztm {
// Plan helpers (gfm_abi.hip).
// Reference geometry (GFM_GPU.cu:16-24): argument check, and the problem and
// shape of its plan set in p without allocating (no device buffers).
int ref_check(int bfr, int di, int dj, int dk, int fh, int fv);
int ref_plan_init(ztm_plan p, int dev, int bfr, int di, int dj, int dk, int fh, int fv);
// Device address of this device's fixed scratch array (gfm_noalloc.hip),
// looked up once per device with that device current.
int noalloc_setup(int dev, char** out);
// While a ProfTimingOff lives on a thread, ProfScope records no events there
// (calls that must not create runtime objects).
struct ProfTimingOff {
    ProfTimingOff();
    ~ProfTimingOff();
};
// Clears bytes (a multiple of 4, 4-byte aligned) with a kernel: graph-safe.
int zero_fill(void* p, size_t bytes, hipStream_t stream);
int grid_check(const ztm_grid& g);
int ensure_workspace(ztm_plan p, size_t doubles);
void plan_free(ztm_plan p);
int set_shape(ztm_plan p, int64_t m);
NearArgs near_args(const ztm_plan p);
NearArgs near_args64(const ztm_plan p);   // ZTM_FP64_ZONED prism zones
double pm_scale(const ztm_plan p);
// The split the batched adjoint runs: the plan's, except the FP16 split on a
// plan whose station bounds spread too far for it (then the BF16 three-way).
int batched_adjoint_split(const ztm_plan p);
// The kernel setting a batched call of n_models runs in `direction` (0
// forward, 1 adjoint): the class entry, with variant 0 / 4 waves for the
// BF16 three-way split.
BatchedKernel batched_kernel_in_force(const ztm_plan p, int direction, int n_models);
PmTuning effective_tuning(ztm_plan plan, int precision);
}  // namespace ztm
