/*
 * ztm_gfm.h -- C ABI of the MI355X (gfx950) gravity forward/adjoint library.
 *
 * The library computes the vertical attraction g_z at every station as the sum,
 * over all cells of a regular 3-D density-contrast grid, of each cell's
 * attraction, and the transpose of that map (the adjoint G^T r).
 *
 * Reference interfaces replaced (abatchev/ZTM):
 *   ztm_gf        <- kernel GF(int bfr,int di,int dj,int dk,int fh,int fv,
 *                            float*g,float*rho,float*s)        GFM_GPU.cu:16-22,
 *                    launched as GF<<<N/ii,ii>>>(...)          GFM_GPU.cu:31
 *   ztm_gfm_host  <- the GFM program body (file buffers -> device -> kernel ->
 *                    host buffers)                              GFM_GPU.cu:24-34
 *   ztm_plan_*    <- the CPU gravity lambda GFM(j0) of FM::FM   TFM.cpp:95-99
 *                    generalised to explicit geometry, SI units, FP32/FP64,
 *                    the point-mass and the rectangular-prism cell formula,
 *                    and its adjoint (no reference counterpart for the adjoint).
 *   ztm_tfm_*     <- the gravity part of FM::FM: the drho producer and the
 *                    file/popen hop to GFM                     TFM.cpp:83-94
 *
 * Conventions
 *   - All functions return ZTM_OK (0) or a positive ZTM_E* code; the detail of
 *     the last failure on the calling thread is returned by ztm_last_error().
 *   - Pointers named *_dev are device pointers (hipMalloc'd, or a torch CUDA
 *     tensor's data_ptr()); pointers named *_host are host pointers.
 *   - `stream` is a hipStream_t passed as void*; NULL means the null stream.
 *     Every device entry point is stream-ordered and returns without waiting
 *     for the GPU, like a kernel launch.
 *   - Cell raster: n = i + nx*(j + ny*k), i fastest (GFM_GPU.cu:18 order).
 *
 * Cell formula (point mass, GFM_GPU.cu:22 / TFM.cpp:98):
 *     g[m] = kappa * sum_n rho[n] * V * (z_m - z_n) / |x_m - x_n|^3
 *   V = |dx*dy*dz| is the cell volume, lengths in the grid's length unit.
 *   With SI lengths (m), rho in kg/m^3 and kappa = ZTM_KAPPA_SI the output is
 *   in mGal. The reference's index-unit constant 66.7/(fh^2 fv) is kappa = 66.7
 *   with V = 1/(fh^2 fv) (GFM_GPU.cu:17, TFM.cpp:190 h = 10 km).
 * Cell formula (prism): V * (z_m - z_n)/r^3 is replaced by the exact volume
 *   integral of (z_m - z)/r^3 over the cell (Plouff 1976; Nagy, Papp & Benedek
 *   2000), evaluated with a station-local origin.
 */
#ifndef ZTM_GFM_H
#define ZTM_GFM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ZTM_GFM_VERSION 100

/* status codes */
#define ZTM_OK      0
#define ZTM_EINVAL  1   /* bad argument (sizes, null pointers, flags)        */
#define ZTM_EHIP    2   /* a HIP runtime call or kernel launch failed          */
#define ZTM_ENOMEM  3   /* device or host allocation failed                    */
#define ZTM_ENODEV  4   /* no usable gfx950 device                             */
#define ZTM_EIO     5   /* file missing, short or oversized (GFM program)      */

/* precisions */
#define ZTM_FP32    0   /* FP32 bulk, FP64 per-row/per-tile and final sums     */
#define ZTM_FP64    1   /* FP64 everywhere                                     */
/* FP64 arithmetic with the FP32 path's zones: for prism plans, the point-mass
 * far field (plus the quadrupole term of non-cubic cells) beyond the
 * correction box, 3x3x3 Gauss-Legendre within it and the closed form in the
 * near box, all in FP64 -- the FP32 path's algorithm at FP64 precision, for
 * comparing the two. Point-mass plans: the same as ZTM_FP64. Arrays are
 * double as for ZTM_FP64. */
#define ZTM_FP64_ZONED 2

/* cell formulas */
#define ZTM_POINT_MASS 0
#define ZTM_PRISM      1

/* 1e5 (m/s^2 -> mGal) * G (6.674e-11 m^3 kg^-1 s^-2) */
#define ZTM_KAPPA_SI   6.674e-6
/* the reference's constant: G * 1e3 (g/cc -> kg/m^3) * 1e4 (h = 10 km) * 1e5 */
#define ZTM_KAPPA_REF  66.7

typedef struct ztm_grid {
    int32_t nx, ny, nz;   /* cells along x (fastest), y, z                    */
    double  x0, y0, z0;   /* centre of cell (0,0,0)                            */
    double  dx, dy, dz;   /* signed centre-to-centre spacing (= cell edges)    */
} ztm_grid;

typedef struct ztm_problem {
    ztm_grid grid;
    int32_t  formula;     /* ZTM_POINT_MASS or ZTM_PRISM                       */
    double   kappa;       /* output scale, e.g. ZTM_KAPPA_SI                   */
    double   threshold;   /* forward skips cells with |rho| <= threshold       */
} ztm_problem;

typedef struct ztm_plan_s* ztm_plan;

/* ---------------------------------------------------------------- library */
int         ztm_version(void);
const char* ztm_strerror(int code);
const char* ztm_last_error(void);
/* Initialise the current HIP device (checks it is gfx950). Optional: every
 * entry point initialises lazily. */
int ztm_init(void);

/* ------------------------------------------- Level 1: reference kernel ABI */
/* Replaces GF (GFM_GPU.cu:16-22) and its launch (GFM_GPU.cu:31).
 * g_dev:   float[di*dj]  accumulated into (+=) at interior stations
 *          bfr <= ni < di-bfr, bfr <= nj < dj-bfr, m = ni + nj*di; the others
 *          are left untouched (the caller zeroes g, GFM_GPU.cu:27,29).
 * rho_dev: float[N], N = ii*jj*kk with ii=(di-1)*fh+1, jj=(dj-1)*fh+1,
 *          kk=(dk-1)*fv+1 (the producer's raster, TFM.cpp:83); cells with
 *          |rho| <= .001f are skipped (GFM_GPU.cu:19).
 * s_dev:   float[di*dj]  station heights k0 in k-index units (GFM_GPU.cu:21).
 * Cell n sits at (i/fh, j/fh, k/fv) and station m at (ni, nj, s[m]) in index
 * units; the constant is 66.7/(fh*fh*fv) (GFM_GPU.cu:17).
 * Unlike the reference kernel, dk is used: it fixes N, so a rho buffer of the
 * wrong size cannot be read past its end. ztm_gf runs on the null stream. */
int ztm_gf(int bfr, int di, int dj, int dk, int fh, int fv,
           float* g_dev, const float* rho_dev, const float* s_dev);
int ztm_gf_async(int bfr, int di, int dj, int dk, int fh, int fv,
                 float* g_dev, const float* rho_dev, const float* s_dev,
                 void* stream);
/* No reference counterpart: ztm_gf_async for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments, kernels and results (bit for bit), but no call
 * allocates device or host memory: the station records and split-K partials
 * live in a fixed 16 MiB device array per GPU, held until the process ends
 * and looked up once per device (by ztm_init or the first call). ZTM_EINVAL
 * when the geometry needs more than 16 MiB of it (about 2e5 interior
 * stations). Calls that overlap in time on one GPU share that array and race:
 * serialise them. */
int ztm_gf_noalloc(int bfr, int di, int dj, int dk, int fh, int fv,
                   float* g_dev, const float* rho_dev, const float* s_dev,
                   void* stream);

/* ---------------------------------------- Level 2: GFM program, host side */
/* Replaces the body of the GFM program (GFM_GPU.cu:24-34): host float
 * buffers in, g_host[di*dj] out (exterior stations 0, interior = the sum),
 * synchronous. n_rho must equal ii*jj*kk for the given dk (checked). */
int ztm_gfm_host(int bfr, int di, int dj, int dk, int fh, int fv,
                 float* g_host, const float* rho_host, int64_t n_rho,
                 const float* s_host);

/* ------------------------------------------------ general-geometry plans */
/* A plan holds one grid + one station set, prepared on the device once and
 * reused by every forward/adjoint call (the geometry is fixed through an
 * inversion, TFM.cpp:124-135). stations_host: double[n_stations][3] (x,y,z)
 * in the grid's length unit. Plans are bound to the device current at create
 * time; calls on one plan from several host threads must be serialised. */
int ztm_plan_create(ztm_plan* plan, const ztm_problem* problem,
                    int64_t n_stations, const double* stations_host);
/* A plan is reference-counted: create returns it with one reference,
 * ztm_plan_retain adds one, and ztm_plan_destroy drops one and frees the plan
 * when none is left. Holders that outlive the creator's reference (the
 * autograd nodes of the libtorch binding, which run the adjoint after the
 * caller may have destroyed its plan) retain it. */
int ztm_plan_retain(ztm_plan plan);
int ztm_plan_destroy(ztm_plan plan);
int64_t ztm_plan_num_cells(ztm_plan plan);
int64_t ztm_plan_num_stations(ztm_plan plan);
/* The problem the plan was created with (grid, formula, kappa, threshold). */
int ztm_plan_get_problem(ztm_plan plan, ztm_problem* problem);

/* Forward: g_dev[m] = kappa * sum_n rho[n] K(m,n)   (FP64 output)
 * precision ZTM_FP32: rho_dev is float[N];
 *           ZTM_FP64, ZTM_FP64_ZONED: rho_dev is double[N]. */
int ztm_plan_forward(ztm_plan plan, int precision, const void* rho_dev,
                     double* g_dev, void* stream);

/* Adjoint: out_dev[n] = kappa * sum_m K(m,n) r[m]   (threshold not applied)
 * precision ZTM_FP32: r_dev float[M], out_dev float[N];
 *           ZTM_FP64, ZTM_FP64_ZONED: r_dev double[M], out_dev double[N]. */
int ztm_plan_adjoint(ztm_plan plan, int precision, const void* r_dev,
                     void* out_dev, void* stream);

/* Stored-matrix variant (no reference counterpart; serves the reuse of one
 * geometry's G across the forward models of an inversion step,
 * TFM.cpp:124-135). ztm_plan_store materialises G once as FP32
 * [M][ceil4(N)] in device memory owned by the plan (ztm_plan_store_bytes
 * says how much; ZTM_ENOMEM if it does not fit), each entry evaluated in FP64
 * with the plan's formula and rounded once. The stored forward/adjoint then
 * stream G from HBM with FP64 accumulation:
 *   forward: rho_dev float[N] (16-byte aligned) -> g_dev double[M]
 *   adjoint: r_dev float[M] -> out_dev float[N]
 * The forward applies the plan's threshold; the store is freed by
 * ztm_plan_store_release or ztm_plan_destroy. */
int64_t ztm_plan_store_bytes(ztm_plan plan);
int ztm_plan_store(ztm_plan plan, void* stream);
int ztm_plan_store_release(ztm_plan plan);
int ztm_plan_forward_stored(ztm_plan plan, const float* rho_dev, double* g_dev,
                            void* stream);
int ztm_plan_adjoint_stored(ztm_plan plan, const float* r_dev, float* out_dev,
                            void* stream);

/* Batched (ensemble) forward on MFMA, point-mass formula (no reference
 * counterpart; many density models on one geometry, e.g. the finite-
 * difference Jacobian's columns, TFM.cpp:124-135, or an ensemble).
 * rho_dev float[n_models][N] (each model contiguous) -> g_dev double
 * [n_models][M]. G is generated on the fly in FP32, G and rho are split into
 * BF16 parts -- 3-way (default, ~2^-24, 6 MFMA products) or 2-way hi/lo
 * (~2^-16: hi*hi + hi*lo + lo*hi) -- run on v_mfma_f32_32x32x16_bf16 with
 * FP32 accumulation per K-slice and FP64 across slices. n_models <= 256 per
 * call. The plan's threshold is applied to rho. There is no FP64
 * near-field pass: results assume no station lies within about half a cell
 * of a cell centre (true for stations above the grid's top face). Up to 4
 * models instead take one FP32 point-mass pass that forms each pair's
 * r^-3 once for all of them, with the plan's FP64 near-field correction
 * (the result of ztm_plan_forward per model, to FP64 rounding of the
 * partial sums), when 4 density rows of the grid fit its LDS tile. */
int ztm_plan_forward_batched(ztm_plan plan, int n_models, const float* rho_dev,
                             double* g_dev, void* stream);
/* Batched (ensemble) adjoint on MFMA, point-mass formula (no reference
 * counterpart; the transpose of ztm_plan_forward_batched, e.g. misfit
 * gradients of an ensemble, or J^T r over many residuals): r_dev float
 * [n_vectors][M] -> out_dev float[n_vectors][N], out[b] = G^T r[b] with G
 * the plan's kernel (scale included, no threshold: the exact transpose of
 * G). Same split-BF16 scheme and split setting as the batched forward, the
 * same G entries (bit-identical FP32 evaluation); K-slices over stations
 * are chosen from the problem shape. n_vectors <= 256 per call; no FP64
 * near-field pass (same assumption as the batched forward). Up to 4 vectors,
 * where the batched forward runs its FP32 pass, instead run
 * ztm_plan_adjoint (FP32) per vector: the transpose of that forward. */
int ztm_plan_adjoint_batched(ztm_plan plan, int n_vectors, const float* r_dev,
                             float* out_dev, void* stream);
/* Dense LA of the Gauss-Newton step, FP64 on the GPU. The reference runs it
 * on the CPU in zmath: MTM (zmath.cpp:260-262) and the Gauss-Jordan inv
 * (zmath.cpp:129-140), called by LMA::fit (TFM.cpp:159-161), create
 * (TFM.cpp:150) and annihilate (TFM.cpp:173). Row-major device arrays; work_dev
 * is a caller-provided device workspace of ztm_gn_workspace_bytes(rows, n, nd)
 * bytes (no allocation inside the calls). */
int64_t ztm_gn_workspace_bytes(int64_t rows, int n, int nd);
/* JJ = J^T J [n][n] (MTM(J), TFM.cpp:159) and, when r_dev is given,
 * jtr = J^T r [n] (T(J)*r, TFM.cpp:160), for J [rows][n]. With rows = 0 both
 * outputs are zero (jtr_dev, when given, is cleared even if r_dev is null). */
int ztm_gn_normal(int64_t rows, int n, const double* J_dev, const double* r_dev,
                  double* JJ_dev, double* jtr_dev, void* work_dev, void* stream);
/* The damped Gauss-Newton steps of LMA::fit (TFM.cpp:160-161) for nd damping
 * conditions dm [nd][2]: B[k] = b + inv(JJ + dm[k][0] diag(diag(JJ)) +
 * dm[k][1] diag(1/sig^2)) jtr, B [nd][n]. Solved by Cholesky, the systems
 * being symmetric positive definite when a damping term is positive; a system
 * whose factorisation meets a pivot that is not positive gives B[k] = b, as
 * the reference's zero inverse of a singular matrix does. n <= 7680. */
int ztm_gn_damped_steps(int n, int nd, const double* JJ_dev, const double* jtr_dev,
                        const double* sig_dev, const double* dm_dev, const double* b_dev,
                        double* B_dev, void* work_dev, void* stream);
/* out = diag(inv(JJ + I d)) (TFM.cpp:150, :173); all 0 when JJ + I d is not
 * positive definite (the reference's loop then raises d). n <= 7680. */
int ztm_gn_inverse_diag(int n, const double* JJ_dev, double d, double* out_dev,
                        void* work_dev, void* stream);

/* MFMA pipe calibration (no reference counterpart; a measurement harness for
 * the MFMA kernels' rooflines): the FLOP rate this chip sustains on one MFMA
 * shape with nothing else in the loop -- every wave streams MFMAs on
 * register operands into independent accumulators over 8 workgroups of 4
 * waves per CU -- with operands of real entropy (low_entropy = 0: uniform
 * random values) or of low entropy (1: values 1 and -0.5). op 0:
 * v_mfma_f32_32x32x16_f16, 1: v_mfma_f32_32x32x16_bf16, 2:
 * v_mfma_f32_16x16x32_f16. Under the power limit the sustained rate depends
 * on the operands. Allocates and frees its own buffers and synchronises the
 * stream (a measurement call). */
int ztm_calib_mfma(int op, int low_entropy, int iters, double* tflops, void* stream);

/* FP64 ensemble (no reference counterpart; the FP64 form of the batched
 * ensemble, for callers that need FP64 and as its same-algorithm FP64
 * comparison): G generated once per tile in FP64, reused across the whole
 * batch on FP64 MFMA, double inputs, accumulation and outputs. Point-mass
 * plans; n <= 256 per call; no threshold on r (the transpose), the plan's
 * threshold on rho. Forward: rho_dev double [n_models][N] -> g_dev double
 * [n_models][M]. Adjoint: r_dev double [n_vectors][M] -> out_dev double
 * [n_vectors][N]. G is the exact point mass dz/r^3 on every pair, as in the
 * FP64 single-model calls, which these match to FP64 rounding. */
int ztm_plan_forward_batched64(ztm_plan plan, int n_models, const double* rho_dev,
                               double* g_dev, void* stream);
int ztm_plan_adjoint_batched64(ztm_plan plan, int n_vectors, const double* r_dev,
                               double* out_dev, void* stream);
/* Batched tuning: K-slices (0 = choose from the problem shape), in both
 * directions and every class: each slice accumulates in FP32 on the MFMAs, whose accumulate
 * loses a biased fraction of an ulp per 16-cell step, so the error grows
 * with the run length (about 3e-9 of max |.| per step; the default keeps a
 * run at most 384 steps), and the slices are summed in FP64; and
 * split (0 = keep): 2 or 3 BF16 parts, or 4 = two FP16 parts, G scaled by
 * powers of two from the plan's bounds on |G| and each model / residual vector
 * by a power of two from its largest entry (22 significant bits in 3 MFMA
 * products, against 24 in the BF16 three-way split's 6). Plans made by
 * ztm_plan_create start at split 4; the reference-geometry plans of the
 * ztm_tfm_* calls, whose stations change per call, at 3. The forward
 * scales G per station. The adjoint takes one G scale per plan (stations are
 * its MFMA K dimension), so it runs split 4 only where the per-station bounds
 * on |G| lie within 2^19 / M of each other (stations at similar heights over
 * the grid; the FP16 error then stays near 2^-20 of max |G^T r|), and the
 * BF16 three-way split on other plans. ZTM_EINVAL for other values. No
 * reference counterpart (the reference has no batched path). */
int ztm_plan_set_batched_tuning(ztm_plan plan, int slices, int split);
/* Kernel variant of the batched forward and adjoint for the two-part splits
 * (FP16 hi/lo and BF16 2-way), in every class: 0 or 1, two issue orders of
 * the same MFMA products in the same order per accumulator, so the results
 * are equal bit for bit (the FP16 adjoint with one K-slice to ~1e-7:
 * variant 1 applies the per-vector factors in FP32 in its epilogue). No
 * reference counterpart. */
int ztm_plan_set_batched_variant(ztm_plan plan, int variant);
/* One batched kernel setting, per direction (0 forward, 1 adjoint) and
 * model-count class (the class of n_models: up to 32, 64, 128 or 256
 * models): kernel variant (0, 1, or 2 = variant 1 with the forward's P
 * loads two chunks ahead; the adjoint runs 2 as 1), waves per workgroup (4,
 * or 8 with variants 1 and 2: 128 or 256 stations / cells per workgroup) and
 * K-slices (0 = automatic). Variant and waves apply to the two-part splits;
 * the BF16 three-way split always runs variant 0 with 4 waves. All variants
 * give the same products in the same order. No reference counterpart. */
int ztm_plan_set_batched_kernel(ztm_plan plan, int direction, int n_models, int variant,
                                int waves, int slices);
int ztm_plan_get_batched_kernel(ztm_plan plan, int direction, int n_models, int* variant,
                                int* waves, int* slices);
/* Autotuner of the batched (ensemble) kernels: times every candidate setting
 * of the forward and of the adjoint for n_models models on this plan, reps
 * calls each on `stream` (variant 0; variants 1 and, for the forward, 2 with
 * 4 and 8 waves per workgroup; the automatic K-slice count and twice it; fewer slices would
 * lengthen the FP32 accumulation runs), and keeps the fastest per direction
 * for the class of n_models (ztm_plan_get_batched_kernel reports it). The
 * split is not a candidate: tuning changes speed, not precision. Allocates
 * and frees its own inputs (a set-up call; synchronises the stream). */
int ztm_plan_autotune_batched(ztm_plan plan, int n_models, int reps, void* stream);
/* The batched tuning in force for 256 models: K-slices (0 = automatic), the
 * split of the forward, the split the adjoint runs on this plan, and the
 * forward's kernel variant. */
int ztm_plan_get_batched_tuning(ztm_plan plan, int* slices, int* split, int* adjoint_split,
                                int* variant);

/* Smoothness operator of the regularised misfit 0.5||G rho-d||^2 +
 * 0.5 lambda ||grad rho||^2 (forward differences on each axis, none across
 * the boundary): out_dev = L^T L rho, so the misfit gradient is
 * G^T (G rho - d) + lambda * out. Same precision convention as the adjoint
 * (float or double rho/out). No reference counterpart (TFM.cpp:124-135 uses
 * a finite-difference Jacobian and no smoothness prior). */
int ztm_plan_smoothness(ztm_plan plan, int precision, const void* rho_dev,
                        void* out_dev, void* stream);

/* Tuning of the point-mass kernels (0 = keep the plan's current value).
 * st_per_lane: stations held in registers per lane of the forward kernel
 *              (1, 2, 4 or 8); rows_per_slice: grid rows (fixed j,k) staged in
 *              LDS per forward workgroup; cells_per_lane: cells per lane of the
 *              adjoint kernel (4, 8 or 16); adj_splits: station slices of the
 *              adjoint (split-K over stations). */
int ztm_plan_set_tuning(ztm_plan plan, int st_per_lane, int rows_per_slice,
                        int cells_per_lane, int adj_splits);
int ztm_plan_get_tuning(ztm_plan plan, int* st_per_lane, int* rows_per_slice,
                        int* cells_per_lane, int* adj_splits);

/* ------------------------------------ gravity producer (TFM.cpp:83-94) */
/* The density perturbation ZTM renders before every gravity call
 * (TFM.cpp:83-89), on the device: fine raster ii x jj x kk (ii=(di-1)fh+1,
 * jj=(dj-1)fh+1, kk=(dk-1)fv+1), coarse grids sea[di*dj] (sea-level k),
 * moho[di*dj] (rendered Moho k, TFM.cpp:66 M), vel[di*dj*dk] (velocity,
 * TFM.cpp:65 v), parameters vm, ro (already negated, TFM.cpp:84), dr, and
 * rc = RC[0..6] of TFM.cpp:211 (lat_ref, lon_ref, d_ref, ENU offsets, h).
 * Cells failing |r0-r1| > .001 or depth > 10 km are 0. drho_dev[N] is float,
 * equal to (float) of the reference's double value. Device pointers. */
int ztm_tfm_drho(int di, int dj, int dk, int fh, int fv, const double* rc, const double* sea_dev,
                 const double* moho_dev, const double* vel_dev, double vm, double ro, double dr,
                 float* drho_dev, void* stream);
/* The whole gravity forward model of one FM call, TFM.cpp:83-94 (GPU
 * branch) in process: producer and GF on the device, no files, no popen.
 * Host arrays in; G[di*dj] out (gravity at interior stations, 0 outside, as
 * the reference's G after :94); drho_out (optional, float[N]) receives the
 * rendered field. surf[di*dj] = topo + sea (station k heights, TFM.cpp:83).
 * Not re-entrant per geometry (cached device buffers). */
int ztm_tfm_gravity_host(int bfr, int di, int dj, int dk, int fh, int fv, const double* rc,
                         const double* sea, const double* moho, const double* vel,
                         const double* surf, double vm, double ro, double dr, double* G,
                         float* drho_out);

/* The gravity of n_models <= 256 FM variants in one call: the columns a
 * finite-difference Jacobian (TFM.cpp:124-135) needs for the parameters the
 * gravity depends on. Variant b renders drho from moho_dev + b*moho_stride,
 * vel_dev + b*vel_stride (strides in elements; 0 = shared) and
 * params_dev[3b..3b+2] = (vm, ro, dr), into drho_dev[b][N] (float, as
 * ztm_tfm_drho); its gravity at the interior stations (station heights s_dev,
 * float[di*dj]) goes to g_dev[b][di*dj] (FP64; 0 outside the interior),
 * through the split-BF16 MFMA ensemble forward: no near-field pass, so its
 * ~1e-6 relative accuracy holds for stations at least half a cell from any
 * kept cell; the 10 km depth filter keeps cells that far below stations at
 * or above the -0.5 topography floor (TFM.cpp:89,225). No reference
 * counterpart. Device pointers. */
int ztm_tfm_gravity_batched(int bfr, int di, int dj, int dk, int fh, int fv, int n_models,
                            const double* rc, const double* sea_dev, const double* moho_dev,
                            int64_t moho_stride, const double* vel_dev, int64_t vel_stride,
                            const double* params_dev, const float* s_dev, float* drho_dev,
                            double* g_dev, void* stream);

/* Exact gravity columns of the Jacobian for the three gravity parameters.
 * drho is piecewise linear in (vm, ro, dr) (TFM.cpp:86-88), so its
 * derivative at one parameter point is rendered directly:
 * ddrho_dev[3][N] = d drho / d(vm, ro, dr) (float), on the cells the producer
 * keeps (TFM.cpp:89) and GF does not skip (|float drho| > .001,
 * GFM_GPU.cu:19), decided at that point; 0 elsewhere. ro is the value the
 * producer uses (ro = -B[4nl+nt+2], TFM.cpp:84), so the column for
 * B[4nl+nt+2] is minus the ro column. No reference counterpart: TFM.cpp:135
 * builds these columns by finite differences, which differ only by the cells
 * whose piece or filter changes within the step. Device pointers. */
int ztm_tfm_drho_jvp(int di, int dj, int dk, int fh, int fv, const double* rc,
                     const double* sea_dev, const double* moho_dev, const double* vel_dev,
                     double vm, double ro, double dr, float* ddrho_dev, void* stream);
/* The gravity of those three fields at the interior stations (station
 * heights s_dev, float[di*dj]): dg_dev[3][di*dj] = dG / d(vm, ro, dr) (FP64;
 * 0 outside the interior), one FP32 GF forward per field with its FP64
 * near-field correction (as ztm_gf, without GF's skip, which the fields have
 * already had); ddrho_dev[3][N] receives the fields. */
int ztm_tfm_gravity_jvp(int bfr, int di, int dj, int dk, int fh, int fv, const double* rc,
                        const double* sea_dev, const double* moho_dev, const double* vel_dev,
                        double vm, double ro, double dr, const float* s_dev, float* ddrho_dev,
                        double* dg_dev, void* stream);
/* Its adjoint: out_dev[3] = sum over interior stations m of r_dev[m] *
 * dG[m]/d(vm, ro, dr) (FP64), the gradient of r . G for a station residual
 * r_dev[di*dj] (exterior entries ignored), computed as the FP64 dot products
 * of G^T r (FP32 adjoint) with the three derivative fields (ddrho_dev[3][N]
 * receives them). One adjoint pass instead of three forwards. */
int ztm_tfm_gravity_vjp(int bfr, int di, int dj, int dk, int fh, int fv, const double* rc,
                        const double* sea_dev, const double* moho_dev, const double* vel_dev,
                        double vm, double ro, double dr, const float* s_dev, const double* r_dev,
                        float* ddrho_dev, double* out_dev, void* stream);

/* Exact gravity columns for tangent Moho and velocity grids. drho depends on
 * the rendered Moho m1 = interp2(M) and velocity v = interp3(V) of each cell
 * (TFM.cpp:86-88), piecewise linearly, and interp2/interp3 are linear in the
 * grid values, so at one point (M, V, vm, ro, dr) the derivative along a
 * tangent direction (dM, dV) is rendered directly: per kept cell (as
 * ztm_tfm_drho_jvp decides them), d drho = dr dv below the band,
 * -(dr (v - vm) + ro) dm1 / 2li + dr (1 + t1) dv in it (t1 = (k - m1 + li)
 * / 2li, li = 1/fv) and 0 above it, with dm1, dv the tangent grids
 * interpolated like M and V. For n_dirs <= 256 directions dmoho_dev
 * [n_dirs][di*dj] and dvel_dev [n_dirs][di*dj*dk] (either may be null: no
 * tangent on that grid): ddrho_dev [n_dirs][N] receives the tangent fields and
 * dg_dev [n_dirs][di*dj] (FP64, 0 outside the interior) their gravity at the
 * stations s_dev: one FP32 GF forward for up to 4 directions (FP64 near-field
 * correction), the split MFMA ensemble forward for more. No reference
 * counterpart: TFM.cpp:134 builds each node's column by a finite-difference
 * forward model. Device pointers. */
int ztm_tfm_gravity_jvp_grids(int bfr, int di, int dj, int dk, int fh, int fv, const double* rc,
                              const double* sea_dev, const double* moho_dev,
                              const double* vel_dev, double vm, double ro, double dr,
                              const float* s_dev, int n_dirs, const double* dmoho_dev,
                              const double* dvel_dev, float* ddrho_dev, double* dg_dev,
                              void* stream);
/* Its adjoint: for a station residual r_dev[di*dj] (FP64, exterior entries
 * ignored), gmoho_dev[di*dj] = d(r . G)/dM and gvel_dev[di*dj*dk] =
 * d(r . G)/dV (FP64; either may be null), from one FP32 adjoint G^T r and the
 * transpose of the tangent-field map (per node, a fixed-order sum over the
 * fine cells its interpolation stencil reaches). The gradient of the
 * misfit over every Moho and velocity node in one pass, where TFM.cpp:134
 * runs one forward model per node. */
int ztm_tfm_gravity_vjp_grids(int bfr, int di, int dj, int dk, int fh, int fv, const double* rc,
                              const double* sea_dev, const double* moho_dev,
                              const double* vel_dev, double vm, double ro, double dr,
                              const float* s_dev, const double* r_dev, double* gmoho_dev,
                              double* gvel_dev, void* stream);

/* ------------------------------------------------- uneven batches */
/* A batch of independent point-mass problems (each its own grid, stations and
 * kappa; FP32 path with the FP64 near-field corrections). Every forward or
 * adjoint pass over the whole batch is one persistent launch that takes
 * (problem, tile) tickets from a device counter, plus one reduce launch.
 * Vectors are concatenated: problem s's cells start at cell_offsets[s], its
 * stations at station_offsets[s] (ztm_batch_layout; n_problems + 1 entries,
 * the last is the total length). Entries between one problem's end and the
 * next start are padding: never read, never written.
 * No reference counterpart: ZTM runs one CPU forward model per Jacobian
 * column (TFM.cpp:124-135). Calls on one batch must not overlap in time. */
typedef struct ztm_batch_s* ztm_batch;

int ztm_batch_create(ztm_batch* batch, int n_problems, const ztm_problem* problems,
                     const int64_t* n_stations, const double* const* stations_host);
int ztm_batch_destroy(ztm_batch batch);
int ztm_batch_num_problems(ztm_batch batch);
int ztm_batch_layout(ztm_batch batch, int64_t* cell_offsets, int64_t* station_offsets);
/* g_dev (FP64) at station_offsets; rho_dev (FP32) at cell_offsets. */
int ztm_batch_forward(ztm_batch batch, const float* rho_dev, double* g_dev, void* stream);
/* out_dev (FP32) at cell_offsets; r_dev (FP32) at station_offsets. */
int ztm_batch_adjoint(ztm_batch batch, const float* r_dev, float* out_dev, void* stream);
/* Conjugate gradients on (G^T G + lambda_s L^T L) x = G^T d for every problem
 * at once (L = forward differences inside the grid), x(0) = 0, `iters`
 * iterations; lambda is host double[n_problems]. rr_dev (optional, device
 * double[(iters + 1) * n_problems]) receives ||r_k||^2 of the normal-equation
 * residual per iteration k and problem. use_graph != 0 captures the whole
 * solve into a HIP graph on the first call and replays it while (d_dev,
 * x_dev, rr_dev, iters) are unchanged. */
int ztm_batch_solve(ztm_batch batch, const float* d_dev, const double* lambda, int iters,
                    float* x_dev, double* rr_dev, int use_graph, void* stream);
/* Tiles: per problem (problem = -1: all). st_per_lane 2/4/8, cells_per_lane
 * 8/16, rows_per_slice / adj_splits 0 = automatic. */
int ztm_batch_set_tuning(ztm_batch batch, int problem, int st_per_lane, int rows_per_slice,
                         int cells_per_lane, int adj_splits);
int ztm_batch_get_tuning(ztm_batch batch, int problem, int* st_per_lane, int* rows_per_slice,
                         int* cells_per_lane, int* adj_splits, int* size_class);
/* Time the grouped kernels per size class (station count x cell count bands)
 * over the tile candidates, `reps` passes each, and keep the fastest. */
int ztm_batch_autotune(ztm_batch batch, int reps, void* stream);

/* ------------------------------------------------------ kernel timing */
/* When enabled, HIP events are recorded on the launch stream around the
 * dominant kernel of every entry point; ztm_profile_read() waits for them and
 * returns the summed duration (ms) and count for `kernel` ("pm_fwd", "pm_adj",
 * ...) since the last reset. */
int ztm_profile_enable(int on);
int ztm_profile_read(const char* kernel, double* total_ms, int64_t* count);
int ztm_profile_reset(void);

#ifdef __cplusplus
}
#endif
#endif /* ZTM_GFM_H */
