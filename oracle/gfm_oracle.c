/*
 * gfm_oracle.c -- FP64 CPU oracle for the gravity forward/adjoint path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker: only tests/, the smoke
 * check and bench.py's CPU-baseline leg load it. The product (libztm_gfm.so,
 * the GFM program, the ztm_amd package) never links, loads or calls it.
 *
 * What each function restates (abatchev/ZTM):
 *   oracle_gf            GFM_GPU.cu:16-22 -- kernel GF on its own float
 *                        inputs (raster decode :17-18, |rho|>.001f skip :19,
 *                        interior-station loop :20-21, point-mass term :22),
 *                        evaluated in FP64 with pow(r^2,1.5); the raster width
 *                        is the producer's ii=(di-1)*fh+1 (TFM.cpp:83,
 *                        GFM_GPU.cu:24), not the kernel's di*fh (:17), which
 *                        agree for the only fh the reference uses (fh=1).
 *   oracle_tfm_gravity   TFM.cpp:95-99 -- the CPU gravity lambda GFM(j0):
 *                        FP64 drho/surf, the sparse cell list nn, G[m]=0 then
 *                        += (66.7*drho*(k0-k)/pow(r2,1.5))/(fh*fh*fv), one
 *                        std::thread per interior station row. Used as the
 *                        timed CPU baseline ("port" of the reference path).
 *   oracle_pm_forward    the same point-mass sum in explicit geometry:
 *                        g[m] = kappa * sum_n [|rho_n|>thr] rho_n V dz/r^3.
 *   oracle_pm_adjoint    its transpose, looped cells-outer (an independent
 *                        loop nest from the forward, so a transposed index
 *                        fails the dot-product test).
 *   oracle_pm_forward_cells / oracle_pm_adjoint_cells
 *                        the same sums restricted to a list of cells (bounded
 *                        samples of large grids); pinned against the full
 *                        forward on a density masked to the list, and against
 *                        the full adjoint's entries, with partition additivity.
 *   oracle_tfm_drho      TFM.cpp:83-89 -- the producer of the gravity path:
 *                        the density perturbation on the fine raster from the
 *                        surface (sea), Moho (M) and velocity (v) grids and
 *                        (vm, ro, dr), with the |r0-r1|>.001 and depth>10 km
 *                        cell filter. Its parts are separate functions:
 *   oracle_interp2       zmath.cpp:225-233, bilinear interp, unit spacing;
 *   oracle_interp3       zmath.cpp:236-249, trilinear interp, unit spacing;
 *   oracle_c2g_depth     TFM.cpp:56 + SET.cpp:86-99, the depth (km) of the
 *                        ENU grid point c, with the reference's sqrtf.
 *                        Both FP64 expressions keep the reference's operation
 *                        order, and the file is built without FMA contraction,
 *                        so results are the reference's bit for bit.
 *   oracle_prism_*       right rectangular prism, exact volume integral of
 *                        (z_m - z)/r^3 (Plouff 1976; Nagy, Papp & Benedek
 *                        2000, eq. for g_z), no reference counterpart.
 *   oracle_mtm           zmath.cpp:260-262, MTM(J) = J^T J.
 *   oracle_jtr           TFM.cpp:160, T(J)*r.
 *   oracle_inv           zmath.cpp:129-140, Gauss-Jordan inverse with row
 *                        pivoting (zero matrix on an exact zero pivot).
 *   oracle_gn_steps      TFM.cpp:159-161, the damped Gauss-Newton steps of
 *                        LMA::fit over a grid of damping conditions.
 *   oracle_inv_diag      TFM.cpp:150,173, diag(inv(J^T J + I d)).
 *
 * Deliberate deviations: (1) oracle_interp2 reads a neighbour outside the
 * grid (i = di-1 or j = dj-1, where its weight is exactly 0) as 0; the
 * reference reads past the row or the array there (zmath.cpp:229-230) and
 * multiplies by 0, which equals ours whenever that value is finite.
 * (2) a station exactly at a cell centre (r = 0) gives
 * 0*inf = NaN in the reference (GFM_GPU.cu:22, TFM.cpp:98); here that pair
 * contributes 0, as in the HIP kernels.
 *
 * How each function is pinned (tests/test_oracle.py): single-cell closed
 * forms, the 66.7 unit derivation, the solid-angle formula of a thin square
 * layer, mirror symmetry, brute force on tiny grids, oracle_gf ==
 * oracle_tfm_gravity on shared inputs, the dot-product identity and central
 * finite differences for the adjoint; the prism against Gauss-Legendre
 * quadrature of the point-mass kernel, its far-field point-mass limit,
 * 8-subprism additivity and the infinite-slab limit 2*pi*G*rho*t.
 * The dense LA functions (tests/test_oracle_gn.py): against numpy's LAPACK
 * inverse, solve and products on random and reference-shaped systems, the
 * identities inv(A) A = I and (A^T A) symmetric, a matrix that needs row
 * swaps, exact-zero-pivot singularity, and -- where the reference's own
 * zmath.cpp builds (oracle/_ref, zmath_ref.cpp) -- its MTM and inv bit for
 * bit.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct og_grid {
    int32_t nx, ny, nz;
    double x0, y0, z0;
    double dx, dy, dz;
} og_grid;

/* ------------------------------------------------------------ threading */
typedef void (*range_fn)(void* ctx, int64_t lo, int64_t hi);
typedef struct { range_fn fn; void* ctx; int64_t lo, hi; } range_job;

static void* range_tramp(void* a) {
    range_job* j = (range_job*)a;
    j->fn(j->ctx, j->lo, j->hi);
    return NULL;
}

/* Split [0,n) into nthreads contiguous chunks and run fn on each. */
static void parallel_for(int64_t n, int nthreads, range_fn fn, void* ctx) {
    if (nthreads <= 1 || n < 2) { fn(ctx, 0, n); return; }
    if (nthreads > n) nthreads = (int)n;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * nthreads);
    range_job* jb = (range_job*)malloc(sizeof(range_job) * nthreads);
    for (int t = 0; t < nthreads; ++t) {
        jb[t].fn = fn; jb[t].ctx = ctx;
        jb[t].lo = n * t / nthreads; jb[t].hi = n * (t + 1) / nthreads;
        pthread_create(&th[t], NULL, range_tramp, &jb[t]);
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(th); free(jb);
}

/* ---------------------------------------------------- oracle_gf (GF kernel) */
typedef struct {
    int bfr, di, dj, fh, fv, ii, jj, kk;
    const float* rho; const float* s; double* g;
} gf_ctx;

static void gf_rows(void* p, int64_t lo, int64_t hi) {
    const gf_ctx* c = (const gf_ctx*)p;
    const int64_t N = (int64_t)c->ii * c->jj * c->kk;
    const double cst = 66.7 / ((double)c->fh * c->fh * c->fv);
    const int ni_count = c->di - 2 * c->bfr;
    for (int64_t t = lo; t < hi; ++t) {
        const int ni = c->bfr + (int)(t % ni_count), nj = c->bfr + (int)(t / ni_count);
        const int m = ni + nj * c->di;
        const double i0 = ni, j0 = nj, k0 = c->s[m];
        double acc = 0.0;
        for (int64_t n = 0; n < N; ++n) {
            const float rf = c->rho[n];
            if (!(fabsf(rf) > .001f)) continue;                 /* GFM_GPU.cu:19 */
            const double i = (double)(n % c->ii) / c->fh;
            const double j = (double)((n / c->ii) % c->jj) / c->fh;
            const double k = (double)(n / ((int64_t)c->ii * c->jj)) / c->fv;
            const double r2 = (i0 - i) * (i0 - i) + (j0 - j) * (j0 - j) + (k0 - k) * (k0 - k);
            if (r2 == 0.0) continue;
            acc += cst * (double)rf * (k0 - k) / pow(r2, 1.5); /* GFM_GPU.cu:22 */
        }
        c->g[m] += acc;
    }
}

/* g[di*dj] is accumulated into at interior stations, like the kernel. */
void oracle_gf(int bfr, int di, int dj, int dk, int fh, int fv, const float* rho,
               const float* s, double* g, int nthreads) {
    gf_ctx c = {bfr, di, dj, fh, fv, (di - 1) * fh + 1, (dj - 1) * fh + 1, (dk - 1) * fv + 1,
                rho, s, g};
    const int ni = di - 2 * bfr, nj = dj - 2 * bfr;
    if (ni <= 0 || nj <= 0) return;
    parallel_for((int64_t)ni * nj, nthreads, gf_rows, &c);
}

/* ------------------------------------------ oracle_tfm_gravity (CPU path) */
typedef struct {
    int bfr, di, fh, fv, ii, jj;
    const double* drho; const double* surf; const int32_t* nn; int64_t n_nn; double* G;
} tfm_ctx;

/* One interior station row j0 (TFM.cpp:96-98). */
static void tfm_row(const tfm_ctx* c, int j0) {
    for (int i0 = c->bfr; i0 < c->di - c->bfr; i0 += 1) {
        const int m = i0 + j0 * c->di;
        const double k0 = c->surf[m];
        c->G[m] = 0;
        for (int64_t q = 0; q < c->n_nn; ++q) {
            const int n = c->nn[q];
            const double i = (n % c->ii) / (1. * c->fh), j = ((n / c->ii) % c->jj) / (1. * c->fh),
                         k = (n / (c->ii * c->jj)) / (1. * c->fv);
            const double r2 = (i0 - i) * (i0 - i) + (j0 - j) * (j0 - j) + (k0 - k) * (k0 - k);
            if (r2 == 0.0) continue;
            c->G[m] += (66.7 * c->drho[n] * (k0 - k) / pow(r2, 1.5)) / (c->fh * c->fh * c->fv);
        }
    }
}

typedef struct { const tfm_ctx* c; int j0; } tfm_job;
static void* tfm_tramp(void* a) { tfm_job* j = (tfm_job*)a; tfm_row(j->c, j->j0); return NULL; }
static void tfm_range(void* p, int64_t lo, int64_t hi) {
    const tfm_ctx* c = (const tfm_ctx*)p;
    for (int64_t r = lo; r < hi; ++r) tfm_row(c, c->bfr + (int)r);
}

/* nthreads <= 0: one thread per interior station row, as TFM.cpp:99;
 * otherwise rows are split over nthreads threads. Returns rows processed. */
int oracle_tfm_gravity(int bfr, int di, int dj, int dk, int fh, int fv, const double* drho,
                       const double* surf, const int32_t* nn, int64_t n_nn, double* G,
                       int nthreads) {
    (void)dk;
    tfm_ctx c = {bfr, di, fh, fv, (di - 1) * fh + 1, (dj - 1) * fh + 1, drho, surf, nn, n_nn, G};
    const int rows = dj - 2 * bfr;
    if (rows <= 0 || di - 2 * bfr <= 0) return 0;
    if (nthreads <= 0) {
        pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * rows);
        tfm_job* jb = (tfm_job*)malloc(sizeof(tfm_job) * rows);
        for (int j = 0; j < rows; ++j) {
            jb[j].c = &c; jb[j].j0 = j + bfr;
            pthread_create(&th[j], NULL, tfm_tramp, &jb[j]);
        }
        for (int j = 0; j < rows; ++j) pthread_join(th[j], NULL);
        free(th); free(jb);
    } else {
        parallel_for(rows, nthreads, tfm_range, &c);
    }
    return rows;
}

/* -------------------------------------------- oracle_tfm_drho (producer) */
/* zmath.cpp:225-233 with hi = 1 (hj = hi), x0 = y0 = 0. */
double oracle_interp2(const double* f, double x, double y, int di, int dj) {
    const double ii = fmin(fmax(x / 1.0, 0), (di - 1)), jj = fmin(fmax(y / 1.0, 0), (dj - 1));
    const int i = (int)ii, j = (int)jj;
    const double si = ii - i, sj = jj - j;
    const double c00 = f[i + (j + 0) * di];
    const double c10 = i + 1 < di ? f[i + 1 + (j + 0) * di] : 0.0;
    const double c01 = j + 1 < dj ? f[i + (j + 1) * di] : 0.0;
    const double c11 = (i + 1 < di && j + 1 < dj) ? f[i + 1 + (j + 1) * di] : 0.0;
    const double c0 = c00 * (1 - si) + c10 * si, c1 = c01 * (1 - si) + c11 * si;
    return c0 * (1 - sj) + c1 * sj;
}

/* zmath.cpp:236-249 with hi = 1 (hj = hk = hi), zero offsets. */
double oracle_interp3(const double* f, double x, double y, double z, int di, int dj, int dk) {
    const double ii = fmin(fmax(x / 1.0, 0), (di - 2)), jj = fmin(fmax(y / 1.0, 0), (dj - 2)),
                 kk = fmin(fmax(z / 1.0, 0), (dk - 2));
    const int i = (int)ii, j = (int)jj, k = (int)kk;
    const double si = ii - i, sj = jj - j, sk = kk - k;
    const int64_t p = (int64_t)di * dj;
    const double c000 = f[i + (j + 0) * di + (k + 0) * p], c100 = f[i + 1 + (j + 0) * di + (k + 0) * p];
    const double c010 = f[i + (j + 1) * di + (k + 0) * p], c110 = f[i + 1 + (j + 1) * di + (k + 0) * p];
    const double c001 = f[i + (j + 0) * di + (k + 1) * p], c101 = f[i + 1 + (j + 0) * di + (k + 1) * p];
    const double c011 = f[i + (j + 1) * di + (k + 1) * p], c111 = f[i + 1 + (j + 1) * di + (k + 1) * p];
    const double c00 = c000 * (1 - si) + c100 * si, c01 = c001 * (1 - si) + c101 * si;
    const double c10 = c010 * (1 - si) + c110 * si, c11 = c011 * (1 - si) + c111 * si;
    const double c0 = c00 * (1 - sj) + c10 * sj, c1 = c01 * (1 - sj) + c11 * sj;
    return c0 * (1 - sk) + c1 * sk;
}

/* TFM.cpp:56 c2g(c) = enu2geo(RC[0..2], c*RC[6] - RC[3..5]), SET.cpp:86-99,
 * depth component only: R - sqrtf(|ECEF point|^2). rc = RC[0..6]. */
double oracle_c2g_depth(const double* rc, double ci, double cj, double ck) {
    const double rad = 57.295, R = 6371;
    const double lat_ref = rc[0], lon_ref = rc[1], d_ref = rc[2];
    const double x = ci * rc[6] - rc[3], y = cj * rc[6] - rc[4], z = ck * rc[6] - rc[5];
    const double xr = (R - d_ref) * cos(lon_ref / rad) * cos(lat_ref / rad);
    const double yr = (R - d_ref) * sin(lon_ref / rad) * cos(lat_ref / rad);
    const double zr = (R - d_ref) * sin(lat_ref / rad);
    const double phi = lat_ref / rad, theta = lon_ref / rad;
    const double m[9] = {-sin(theta), -sin(phi) * cos(theta), cos(phi) * cos(theta),
                         cos(theta),  -sin(phi) * sin(theta), cos(phi) * sin(theta),
                         0,           cos(phi),               sin(phi)};
    const double xi = m[0] * x + m[1] * y + m[2] * z + xr;
    const double yi = m[3] * x + m[4] * y + m[5] * z + yr;
    const double zi = m[6] * x + m[7] * y + m[8] * z + zr;
    const double r = sqrtf(xi * xi + yi * yi + zi * zi);
    return R - r;
}

/* TFM.cpp:83-89: drho[N] (zero outside the kept set) and the kept-cell list
 * nn (ascending n); returns the number of kept cells. nn may be NULL. */
int64_t oracle_tfm_drho(int di, int dj, int dk, int fh, int fv, const double* rc,
                        const double* sea, const double* M, const double* v, double vm,
                        double ro, double dr, double* drho, int32_t* nn) {
    const int ii = (di - 1) * fh + 1, jj = (dj - 1) * fh + 1, kk = (dk - 1) * fv + 1;
    const int64_t N = (int64_t)ii * jj * kk;
    int64_t count = 0;
    for (int64_t n = 0; n < N; ++n) {
        const double i = (n % ii) / (1. * fh), j = ((n / ii) % jj) / (1. * fh),
                     k = (n / ((int64_t)ii * jj)) / (1. * fv), li = 1. / fv;
        double m0, m1, r0, r1;
        m0 = oracle_interp2(sea, i, j, di, dj) - 3.5;
        m1 = oracle_interp2(M, i, j, di, dj);
        r0 = 0;
        r1 = dr * (oracle_interp3(v, i, j, k, di, dj, dk) - vm);
        if (k > m0 + li) { r0 = ro; } else if (fabs(k - m0) < li) { r0 = r0 + (k - m0 + li) * (r0 + ro) / (2 * li); }
        if (k > m1 + li) { r1 = ro; } else if (fabs(k - m1) < li) { r1 = r1 + (k - m1 + li) * (r1 + ro) / (2 * li); }
        drho[n] = 0;
        if (fabs(r0 - r1) > .001 && oracle_c2g_depth(rc, i, j, k) > 10) {
            drho[n] = r1 - r0;
            if (nn) nn[count] = (int32_t)n;
            ++count;
        }
    }
    return count;
}

/* ------------------------------------------------- general point mass */
static inline void cell_centre(const og_grid* g, int64_t n, double* x, double* y, double* z) {
    const int64_t i = n % g->nx, j = (n / g->nx) % g->ny, k = n / ((int64_t)g->nx * g->ny);
    *x = g->x0 + i * g->dx; *y = g->y0 + j * g->dy; *z = g->z0 + k * g->dz;
}

typedef struct {
    const og_grid* g; double kappa, thr; const double* rho; const double* xyz;
    const double* r; double* out; int64_t m;
} pm_ctx;

static void pm_fwd_range(void* p, int64_t lo, int64_t hi) {
    const pm_ctx* c = (const pm_ctx*)p;
    const og_grid* g = c->g;
    const int64_t N = (int64_t)g->nx * g->ny * g->nz;
    const double V = fabs(g->dx * g->dy * g->dz);
    for (int64_t m = lo; m < hi; ++m) {
        const double xs = c->xyz[3 * m], ys = c->xyz[3 * m + 1], zs = c->xyz[3 * m + 2];
        double acc = 0.0;
        for (int64_t n = 0; n < N; ++n) {
            const double rho = c->rho[n];
            if (!(fabs(rho) > c->thr)) continue;
            double x, y, z;
            cell_centre(g, n, &x, &y, &z);
            const double r2 = (xs - x) * (xs - x) + (ys - y) * (ys - y) + (zs - z) * (zs - z);
            if (r2 == 0.0) continue;
            acc += rho * V * (zs - z) / pow(r2, 1.5);
        }
        c->out[m] = c->kappa * acc;
    }
}

void oracle_pm_forward(const og_grid* g, double kappa, double thr, const double* rho,
                       int64_t m, const double* xyz, double* out, int nthreads) {
    pm_ctx c = {g, kappa, thr, rho, xyz, NULL, out, m};
    parallel_for(m, nthreads, pm_fwd_range, &c);
}

static void pm_adj_range(void* p, int64_t lo, int64_t hi) {
    const pm_ctx* c = (const pm_ctx*)p;
    const og_grid* g = c->g;
    const double V = fabs(g->dx * g->dy * g->dz);
    for (int64_t n = lo; n < hi; ++n) {
        double x, y, z;
        cell_centre(g, n, &x, &y, &z);
        double acc = 0.0;
        for (int64_t m = 0; m < c->m; ++m) {
            const double dx = c->xyz[3 * m] - x, dy = c->xyz[3 * m + 1] - y,
                         dz = c->xyz[3 * m + 2] - z;
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 == 0.0) continue;
            acc += c->r[m] * V * dz / pow(r2, 1.5);
        }
        c->out[n] = c->kappa * acc;
    }
}

void oracle_pm_adjoint(const og_grid* g, double kappa, int64_t m, const double* xyz,
                       const double* r, double* out, int nthreads) {
    pm_ctx c = {g, kappa, 0.0, NULL, xyz, r, out, m};
    parallel_for((int64_t)g->nx * g->ny * g->nz, nthreads, pm_adj_range, &c);
}

/* Adjoint restricted to the listed cells: out[q] for cell cells[q]. */
typedef struct {
    const og_grid* g; double kappa; const int64_t* cells; const double* xyz; const double* r;
    double* out; int64_t m;
} pm_adj_sub_ctx;

static void pm_adj_sub_range(void* p, int64_t lo, int64_t hi) {
    const pm_adj_sub_ctx* c = (const pm_adj_sub_ctx*)p;
    const double V = fabs(c->g->dx * c->g->dy * c->g->dz);
    for (int64_t q = lo; q < hi; ++q) {
        double x, y, z;
        cell_centre(c->g, c->cells[q], &x, &y, &z);
        double acc = 0.0;
        for (int64_t m = 0; m < c->m; ++m) {
            const double dx = c->xyz[3 * m] - x, dy = c->xyz[3 * m + 1] - y,
                         dz = c->xyz[3 * m + 2] - z;
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 == 0.0) continue;
            acc += c->r[m] * V * dz / pow(r2, 1.5);
        }
        c->out[q] = c->kappa * acc;
    }
}

void oracle_pm_adjoint_cells(const og_grid* g, double kappa, int64_t m, const double* xyz,
                             const double* r, const int64_t* cells, int64_t nc, double* out,
                             int nthreads) {
    pm_adj_sub_ctx c = {g, kappa, cells, xyz, r, out, m};
    parallel_for(nc, nthreads, pm_adj_sub_range, &c);
}

/* Forward over a subset of cells (cells[q], q < nc) for all stations: the
 * unit of the CPU baseline on a bounded sample. */
typedef struct {
    const og_grid* g; double kappa; const double* rho; const int64_t* cells; int64_t nc;
    const double* xyz; double* out;
} pm_sub_ctx;

static void pm_sub_range(void* p, int64_t lo, int64_t hi) {
    const pm_sub_ctx* c = (const pm_sub_ctx*)p;
    const double V = fabs(c->g->dx * c->g->dy * c->g->dz);
    for (int64_t m = lo; m < hi; ++m) {
        const double xs = c->xyz[3 * m], ys = c->xyz[3 * m + 1], zs = c->xyz[3 * m + 2];
        double acc = 0.0;
        for (int64_t q = 0; q < c->nc; ++q) {
            double x, y, z;
            cell_centre(c->g, c->cells[q], &x, &y, &z);
            const double r2 = (xs - x) * (xs - x) + (ys - y) * (ys - y) + (zs - z) * (zs - z);
            if (r2 == 0.0) continue;
            acc += c->rho[c->cells[q]] * V * (zs - z) / pow(r2, 1.5);
        }
        c->out[m] = c->kappa * acc;
    }
}

void oracle_pm_forward_cells(const og_grid* g, double kappa, const double* rho,
                             const int64_t* cells, int64_t nc, int64_t m, const double* xyz,
                             double* out, int nthreads) {
    pm_sub_ctx c = {g, kappa, rho, cells, nc, xyz, out};
    parallel_for(m, nthreads, pm_sub_range, &c);
}

/* ------------------------------------------------------------ prism */
/* Antiderivative of the g_z kernel over a box in station-relative
 * coordinates (u,v,w) = (x,y,z) - station:
 *   F = u ln(v + r) + v ln(u + r) - w atan(u v / (w r))
 * with the principal-value arctangent (an atan2 would jump by pi across
 * w r < 0). The 8-corner alternating sum sum (-1)^(a+b+c) F(u_a,v_b,w_c)
 * equals the integral of w/r^3 over the box (Nagy et al. 2000), so the
 * integral of (z_m - z)/r^3 = -w/r^3 is minus that sum.
 * Terms whose prefactor is exactly 0 are dropped (their log may diverge). */
/* ln(v + r) for r = |(u,v,w)|; for v < 0, v + r = (u^2 + w^2)/(r - v) avoids
 * the cancellation of r and -v (stations beside the box, far away). */
static double log_vr(double v, double r, double rest2) {
    return v >= 0.0 ? log(v + r) : log(rest2 / (r - v));
}

static double prism_F(double u, double v, double w) {
    const double u2 = u * u, v2 = v * v, w2 = w * w;
    const double r = sqrt(u2 + v2 + w2);
    double f = 0.0;
    if (u != 0.0) f += u * log_vr(v, r, u2 + w2);
    if (v != 0.0) f += v * log_vr(u, r, v2 + w2);
    if (w != 0.0) f -= w * atan(u * v / (w * r));
    return f;
}

/* Integral over the box of (z_s - z)/r^3 for station (xs,ys,zs). */
double oracle_prism_kernel(double xs, double ys, double zs, double x1, double x2, double y1,
                           double y2, double z1, double z2) {
    const double u[2] = {x1 - xs, x2 - xs}, v[2] = {y1 - ys, y2 - ys}, w[2] = {z1 - zs, z2 - zs};
    double s = 0.0;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int c = 0; c < 2; ++c) {
                const double sg = ((a + b + c) & 1) ? -1.0 : 1.0;
                s += sg * prism_F(u[a], v[b], w[c]);
            }
    return -s;
}

static void prism_fwd_range(void* p, int64_t lo, int64_t hi) {
    const pm_ctx* c = (const pm_ctx*)p;
    const og_grid* g = c->g;
    const int64_t N = (int64_t)g->nx * g->ny * g->nz;
    const double hx = 0.5 * fabs(g->dx), hy = 0.5 * fabs(g->dy), hz = 0.5 * fabs(g->dz);
    for (int64_t m = lo; m < hi; ++m) {
        const double xs = c->xyz[3 * m], ys = c->xyz[3 * m + 1], zs = c->xyz[3 * m + 2];
        double acc = 0.0;
        for (int64_t n = 0; n < N; ++n) {
            const double rho = c->rho[n];
            if (!(fabs(rho) > c->thr)) continue;
            double x, y, z;
            cell_centre(g, n, &x, &y, &z);
            acc += rho * oracle_prism_kernel(xs, ys, zs, x - hx, x + hx, y - hy, y + hy, z - hz,
                                             z + hz);
        }
        c->out[m] = c->kappa * acc;
    }
}

void oracle_prism_forward(const og_grid* g, double kappa, double thr, const double* rho,
                          int64_t m, const double* xyz, double* out, int nthreads) {
    pm_ctx c = {g, kappa, thr, rho, xyz, NULL, out, m};
    parallel_for(m, nthreads, prism_fwd_range, &c);
}

static void prism_adj_range(void* p, int64_t lo, int64_t hi) {
    const pm_ctx* c = (const pm_ctx*)p;
    const og_grid* g = c->g;
    const double hx = 0.5 * fabs(g->dx), hy = 0.5 * fabs(g->dy), hz = 0.5 * fabs(g->dz);
    for (int64_t n = lo; n < hi; ++n) {
        double x, y, z;
        cell_centre(g, n, &x, &y, &z);
        double acc = 0.0;
        for (int64_t m = 0; m < c->m; ++m)
            acc += c->r[m] * oracle_prism_kernel(c->xyz[3 * m], c->xyz[3 * m + 1],
                                                 c->xyz[3 * m + 2], x - hx, x + hx, y - hy,
                                                 y + hy, z - hz, z + hz);
        c->out[n] = c->kappa * acc;
    }
}

void oracle_prism_adjoint(const og_grid* g, double kappa, int64_t m, const double* xyz,
                          const double* r, double* out, int nthreads) {
    pm_ctx c = {g, kappa, 0.0, NULL, xyz, r, out, m};
    parallel_for((int64_t)g->nx * g->ny * g->nz, nthreads, prism_adj_range, &c);
}

/* ------------------------------------------- dense LA of the GN step */
/* oracle_mtm: zmath.cpp:260-262, MTM(a) = a^T a for a [rows][cols] row-major
 * (the reference's mat J[m][n]): out[i][j] = out[j][i] = sum_n a[n][i] a[n][j]
 * for j >= i, the sum in row order n = 0, 1, ... (the reference transposes a
 * first, then sums a^T[i][n] a^T[j][n] over n; one thread per column i there,
 * a parallel_for over i here). */
typedef struct { int64_t rows; int cols; const double* a; double* out; } mtm_ctx;

static void mtm_range(void* p, int64_t lo, int64_t hi) {
    const mtm_ctx* c = (const mtm_ctx*)p;
    for (int64_t i = lo; i < hi; ++i)
        for (int j = (int)i; j < c->cols; ++j) {
            double s = 0.0;
            for (int64_t n = 0; n < c->rows; ++n)
                s += c->a[n * c->cols + i] * c->a[n * c->cols + j];
            c->out[i * c->cols + j] = c->out[(int64_t)j * c->cols + i] = s;
        }
}

void oracle_mtm(int64_t rows, int cols, const double* a, double* out, int nthreads) {
    mtm_ctx c = {rows, cols, a, out};
    parallel_for(cols, nthreads, mtm_range, &c);
}

/* oracle_jtr: J^T r of TFM.cpp:160 (T(J)*r), out[i] = sum_m J[m][i] r[m]. */
void oracle_jtr(int64_t rows, int cols, const double* J, const double* r, double* out) {
    for (int i = 0; i < cols; ++i) {
        double s = 0.0;
        for (int64_t m = 0; m < rows; ++m) s += J[m * cols + i] * r[m];
        out[i] = s;
    }
}

/* oracle_inv: zmath.cpp:129-140, Gauss-Jordan inversion of a [n][n]: for each
 * column i the row with the largest |a[.][i]| at or below i (the first one on
 * ties) is swapped into place in a and in m (which starts as I); an exact zero
 * pivot returns the zero matrix (the reference's `return arr(l,l)`, here with
 * return value 0, else 1); row i of a and m is divided by the pivot, then
 * c = a[r][i] times row i is subtracted from every row r below; finally, for
 * i = 1..n-1, c = a[r][i] times row i is subtracted from every row r above.
 * The same operations in the same order. */
int oracle_inv(int n, const double* a_in, double* m) {
    const size_t nn = (size_t)n * n;
    double* a = (double*)malloc(nn * sizeof(double));
    double* tmp = (double*)malloc((size_t)n * sizeof(double));
    memcpy(a, a_in, nn * sizeof(double));
    memset(m, 0, nn * sizeof(double));
    for (int i = 0; i < n; ++i) m[(size_t)i * n + i] = 1.0;
    int ok = 1;
    for (int i = 0; i < n && ok; ++i) {
        int im = i;
        for (int r = i; r < n; ++r)
            if (fabs(a[(size_t)r * n + i]) > fabs(a[(size_t)im * n + i])) im = r;
        if (im != i) {
            memcpy(tmp, a + (size_t)i * n, n * sizeof(double));
            memcpy(a + (size_t)i * n, a + (size_t)im * n, n * sizeof(double));
            memcpy(a + (size_t)im * n, tmp, n * sizeof(double));
            memcpy(tmp, m + (size_t)i * n, n * sizeof(double));
            memcpy(m + (size_t)i * n, m + (size_t)im * n, n * sizeof(double));
            memcpy(m + (size_t)im * n, tmp, n * sizeof(double));
        }
        if (a[(size_t)i * n + i] == 0.0) {
            ok = 0;
            break;
        }
        const double f = a[(size_t)i * n + i];
        for (int j = 0; j < n; ++j) {
            a[(size_t)i * n + j] /= f;
            m[(size_t)i * n + j] /= f;
        }
        for (int r = i + 1; r < n; ++r) {
            const double c = a[(size_t)r * n + i];
            for (int j = 0; j < n; ++j) {
                a[(size_t)r * n + j] -= c * a[(size_t)i * n + j];
                m[(size_t)r * n + j] -= c * m[(size_t)i * n + j];
            }
        }
    }
    if (ok)
        for (int i = 1; i < n; ++i)
            for (int r = 0; r < i; ++r) {
                const double c = a[(size_t)r * n + i];
                for (int j = 0; j < n; ++j) {
                    m[(size_t)r * n + j] -= c * m[(size_t)i * n + j];
                    a[(size_t)r * n + j] -= c * a[(size_t)i * n + j];
                }
            }
    else
        memset(m, 0, nn * sizeof(double));
    free(a);
    free(tmp);
    return ok;
}

/* oracle_gn_steps: TFM.cpp:159-161 (LMA::fit): with M = diag(diag(JJ)) and
 * A = diag(1/sig^2), for each damping condition k, B[k] = b + inv(JJ +
 * dm[k][0] M + dm[k][1] A) (J^T r); jtr is J^T r. The reference evaluates
 * inv(...)*T(J)*r left to right, (inv * J^T) * r; here inv * (J^T r), the
 * same product reassociated (FP64 rounding apart), which saves the n x rows
 * intermediate. A singular system (exact zero pivot) gives B[k] = b, as the
 * reference's zero inverse does. One damping condition per thread, as the
 * reference runs them. */
typedef struct {
    int n; const double* JJ; const double* jtr; const double* sig; const double* dm;
    const double* b; double* B;
} gn_ctx;

static void gn_range(void* p, int64_t lo, int64_t hi) {
    const gn_ctx* c = (const gn_ctx*)p;
    const int n = c->n;
    const size_t nn = (size_t)n * n;
    double* K = (double*)malloc(nn * sizeof(double));
    double* Ki = (double*)malloc(nn * sizeof(double));
    for (int64_t k = lo; k < hi; ++k) {
        const double d0 = c->dm[2 * k], d1 = c->dm[2 * k + 1];
        memcpy(K, c->JJ, nn * sizeof(double));
        for (int i = 0; i < n; ++i)
            K[(size_t)i * n + i] = c->JJ[(size_t)i * n + i] + d0 * c->JJ[(size_t)i * n + i] +
                                   d1 * (1.0 / (c->sig[i] * c->sig[i]));
        oracle_inv(n, K, Ki);
        for (int i = 0; i < n; ++i) {
            double s = 0.0;
            for (int j = 0; j < n; ++j) s += Ki[(size_t)i * n + j] * c->jtr[j];
            c->B[k * n + i] = c->b[i] + s;
        }
    }
    free(K);
    free(Ki);
}

void oracle_gn_steps(int n, int nd, const double* JJ, const double* jtr, const double* sig,
                     const double* dm, const double* b, double* B, int nthreads) {
    gn_ctx c = {n, JJ, jtr, sig, dm, b, B};
    parallel_for(nd, nthreads, gn_range, &c);
}

/* oracle_inv_diag: TFM.cpp:150 (create) and :173 (annihilate), the diagonal
 * of inv(JJ + I d): out[i] = inv(...)[i][i]; all 0 when singular. */
void oracle_inv_diag(int n, const double* JJ, double d, double* out) {
    const size_t nn = (size_t)n * n;
    double* K = (double*)malloc(nn * sizeof(double));
    double* Ki = (double*)malloc(nn * sizeof(double));
    memcpy(K, JJ, nn * sizeof(double));
    for (int i = 0; i < n; ++i) K[(size_t)i * n + i] += d;
    oracle_inv(n, K, Ki);
    for (int i = 0; i < n; ++i) out[i] = Ki[(size_t)i * n + i];
    free(K);
    free(Ki);
}
