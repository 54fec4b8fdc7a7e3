// Producer of the gravity path on the device: the density perturbation drho
// that ZTM's forward model renders on the fine raster before every gravity
// call (TFM.cpp:83-89), so the rendered field never leaves the GPU.
//
// Per fine cell n (ii x jj x kk raster, coarse-index coordinates i, j, k):
//   m0 = interp2(sea, i, j) - 3.5           reference Moho (k units)
//   m1 = interp2(M, i, j)                   rendered Moho
//   r0 = 0 below m0, ro above, linear band of half-width 1/fv between
//   r1 = dr (interp3(v, i, j, k) - vm) below m1, ro above, band as written
//        in TFM.cpp:88 (r1 + t (r1 + ro), t in [0, 1))
//   drho = r1 - r0 if |r0 - r1| > .001 and depth(i, j, k) > 10 km, else 0
// with interp2/interp3 the unit-spacing bilinear/trilinear interpolation of
// zmath.cpp:225-249 and depth = R - sqrtf(|ECEF|^2) of the ENU point
// (TFM.cpp:56, SET.cpp:86-99). All FP64 in the reference's operation order,
// FMA contraction off, so drho equals the reference's value bit for bit; it
// is written as float, the narrowing the reference applies before its GPU
// call (TFM.cpp:91). One thread per fine cell (and, batched, per group of
// variants); the coarse grids are read through the cache.
#include "gfm_internal.h"

#include <cmath>

namespace ztm {

// Unit-spacing bilinear interpolation (zmath.cpp:225-236) in two parts: the
// weights and corner offsets of a point, which depend on the point only, and
// the evaluation on one grid, so that several grids sampled at the same point
// (the variants of a batch) share the first part. The operations and their
// order are those of one plain interpolation.
struct Interp2W {
    int64_t b;                  // offset of the (i, j) corner
    int di;                     // row stride
    double si, sj, ri, rj;      // fractions and 1 - fractions
    bool hi, hj;                // the i+1 / j+1 neighbours lie inside the grid
};

__device__ __forceinline__ Interp2W interp2_weights(double x, double y, int di, int dj) {
#pragma clang fp contract(off)
    const double ii = fmin(fmax(x / 1.0, 0.0), (double)(di - 1));
    const double jj = fmin(fmax(y / 1.0, 0.0), (double)(dj - 1));
    const int i = (int)ii, j = (int)jj;
    Interp2W w;
    w.b = i + (int64_t)j * di;
    w.di = di;
    w.si = ii - i;
    w.sj = jj - j;
    w.ri = 1 - w.si;
    w.rj = 1 - w.sj;
    w.hi = i + 1 < di;
    w.hj = j + 1 < dj;
    return w;
}

__device__ __forceinline__ double interp2_eval(const double* __restrict__ f, const Interp2W& w) {
#pragma clang fp contract(off)
    // a neighbour outside the grid has weight exactly 0: read as 0
    const double* q = f + w.b;
    const double c00 = q[0];
    const double c10 = w.hi ? q[1] : 0.0;
    const double c01 = w.hj ? q[w.di] : 0.0;
    const double c11 = (w.hi && w.hj) ? q[w.di + 1] : 0.0;
    const double c0 = c00 * w.ri + c10 * w.si, c1 = c01 * w.ri + c11 * w.si;
    return c0 * w.rj + c1 * w.sj;
}

__device__ __forceinline__ double interp2_unit(const double* __restrict__ f, double x, double y,
                                               int di, int dj) {
    return interp2_eval(f, interp2_weights(x, y, di, dj));
}

// Trilinear counterpart (zmath.cpp:238-249); the base corner is clamped to
// d-2 per axis, so all 8 corners lie inside the grid.
struct Interp3W {
    int64_t b, p;               // base corner offset, plane stride
    int di;
    double si, sj, sk, ri, rj, rk;
};

__device__ __forceinline__ Interp3W interp3_weights(double x, double y, double z, int di, int dj,
                                                   int dk) {
#pragma clang fp contract(off)
    const double ii = fmin(fmax(x / 1.0, 0.0), (double)(di - 2));
    const double jj = fmin(fmax(y / 1.0, 0.0), (double)(dj - 2));
    const double kk = fmin(fmax(z / 1.0, 0.0), (double)(dk - 2));
    const int i = (int)ii, j = (int)jj, k = (int)kk;
    Interp3W w;
    w.p = (int64_t)di * dj;
    w.b = i + (int64_t)j * di + k * w.p;
    w.di = di;
    w.si = ii - i;
    w.sj = jj - j;
    w.sk = kk - k;
    w.ri = 1 - w.si;
    w.rj = 1 - w.sj;
    w.rk = 1 - w.sk;
    return w;
}

__device__ __forceinline__ double interp3_eval(const double* __restrict__ f, const Interp3W& w) {
#pragma clang fp contract(off)
    const double* q = f + w.b;
    const int64_t di = w.di, p = w.p;
    const double c000 = q[0], c100 = q[1];
    const double c010 = q[di], c110 = q[di + 1];
    const double c001 = q[p], c101 = q[p + 1];
    const double c011 = q[p + di], c111 = q[p + di + 1];
    const double c00 = c000 * w.ri + c100 * w.si, c01 = c001 * w.ri + c101 * w.si;
    const double c10 = c010 * w.ri + c110 * w.si, c11 = c011 * w.ri + c111 * w.si;
    const double c0 = c00 * w.rj + c10 * w.sj, c1 = c01 * w.rj + c11 * w.sj;
    return c0 * w.rk + c1 * w.sk;
}

__device__ __forceinline__ double interp3_unit(const double* __restrict__ f, double x, double y,
                                               double z, int di, int dj, int dk) {
    return interp3_eval(f, interp3_weights(x, y, z, di, dj, dk));
}

__device__ __forceinline__ double depth_km(const TfmFrame& fr, double ci, double cj, double ck) {
#pragma clang fp contract(off)
    const double x = ci * fr.h - fr.off[0], y = cj * fr.h - fr.off[1], z = ck * fr.h - fr.off[2];
    const double xi = fr.m[0] * x + fr.m[1] * y + fr.m[2] * z + fr.ref[0];
    const double yi = fr.m[3] * x + fr.m[4] * y + fr.m[5] * z + fr.ref[1];
    const double zi = fr.m[6] * x + fr.m[7] * y + fr.m[8] * z + fr.ref[2];
    const double r = sqrtf((float)(xi * xi + yi * yi + zi * zi));
    return 6371.0 - r;
}

// The per-cell terms that do not depend on the model parameters: the
// coarse-index coordinates, the reference Moho m0 and the depth filter.
struct DrhoCell {
    double i, j, k, li, m0;
    bool deep;
};

__device__ __forceinline__ DrhoCell drho_cell(int64_t n, int di, int dj, int fh, int fv,
                                              const TfmFrame& fr,
                                              const double* __restrict__ sea) {
#pragma clang fp contract(off)
    const int ii = (di - 1) * fh + 1, jj = (dj - 1) * fh + 1;
    DrhoCell c;
    c.i = (double)(n % ii) / (1. * fh);
    c.j = (double)((n / ii) % jj) / (1. * fh);
    c.k = (double)(n / ((int64_t)ii * jj)) / (1. * fv);
    c.li = 1. / fv;
    c.m0 = interp2_unit(sea, c.i, c.j, di, dj) - 3.5;
    c.deep = depth_km(fr, c.i, c.j, c.k) > 10;
    return c;
}

// drho of one cell from its rendered Moho m1 and velocity v (TFM.cpp:86-89).
__device__ __forceinline__ float drho_from(const DrhoCell& c, double m1, double v, double vm,
                                           double ro, double dr) {
#pragma clang fp contract(off)
    const double k = c.k, li = c.li, m0 = c.m0;
    double r0 = 0;
    double r1 = dr * (v - vm);
    if (k > m0 + li) {
        r0 = ro;
    } else if (fabs(k - m0) < li) {
        r0 = r0 + (k - m0 + li) * (r0 + ro) / (2 * li);
    }
    if (k > m1 + li) {
        r1 = ro;
    } else if (fabs(k - m1) < li) {
        r1 = r1 + (k - m1 + li) * (r1 + ro) / (2 * li);
    }
    float out = 0.0f;
    if (fabs(r0 - r1) > .001 && c.deep) out = (float)(r1 - r0);
    return out;
}

__device__ __forceinline__ float drho_at(int64_t n, int di, int dj, int dk, int fh, int fv,
                                         const TfmFrame& fr, const double* __restrict__ sea,
                                         const double* __restrict__ moho,
                                         const double* __restrict__ vel, double vm, double ro,
                                         double dr) {
    const DrhoCell c = drho_cell(n, di, dj, fh, fv, fr, sea);
    const double m1 = interp2_unit(moho, c.i, c.j, di, dj);
    const double v = interp3_unit(vel, c.i, c.j, c.k, di, dj, dk);
    return drho_from(c, m1, v, vm, ro, dr);
}

__global__ __launch_bounds__(256) void tfm_drho_kernel(int di, int dj, int dk, int fh, int fv,
                                                       TfmFrame fr,
                                                       const double* __restrict__ sea,
                                                       const double* __restrict__ moho,
                                                       const double* __restrict__ vel, double vm,
                                                       double ro, double dr,
                                                       float* __restrict__ drho, int64_t N) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    drho[n] = drho_at(n, di, dj, dk, fh, fv, fr, sea, moho, vel, vm, ro, dr);
}

// Variants kDrhoVariants * blockIdx.y onwards, each with its own Moho and
// velocity grids (strides in elements, 0 = shared) and (vm, ro, dr) =
// params[3b .. 3b+2]; drho [B][N]. The parameter-free terms of a cell, the
// interpolation weights and a shared grid's interpolation are formed once per
// thread for all its variants (the same operations, so each variant equals
// drho_at).
constexpr int kDrhoVariants = 16;

__global__ __launch_bounds__(256) void tfm_drho_batched_kernel(
    int di, int dj, int dk, int fh, int fv, TfmFrame fr, const double* __restrict__ sea,
    const double* __restrict__ moho, int64_t moho_stride, const double* __restrict__ vel,
    int64_t vel_stride, const double* __restrict__ params, float* __restrict__ drho, int64_t N,
    int B) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const DrhoCell c = drho_cell(n, di, dj, fh, fv, fr, sea);
    const Interp2W w2 = interp2_weights(c.i, c.j, di, dj);
    const Interp3W w3 = interp3_weights(c.i, c.j, c.k, di, dj, dk);
    const double m1s = moho_stride == 0 ? interp2_eval(moho, w2) : 0.0;
    const double vs = vel_stride == 0 ? interp3_eval(vel, w3) : 0.0;
    const int b0 = blockIdx.y * kDrhoVariants;
    const int b1 = min(b0 + kDrhoVariants, B);
    for (int b = b0; b < b1; ++b) {
        const double m1 = moho_stride == 0 ? m1s : interp2_eval(moho + b * moho_stride, w2);
        const double v = vel_stride == 0 ? vs : interp3_eval(vel + b * vel_stride, w3);
        drho[(size_t)b * N + n] =
            drho_from(c, m1, v, params[3 * b], params[3 * b + 1], params[3 * b + 2]);
    }
}

// Derivatives of drho with respect to (vm, ro, dr) at one parameter point,
// out [3][N] = d drho / d(vm, ro, dr) (float). Within each piece of
// TFM.cpp:86-88 drho is linear in the three parameters:
//   r0 = ro above m0 + li, t0 ro in the band (t0 = (k - m0 + li) / 2li), else 0;
//   r1 = ro above m1 + li, dr u (1 + t1) + t1 ro in the band, else dr u
//   (u = v - vm). The derivative is that of r1 - r0 where the cell is kept:
// the producer's own filter (|r0 - r1| > .001 and depth > 10 km) and GF's skip
// of |float drho| <= .001 (GFM_GPU.cu:19), both decided on the value at this
// point; 0 elsewhere. A finite-difference column differs from it only by
// cells whose piece or filter changes within the step.
__global__ __launch_bounds__(256) void tfm_drho_jvp_kernel(
    int di, int dj, int dk, int fh, int fv, TfmFrame fr, const double* __restrict__ sea,
    const double* __restrict__ moho, const double* __restrict__ vel, double vm, double ro,
    double dr, float* __restrict__ out, int64_t N) {
#pragma clang fp contract(off)
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const DrhoCell c = drho_cell(n, di, dj, fh, fv, fr, sea);
    const double m1 = interp2_unit(moho, c.i, c.j, di, dj);
    const double v = interp3_unit(vel, c.i, c.j, c.k, di, dj, dk);
    const float base = drho_from(c, m1, v, vm, ro, dr);
    double d_vm = 0.0, d_ro = 0.0, d_dr = 0.0;
    if (fabsf(base) > .001f) {
        const double k = c.k, li = c.li, m0 = c.m0;
        // r0
        if (k > m0 + li) d_ro -= 1.0;
        else if (fabs(k - m0) < li) d_ro -= (k - m0 + li) / (2 * li);
        // r1
        const double u = v - vm;
        if (k > m1 + li) {
            d_ro += 1.0;
        } else if (fabs(k - m1) < li) {
            const double t1 = (k - m1 + li) / (2 * li);
            d_vm += -dr * (1 + t1);
            d_ro += t1;
            d_dr += u * (1 + t1);
        } else {
            d_vm += -dr;
            d_dr += u;
        }
    }
    out[n] = (float)d_vm;
    out[N + n] = (float)d_ro;
    out[2 * N + n] = (float)d_dr;
}

// Sensitivity of one cell's drho to its rendered Moho m1 and velocity v at
// the parameter point: only r1 depends on them (TFM.cpp:86-88), so
//   above the band (k > m1 + li):  0, 0;
//   in the band, r1 = dr u + t1 (dr u + ro), t1 = (k - m1 + li) / 2li:
//     d/dm1 = -(dr u + ro) / 2li,  d/dv = dr (1 + t1);
//   below it:                      0, dr;
// and 0 on the cells the producer or GF's skip drops at this point
// (|float drho| <= .001, as tfm_drho_jvp_kernel decides them).
struct DrhoSens {
    double cm, cv;
};

__device__ __forceinline__ DrhoSens drho_sens(const DrhoCell& c, double m1, double v, double vm,
                                              double ro, double dr) {
#pragma clang fp contract(off)
    DrhoSens d{0.0, 0.0};
    if (!(fabsf(drho_from(c, m1, v, vm, ro, dr)) > .001f)) return d;
    const double k = c.k, li = c.li;
    if (k > m1 + li) return d;
    if (fabs(k - m1) < li) {
        const double t1 = (k - m1 + li) / (2 * li);
        d.cm = -(dr * (v - vm) + ro) / (2 * li);
        d.cv = dr * (1 + t1);
    } else {
        d.cv = dr;
    }
    return d;
}

// Tangent fields for directions b = kDrhoVariants * blockIdx.y onwards:
// out[b][n] = d drho[n] / d(M, V) . (dmoho[b], dvel[b]), the tangent grids
// interpolated with the cell's own weights (interp2 / interp3 are linear in
// the grid values). dmoho [nv][di*dj] and dvel [nv][di*dj*dk] may be null
// (no tangent on that grid).
__global__ __launch_bounds__(256) void tfm_drho_jvp_grids_kernel(
    int di, int dj, int dk, int fh, int fv, TfmFrame fr, const double* __restrict__ sea,
    const double* __restrict__ moho, const double* __restrict__ vel, double vm, double ro,
    double dr, int nv, const double* __restrict__ dmoho, const double* __restrict__ dvel,
    float* __restrict__ out, int64_t N) {
#pragma clang fp contract(off)
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const DrhoCell c = drho_cell(n, di, dj, fh, fv, fr, sea);
    const Interp2W w2 = interp2_weights(c.i, c.j, di, dj);
    const Interp3W w3 = interp3_weights(c.i, c.j, c.k, di, dj, dk);
    const DrhoSens d = drho_sens(c, interp2_eval(moho, w2), interp3_eval(vel, w3), vm, ro, dr);
    const int64_t M = (int64_t)di * dj;
    const int b0 = blockIdx.y * kDrhoVariants;
    const int b1 = min(b0 + kDrhoVariants, nv);
    for (int b = b0; b < b1; ++b) {
        const double dm = (dmoho && d.cm != 0.0) ? interp2_eval(dmoho + b * M, w2) : 0.0;
        const double dv = (dvel && d.cv != 0.0) ? interp3_eval(dvel + b * M * dk, w3) : 0.0;
        out[(size_t)b * N + n] = (float)(d.cm * dm + d.cv * dv);
    }
}

// Transpose, first step: per fine column (fi, fj), with a = G^T r on the
// fine raster, wm[fi + ii fj] = sum over the column's cells of a cm (the
// Moho enters per column) and wv[n] = a[n] cv per cell.
__global__ __launch_bounds__(256) void tfm_grid_vjp_cells_kernel(
    int di, int dj, int dk, int fh, int fv, TfmFrame fr, const double* __restrict__ sea,
    const double* __restrict__ moho, const double* __restrict__ vel, double vm, double ro,
    double dr, const float* __restrict__ a, double* __restrict__ wm, double* __restrict__ wv) {
#pragma clang fp contract(off)
    const int ii = (di - 1) * fh + 1, jj = (dj - 1) * fh + 1, kk = (dk - 1) * fv + 1;
    const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= (int64_t)ii * jj) return;
    double sm = 0.0;
    for (int k = 0; k < kk; ++k) {
        const int64_t n = col + (int64_t)k * ii * jj;
        const DrhoCell c = drho_cell(n, di, dj, fh, fv, fr, sea);
        const DrhoSens d = drho_sens(c, interp2_unit(moho, c.i, c.j, di, dj),
                                     interp3_unit(vel, c.i, c.j, c.k, di, dj, dk), vm, ro, dr);
        const double an = (double)a[n];
        sm += an * d.cm;
        wv[n] = an * d.cv;
    }
    wm[col] = sm;
}

// Weight of coarse node c along one axis for a fine point whose stencil has
// base corner b and fraction f (1 - f = r): r at b, f at b + 1, else 0.
__device__ __forceinline__ double axis_weight(int c, int b, double f, double r, bool has_next) {
    return c == b ? r : (c == b + 1 && has_next) ? f : 0.0;
}

// Transpose, second step, Moho: gm[c] = sum over the fine columns whose
// bilinear stencil holds node c of wm * weight (fine points within one
// coarse cell of the node), in a fixed order.
__global__ __launch_bounds__(256) void tfm_grid_vjp_moho_kernel(int di, int dj, int fh,
                                                                const double* __restrict__ wm,
                                                                double* __restrict__ gm) {
    const int ii = (di - 1) * fh + 1, jj = (dj - 1) * fh + 1;
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= di * dj) return;
    const int ci = node % di, cj = node / di;
    const int fi0 = max((ci - 1) * fh, 0), fi1 = min((ci + 1) * fh, ii - 1);
    const int fj0 = max((cj - 1) * fh, 0), fj1 = min((cj + 1) * fh, jj - 1);
    double acc = 0.0;
    for (int fj = fj0; fj <= fj1; ++fj)
        for (int fi = fi0; fi <= fi1; ++fi) {
            const Interp2W w = interp2_weights((double)fi / (1. * fh), (double)fj / (1. * fh), di,
                                               dj);
            const int bi = (int)(w.b % di), bj = (int)(w.b / di);
            const double wx = axis_weight(ci, bi, w.si, w.ri, w.hi);
            const double wy = axis_weight(cj, bj, w.sj, w.rj, w.hj);
            if (wx != 0.0 && wy != 0.0) acc += wm[fi + (int64_t)ii * fj] * (wx * wy);
        }
    gm[node] = acc;
}

// Velocity: gv[c] = sum over the fine cells whose trilinear stencil holds
// node c of wv * weight (the stencil's base is clamped to d - 2 per axis, so
// the last node takes fraction 1 from the last fine points).
__global__ __launch_bounds__(256) void tfm_grid_vjp_vel_kernel(int di, int dj, int dk, int fh,
                                                               int fv,
                                                               const double* __restrict__ wv,
                                                               double* __restrict__ gv) {
    const int ii = (di - 1) * fh + 1, jj = (dj - 1) * fh + 1, kk = (dk - 1) * fv + 1;
    const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= (int64_t)di * dj * dk) return;
    const int ci = (int)(node % di), cj = (int)((node / di) % dj), ck = (int)(node / ((int64_t)di * dj));
    const int fi0 = max((ci - 1) * fh, 0), fi1 = min((ci + 1) * fh, ii - 1);
    const int fj0 = max((cj - 1) * fh, 0), fj1 = min((cj + 1) * fh, jj - 1);
    const int fk0 = max((ck - 1) * fv, 0), fk1 = min((ck + 1) * fv, kk - 1);
    double acc = 0.0;
    for (int fk = fk0; fk <= fk1; ++fk)
        for (int fj = fj0; fj <= fj1; ++fj)
            for (int fi = fi0; fi <= fi1; ++fi) {
                const Interp3W w = interp3_weights((double)fi / (1. * fh), (double)fj / (1. * fh),
                                                   (double)fk / (1. * fv), di, dj, dk);
                const int bi = (int)(w.b % di), bj = (int)((w.b / di) % dj);
                const int bk = (int)(w.b / w.p);
                const double wx = axis_weight(ci, bi, w.si, w.ri, true);
                const double wy = axis_weight(cj, bj, w.sj, w.rj, true);
                const double wz = axis_weight(ck, bk, w.sk, w.rk, true);
                if (wx != 0.0 && wy != 0.0 && wz != 0.0)
                    acc += wv[fi + (int64_t)ii * (fj + (int64_t)jj * fk)] * (wx * wy * wz);
            }
    gv[node] = acc;
}

int tfm_drho_jvp_grids_launch(int di, int dj, int dk, int fh, int fv, const double* rc,
                              const double* sea, const double* moho, const double* vel,
                              double vm, double ro, double dr, int nv, const double* dmoho,
                              const double* dvel, float* out, hipStream_t stream) {
    const int64_t N = ((int64_t)(di - 1) * fh + 1) * ((int64_t)(dj - 1) * fh + 1) *
                      ((int64_t)(dk - 1) * fv + 1);
    if (N == 0 || nv <= 0) return ZTM_OK;
    const TfmFrame fr = tfm_frame(rc);
    const unsigned gy = (unsigned)((nv + kDrhoVariants - 1) / kDrhoVariants);
    hipLaunchKernelGGL(tfm_drho_jvp_grids_kernel, dim3((unsigned)((N + 255) / 256), gy),
                       dim3(256), 0, stream, di, dj, dk, fh, fv, fr, sea, moho, vel, vm, ro, dr,
                       nv, dmoho, dvel, out, N);
    ZTM_HIP(hipGetLastError());
    return ZTM_OK;
}

size_t tfm_grid_vjp_workspace(int di, int dj, int dk, int fh, int fv) {
    const int64_t cols = ((int64_t)(di - 1) * fh + 1) * ((int64_t)(dj - 1) * fh + 1);
    const int64_t N = cols * ((int64_t)(dk - 1) * fv + 1);
    return sizeof(double) * (size_t)(cols + N);
}

int tfm_grid_vjp_launch(int di, int dj, int dk, int fh, int fv, const double* rc,
                        const double* sea, const double* moho, const double* vel, double vm,
                        double ro, double dr, const float* a, double* ws, double* gmoho,
                        double* gvel, hipStream_t stream) {
    const int64_t cols = ((int64_t)(di - 1) * fh + 1) * ((int64_t)(dj - 1) * fh + 1);
    double* wm = ws;
    double* wv = ws + cols;
    const TfmFrame fr = tfm_frame(rc);
    hipLaunchKernelGGL(tfm_grid_vjp_cells_kernel, dim3((unsigned)((cols + 255) / 256)),
                       dim3(256), 0, stream, di, dj, dk, fh, fv, fr, sea, moho, vel, vm, ro, dr,
                       a, wm, wv);
    ZTM_HIP(hipGetLastError());
    if (gmoho) {
        hipLaunchKernelGGL(tfm_grid_vjp_moho_kernel, dim3((unsigned)((di * dj + 255) / 256)),
                           dim3(256), 0, stream, di, dj, fh, wm, gmoho);
        ZTM_HIP(hipGetLastError());
    }
    if (gvel) {
        const int64_t nodes = (int64_t)di * dj * dk;
        hipLaunchKernelGGL(tfm_grid_vjp_vel_kernel, dim3((unsigned)((nodes + 255) / 256)),
                           dim3(256), 0, stream, di, dj, dk, fh, fv, wv, gvel);
        ZTM_HIP(hipGetLastError());
    }
    return ZTM_OK;
}

int tfm_drho_jvp_launch(int di, int dj, int dk, int fh, int fv, const double* rc,
                        const double* sea, const double* moho, const double* vel, double vm,
                        double ro, double dr, float* out, hipStream_t stream) {
    const int64_t N = ((int64_t)(di - 1) * fh + 1) * ((int64_t)(dj - 1) * fh + 1) *
                      ((int64_t)(dk - 1) * fv + 1);
    if (N == 0) return ZTM_OK;
    const TfmFrame fr = tfm_frame(rc);
    hipLaunchKernelGGL(tfm_drho_jvp_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0,
                       stream, di, dj, dk, fh, fv, fr, sea, moho, vel, vm, ro, dr, out, N);
    ZTM_HIP(hipGetLastError());
    return ZTM_OK;
}

// r_int[i] = r[out_map[i]] (float) for the m prepared stations: a residual on
// the di x dj station raster in the order of the reference plan's stations.
__global__ void tfm_gather_kernel(const double* __restrict__ r, const int* __restrict__ out_map,
                                  int m, float* __restrict__ r_int) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) r_int[i] = (float)r[out_map[i]];
}

// Dot products of a[N] with the three fields f[3][N], FP64, in a fixed
// order: block b sums the elements b, b + kDotBlocks*256, ... per thread,
// then through LDS (tree); tfm_dot_final adds the kDotBlocks partials in
// block order.
constexpr int kDotBlocks = 512;

__global__ __launch_bounds__(256) void tfm_dot3_kernel(const float* __restrict__ a,
                                                       const float* __restrict__ f, int64_t N,
                                                       double* __restrict__ part) {
    __shared__ double red[3][256];
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x; n < N;
         n += (int64_t)kDotBlocks * 256) {
        const double an = a[n];
        s0 += an * (double)f[n];
        s1 += an * (double)f[N + n];
        s2 += an * (double)f[2 * N + n];
    }
    red[0][threadIdx.x] = s0;
    red[1][threadIdx.x] = s1;
    red[2][threadIdx.x] = s2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w)
#pragma unroll
            for (int q = 0; q < 3; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 3) part[threadIdx.x * kDotBlocks + blockIdx.x] = red[threadIdx.x][0];
}

__global__ void tfm_dot_final_kernel(const double* __restrict__ part, double* __restrict__ out) {
    const int q = threadIdx.x;
    if (q >= 3) return;
    double s = 0.0;
    for (int b = 0; b < kDotBlocks; ++b) s += part[q * kDotBlocks + b];
    out[q] = s;
}

size_t tfm_vjp_workspace(int64_t N, int m) {
    // a[N] float, r_int[m] float, block partials [3][kDotBlocks] double
    return (((size_t)N + (size_t)m) * sizeof(float) + 255) / 256 * 256 +
           (size_t)3 * kDotBlocks * sizeof(double);
}

int tfm_gather_launch(const double* r, const int* out_map, int m, float* r_int,
                      hipStream_t stream) {
    if (m == 0) return ZTM_OK;
    hipLaunchKernelGGL(tfm_gather_kernel, dim3((m + 255) / 256), dim3(256), 0, stream, r,
                       out_map, m, r_int);
    ZTM_HIP(hipGetLastError());
    return ZTM_OK;
}

int tfm_dot3_launch(const float* a, const float* f, int64_t N, double* part, double* out,
                    hipStream_t stream) {
    hipLaunchKernelGGL(tfm_dot3_kernel, dim3(kDotBlocks), dim3(256), 0, stream, a, f, N, part);
    hipLaunchKernelGGL(tfm_dot_final_kernel, dim3(1), dim3(64), 0, stream, part, out);
    ZTM_HIP(hipGetLastError());
    return ZTM_OK;
}

// g[b][*] = 0, then g[b][out_map[i]] = Y[b][i] for the m prepared stations.
__global__ void tfm_fill_kernel(double* __restrict__ g, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) g[t] = 0.0;
}

__global__ void tfm_scatter_kernel(const double* __restrict__ Y, const int* __restrict__ out_map,
                                   int m, int M, int B, double* __restrict__ g) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)m * B) return;
    const int b = (int)(t / m), i = (int)(t % m);
    g[(size_t)b * M + out_map[i]] = Y[t];
}

TfmFrame tfm_frame(const double* rc) {
    // SET.cpp:86-96: ENU -> ECEF rotation and the reference point, from RC
    // (TFM.cpp:211): lat_ref, lon_ref, d_ref, ENU offsets, cell edge h.
    const double rad = 57.295, R = 6371;
    const double lat_ref = rc[0], lon_ref = rc[1], d_ref = rc[2];
    TfmFrame f;
    f.ref[0] = (R - d_ref) * std::cos(lon_ref / rad) * std::cos(lat_ref / rad);
    f.ref[1] = (R - d_ref) * std::sin(lon_ref / rad) * std::cos(lat_ref / rad);
    f.ref[2] = (R - d_ref) * std::sin(lat_ref / rad);
    const double phi = lat_ref / rad, theta = lon_ref / rad;
    const double m[9] = {-std::sin(theta), -std::sin(phi) * std::cos(theta), std::cos(phi) * std::cos(theta),
                         std::cos(theta),  -std::sin(phi) * std::sin(theta), std::cos(phi) * std::sin(theta),
                         0,                std::cos(phi),                    std::sin(phi)};
    for (int q = 0; q < 9; ++q) f.m[q] = m[q];
    f.off[0] = rc[3];
    f.off[1] = rc[4];
    f.off[2] = rc[5];
    f.h = rc[6];
    return f;
}

int tfm_drho_launch(int di, int dj, int dk, int fh, int fv, const double* rc, const double* sea,
                    const double* moho, const double* vel, double vm, double ro, double dr,
                    float* drho, hipStream_t stream) {
    const int64_t N = ((int64_t)(di - 1) * fh + 1) * ((int64_t)(dj - 1) * fh + 1) *
                      ((int64_t)(dk - 1) * fv + 1);
    if (N == 0) return ZTM_OK;
    const TfmFrame fr = tfm_frame(rc);
    ProfScope prof("tfm_drho", stream);
    hipLaunchKernelGGL(tfm_drho_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, stream,
                       di, dj, dk, fh, fv, fr, sea, moho, vel, vm, ro, dr, drho, N);
    ZTM_HIP(hipGetLastError());
    return ZTM_OK;
}

int tfm_drho_batched_launch(int di, int dj, int dk, int fh, int fv, int B, const double* rc,
                            const double* sea, const double* moho, int64_t moho_stride,
                            const double* vel, int64_t vel_stride, const double* params,
                            float* drho, hipStream_t stream) {
    const int64_t N = ((int64_t)(di - 1) * fh + 1) * ((int64_t)(dj - 1) * fh + 1) *
                      ((int64_t)(dk - 1) * fv + 1);
    if (N == 0 || B <= 0) return ZTM_OK;
    const TfmFrame fr = tfm_frame(rc);
    ProfScope prof("tfm_drho", stream);
    const unsigned gy = (unsigned)((B + kDrhoVariants - 1) / kDrhoVariants);
    hipLaunchKernelGGL(tfm_drho_batched_kernel, dim3((unsigned)((N + 255) / 256), gy),
                       dim3(256), 0, stream, di, dj, dk, fh, fv, fr, sea, moho, moho_stride, vel,
                       vel_stride, params, drho, N, B);
    ZTM_HIP(hipGetLastError());
    return ZTM_OK;
}

int tfm_scatter_launch(const double* Y, const int* out_map, int m, int M, int B, double* g,
                       hipStream_t stream) {
    const int64_t n = (int64_t)M * B;
    if (n == 0) return ZTM_OK;
    hipLaunchKernelGGL(tfm_fill_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream,
                       g, n);
    const int64_t ns = (int64_t)m * B;
    if (ns > 0)
        hipLaunchKernelGGL(tfm_scatter_kernel, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0,
                           stream, Y, out_map, m, M, B, g);
    ZTM_HIP(hipGetLastError());
    return ZTM_OK;
}

}  // namespace ztm
