// Device-side pair terms shared by the point-mass and prism kernels.
//
// Units: lengths in L = |dx|; a station is (X, Y, Z) with cell (i,j,k) at
// (i, j*sy, k*sz). The main FP32 kernels accumulate, per station-cell pair,
//     w * dz * K32,  K32 = v^3 (point mass)  or  v^3 * (1 + v^2 (q v^2 - Qz))
// (point mass plus the quadrupole term of a non-cubic box), v = 1/r.
#pragma once

#include "gfm_internal.h"

namespace ztm {

__device__ __forceinline__ float rsq_f(float x) { return __builtin_amdgcn_rsqf(x); }

// Row term of the FP32 kernels: R = dy^2 + dz^2, clamped below at kRmin2F.
// A station closer than 0.5 cell widths to a row's axis would otherwise put a
// near-singular term into that row's FP32 sum and swamp the rest of the row;
// clamped rows are corrected in FP64 afterwards (row_clamped, corr_pair).
__device__ __forceinline__ float row_R32(float dy, float dz) {
    return fmaxf(fmaf(dy, dy, dz * dz), kRmin2F);
}
__device__ __forceinline__ bool row_clamped(float dy, float dz) {
    return fmaf(dy, dy, dz * dz) < kRmin2F;
}

// Far-field kernel of one pair in FP32, exactly as the main kernels evaluate
// it. a = X - i, R = row_R32(dy, dz), qrow = q2y dy^2 + q2z dz^2.
template <bool QUAD>
__device__ __forceinline__ float pair_k32(float a, float R, float qrow, const QuadArgs& q) {
    const float v = rsq_f(fmaf(a, a, R));
    const float v2 = v * v;
    const float v3 = v2 * v;
    if constexpr (QUAD) {
        const float qq = fmaf(a * q.q2x, a, qrow);
        const float s = fmaf(qq, v2, -q.qz);
        const float f = fmaf(s, v2, 1.0f);
        return v3 * f;
    } else {
        return v3;
    }
}

// ln(v + r) without cancellation when v < 0: v + r = (r^2 - v^2)/(r - v).
__device__ __forceinline__ double log_vr(double v, double r, double rest2) {
    return v >= 0.0 ? log(v + r) : log(rest2 / (r - v));
}

// Antiderivative of the prism g_z kernel (Nagy, Papp & Benedek 2000) in
// station-relative coordinates; terms with an exactly zero prefactor are 0.
__device__ __forceinline__ double prism_F(double u, double v, double w) {
    const double u2 = u * u, v2 = v * v, w2 = w * w;
    const double r = sqrt(u2 + v2 + w2);
    double f = 0.0;
    if (u != 0.0) f += u * log_vr(v, r, u2 + w2);
    if (v != 0.0) f += v * log_vr(u, r, v2 + w2);
    if (w != 0.0) f -= w * atan(u * v / (w * r));
    return f;
}

// Exact integral of (Z - z)/r^3 over cell (i,j,k), in units of L, FP64.
__device__ __forceinline__ double prism_cell(double X, double Y, double Z, int i, int j, int k,
                                             double sy, double sz) {
    const double hy = 0.5 * fabs(sy), hz = 0.5 * fabs(sz);
    const double u0 = (double)i - 0.5 - X, u1 = (double)i + 0.5 - X;
    const double v0 = (double)j * sy - hy - Y, v1 = (double)j * sy + hy - Y;
    const double w0 = (double)k * sz - hz - Z, w1 = (double)k * sz + hz - Z;
    double s = 0.0;
    s += prism_F(u0, v0, w0) - prism_F(u0, v0, w1) - prism_F(u0, v1, w0) + prism_F(u0, v1, w1);
    s -= prism_F(u1, v0, w0) - prism_F(u1, v0, w1) - prism_F(u1, v1, w0) + prism_F(u1, v1, w1);
    return -s;
}

// Mean of the point-mass kernel dz/r^3 over cell (i,j,k) by 3x3x3
// Gauss-Legendre quadrature, FP32 (accurate to ~2e-6 of the cell's 1/r^2
// scale at 3 cells, ~4e-9 beyond 8).
__device__ __forceinline__ float gl3_cell(float X, float Y, float Z, int i, int j, int k,
                                          float sy, float sz) {
    const float n = 0.38729833462074170f;   // sqrt(3/5)/2: nodes at +-n, 0 on a unit edge
    const float wgt[3] = {5.0f / 18.0f, 8.0f / 18.0f, 5.0f / 18.0f};
    const float off[3] = {-n, 0.0f, n};
    const float ay = fabsf(sy), az = fabsf(sz);
    float acc = 0.0f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float dz = Z - ((float)k * sz + off[c] * az);
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const float dy = Y - ((float)j * sy + off[b] * ay);
            const float R = fmaf(dy, dy, dz * dz);
            float row = 0.0f;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const float dx = X - ((float)i + off[a]);
                const float v = rsq_f(fmaf(dx, dx, R));
                row = fmaf(wgt[a], v * v * v, row);
            }
            acc = fmaf(wgt[c] * wgt[b] * dz, row, acc);
        }
    }
    return acc;
}

// FP64 far-field kernel of one pair, as the FP64 main kernels evaluate it
// (the same formula as pair_k32, in double, without the row clamp).
template <bool QUAD>
__device__ __forceinline__ double pair_k64(double a, double R, double qrow, const QuadArgs& q) {
    const double v = rsqrt(fma(a, a, R));
    const double v2 = v * v;
    const double v3 = v2 * v;
    if constexpr (QUAD) {
        const double qq = fma(a * (double)q.q2x, a, qrow);
        const double s = fma(qq, v2, -(double)q.qz);
        return v3 * fma(s, v2, 1.0);
    } else {
        return v3;
    }
}

// gl3_cell in FP64 (the mid zone of ZTM_FP64_ZONED).
__device__ __forceinline__ double gl3_cell64(double X, double Y, double Z, int i, int j, int k,
                                             double sy, double sz) {
    const double n = 0.38729833462074170;   // sqrt(3/5)/2
    const double wgt[3] = {5.0 / 18.0, 8.0 / 18.0, 5.0 / 18.0};
    const double off[3] = {-n, 0.0, n};
    const double ay = fabs(sy), az = fabs(sz);
    double acc = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double dz = Z - ((double)k * sz + off[c] * az);
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const double dy = Y - ((double)j * sy + off[b] * ay);
            const double R = fma(dy, dy, dz * dz);
            double row = 0.0;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const double dx = X - ((double)i + off[a]);
                const double v = rsqrt(fma(dx, dx, R));
                row = fma(wgt[a], v * v * v, row);
            }
            acc = fma(wgt[c] * wgt[b] * dz, row, acc);
        }
    }
    return acc;
}

// Correction of one pair for ZTM_FP64_ZONED: w * (better - far field), all in
// FP64; better = the FP64 prism closed form in the near box, FP64
// Gauss-Legendre in the rest of the correction box (callers visit only
// those cells).
template <bool QUAD>
__device__ __forceinline__ double corr_pair64(double w, const double4& pd, const int4& c, int i,
                                              int j, int k, const NearArgs& a) {
    const double ad = pd.x - (double)i;
    const double dyd = pd.y - (double)j * a.syd;
    const double dzd = pd.z - (double)k * a.szd;
    const double R = fma(dyd, dyd, fma(dzd, dzd, kEpsD));
    const double qrow = QUAD ? fma((double)a.q.q2y, dyd * dyd, (double)a.q.q2z * (dzd * dzd)) : 0.0;
    const double far = dzd * pair_k64<QUAD>(ad, R, qrow, a.q);
    const bool in_near = abs(i - c.x) <= a.nrx && abs(j - c.y) <= a.nry && abs(k - c.z) <= a.nrz;
    const double better = in_near ? prism_cell(pd.x, pd.y, pd.z, i, j, k, a.syd, a.szd) / a.vol
                                  : gl3_cell64(pd.x, pd.y, pd.z, i, j, k, a.syd, a.szd);
    return w * (better - far);
}

// Correction of one station-cell pair for the FP32 path: the better value of
// the pair minus its FP32 far-field term as the main kernels evaluate it (same
// rounded coordinates, same row clamp, same v_rsq_f32), so adding it replaces
// the pair's FP32 value. Terms are  w * dz * K  in cell units (K = 1/r^3 for a
// point mass). The better value depends on where the cell lies relative to
// the station's nearest cell (per-axis index radii, NearArgs):
//   near box (nr*):                FP64 point mass, or FP64 prism closed form;
//   correction box (r*), prism:    FP32 3x3x3 Gauss-Legendre mean of dz/r^3;
//   beyond (cells of a clamped row): FP64 far-field kernel (point mass, plus
//                                 the quadrupole term for non-cubic prisms).
template <bool QUAD>
__device__ __forceinline__ double corr_pair(int mode, float w, const float4& pf,
                                            const double4& pd, const int4& c, int i, int j,
                                            int k, const NearArgs& a) {
    const float dyf = pf.y - (float)j * a.syf;
    const float dzf = pf.z - (float)k * a.szf;
    const float af = pf.x - (float)i;
    const float R = row_R32(dyf, dzf);
    const float qrow = QUAD ? fmaf(a.q.q2y, dyf * dyf, a.q.q2z * (dzf * dzf)) : 0.f;
    const double t32 = (double)dzf * (double)(w * pair_k32<QUAD>(af, R, qrow, a.q));
    const int di = abs(i - c.x), dj = abs(j - c.y), dk = abs(k - c.z);
    const bool in_near = di <= a.nrx && dj <= a.nry && dk <= a.nrz;
    const bool in_box = di <= a.rx && dj <= a.ry && dk <= a.rz;
    const double ad = pd.x - (double)i;
    const double dyd = pd.y - (double)j * a.syd;
    const double dzd = pd.z - (double)k * a.szd;
    double better;
    if (mode == kCorrPrism && in_near) {
        better = prism_cell(pd.x, pd.y, pd.z, i, j, k, a.syd, a.szd) / a.vol;
    } else if (mode == kCorrPrism && in_box) {
        better = (double)gl3_cell(pf.x, pf.y, pf.z, i, j, k, a.syf, a.szf);
    } else {
        const double r2 = ad * ad + dyd * dyd + dzd * dzd + kEpsD;
        const double v2 = 1.0 / r2;
        better = dzd * v2 * sqrt(v2);
        if (QUAD && mode == kCorrPrism) {
            const double q = (double)a.q.q2x * ad * ad + (double)a.q.q2y * dyd * dyd +
                             (double)a.q.q2z * dzd * dzd;
            better *= 1.0 + v2 * (q * v2 - (double)a.q.qz);
        }
    }
    return (double)w * better - t32;
}

// Sum of corr_pair over the cells of the station's clamped rows that lie
// outside its correction box (radii rx, ry, rz); the box itself is corrected by
// the caller. Rows are the (j,k) within (jr, kr) of (c.y, c.z) whose FP32 row
// term was clamped (row_clamped). Lanes [lane0, lane0 + nlanes) split the
// cells of each row.
template <bool QUAD>
__device__ __forceinline__ double clamped_rows_corr(int mode, const float4& pf, const double4& pd,
                                                    const int4& c, const float* __restrict__ rho,
                                                    const NearArgs& a, int lane0, int nlanes) {
    double acc = 0.0;
    const int k0 = max(c.z - a.kr, 0), k1 = min(c.z + a.kr, a.nz - 1);
    const int j0 = max(c.y - a.jr, 0), j1 = min(c.y + a.jr, a.ny - 1);
    for (int kk = k0; kk <= k1; ++kk)
        for (int jj = j0; jj <= j1; ++jj) {
            const float dyf = pf.y - (float)jj * a.syf;
            const float dzf = pf.z - (float)kk * a.szf;
            if (!row_clamped(dyf, dzf)) continue;
            const bool row_in_box = abs(jj - c.y) <= a.ry && abs(kk - c.z) <= a.rz;
            for (int ii = lane0; ii < a.nx; ii += nlanes) {
                if (row_in_box && abs(ii - c.x) <= a.rx) continue;
                const float w = rho[((size_t)kk * a.ny + jj) * a.nx + ii];
                if (!(fabsf(w) > a.thr)) continue;
                acc += corr_pair<QUAD>(mode, w, pf, pd, c, ii, jj, kk, a);
            }
        }
    return acc;
}

}  // namespace ztm
