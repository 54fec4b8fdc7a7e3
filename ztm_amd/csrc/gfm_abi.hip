// extern "C" layer of libztm_gfm: error reporting, device checks, plans, the
// reference-signature entry points and dominant-kernel timing.
#include "gfm_internal.h"

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

namespace ztm {

namespace {
thread_local std::string g_last_error;
std::mutex g_mu;   // guards device-check cache, reference plans and profiling
}  // namespace

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    const int code = (e == hipErrorOutOfMemory) ? ZTM_ENOMEM
                     : (e == hipErrorNoDevice || e == hipErrorInvalidDevice) ? ZTM_ENODEV
                                                                              : ZTM_EHIP;
    return fail(code, "%s failed: %s", what, hipGetErrorString(e));
}

// ------------------------------------------------------------ profiling
namespace {
struct ProfRec {
    std::string name;
    hipEvent_t a, b;
};
bool g_prof_on = false;
std::vector<ProfRec> g_prof;
thread_local int g_prof_off = 0;
}  // namespace

ProfTimingOff::ProfTimingOff() { ++g_prof_off; }
ProfTimingOff::~ProfTimingOff() { --g_prof_off; }

ProfScope::ProfScope(const char* name, hipStream_t s) : name_(name), stream_(s) {
    if (g_prof_off) return;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_prof_on) return;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(stream_, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
        return;   // no timing events inside a graph capture
    if (hipEventCreate(&a_) != hipSuccess || hipEventCreate(&b_) != hipSuccess) {
        a_ = b_ = nullptr;
        return;
    }
    (void)hipEventRecord(a_, stream_);
}

ProfScope::~ProfScope() {
    if (!a_) return;
    (void)hipEventRecord(b_, stream_);
    std::lock_guard<std::mutex> lk(g_mu);
    g_prof.push_back({name_, a_, b_});
}

// ------------------------------------------------------------- devices
namespace {
std::map<int, int> g_dev_ok;   // device -> ZTM_OK / error code

int check_device(int* dev_out) {
    int dev = 0;
    ZTM_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_dev_ok.find(dev);
    if (it == g_dev_ok.end()) {
        hipDeviceProp_t p;
        hipError_t e = hipGetDeviceProperties(&p, dev);
        if (e != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
        const int ok = std::strncmp(p.gcnArchName, "gfx950", 6) == 0 ? ZTM_OK : ZTM_ENODEV;
        it = g_dev_ok.emplace(dev, ok).first;
        if (ok != ZTM_OK)
            return fail(ZTM_ENODEV, "device %d is %s; libztm_gfm is built for gfx950 only",
                        dev, p.gcnArchName);
    }
    if (it->second != ZTM_OK)
        return fail(ZTM_ENODEV, "device %d is not gfx950", dev);
    if (dev_out) *dev_out = dev;
    return ZTM_OK;
}
}  // namespace

}  // namespace ztm

using namespace ztm;


namespace ztm {

namespace {
// Grid-stride clear: 16-byte stores over n16 uint4s at p16, 4-byte stores over
// n4 words at p4.
__global__ void zero_fill_kernel(uint4* __restrict__ p16, int64_t n16,
                                 uint32_t* __restrict__ p4, int64_t n4) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int64_t i = t; i < n16; i += stride) p16[i] = make_uint4(0u, 0u, 0u, 0u);
    for (int64_t i = t; i < n4; i += stride) p4[i] = 0u;
}
}  // namespace

// Clears `bytes` (a multiple of 4) at p with a kernel, not a memset, so the
// call can sit inside a captured HIP graph.
int zero_fill(void* p, size_t bytes, hipStream_t stream) {
    if (bytes == 0) return ZTM_OK;
    if (!p || (bytes & 3) || ((uintptr_t)p & 3))
        return fail(ZTM_EINVAL, "zero_fill needs 4-byte aligned whole words");
    const int64_t n16 = ((uintptr_t)p & 15) ? 0 : (int64_t)(bytes / 16);
    const int64_t n4 = (int64_t)((bytes - (size_t)n16 * 16) / 4);
    const int64_t need = (std::max(n16, n4) + 255) / 256;
    const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>(need, 1 << 16));
    hipLaunchKernelGGL(zero_fill_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, (uint4*)p,
                       n16, (uint32_t*)((char*)p + n16 * 16), n4);
    ZTM_HIP(hipGetLastError());
    return ZTM_OK;
}

int grid_check(const ztm_grid& g) {
    if (g.nx <= 0 || g.ny <= 0 || g.nz <= 0)
        return fail(ZTM_EINVAL, "grid dims must be positive (got %d x %d x %d)", g.nx, g.ny, g.nz);
    if (!(std::isfinite(g.dx) && std::isfinite(g.dy) && std::isfinite(g.dz)) ||
        g.dx == 0.0 || g.dy == 0.0 || g.dz == 0.0)
        return fail(ZTM_EINVAL, "grid spacings must be finite and non-zero");
    const double n = (double)g.nx * g.ny * g.nz;
    if (n > 2147483647.0) return fail(ZTM_EINVAL, "grid has more than 2^31-1 cells");
    return ZTM_OK;
}

int ensure_workspace(ztm_plan p, size_t doubles) {
    if (doubles <= p->part_cap) return ZTM_OK;
    if (p->part) (void)hipFree(p->part);
    p->part = nullptr;
    p->part_cap = 0;
    ZTM_HIP(hipMalloc(&p->part, doubles * sizeof(double)));
    p->part_cap = doubles;
    return ZTM_OK;
}

// Grows the batched (ensemble) workspace to at least `bytes`.
int ensure_batched_workspace(ztm_plan p, size_t bytes) {
    if (bytes <= p->bws_cap) return ZTM_OK;
    if (p->bws) (void)hipFree(p->bws);
    p->bws = nullptr;
    p->bws_cap = 0;
    ZTM_HIP(hipMalloc(&p->bws, bytes));
    p->bws_cap = bytes;
    return ZTM_OK;
}

void plan_free(ztm_plan p) {
    if (!p) return;
    if (p->st_f) (void)hipFree(p->st_f);
    if (p->st_d) (void)hipFree(p->st_d);
    if (p->st_c) (void)hipFree(p->st_c);
    if (p->st_gs) (void)hipFree(p->st_gs);
    if (p->col_start) (void)hipFree(p->col_start);
    if (p->col_st) (void)hipFree(p->col_st);
    if (p->row_start) (void)hipFree(p->row_start);
    if (p->row_st) (void)hipFree(p->row_st);
    if (p->out_map) (void)hipFree(p->out_map);
    if (p->part) (void)hipFree(p->part);
    if (p->G) (void)hipFree(p->G);
    if (p->bws) (void)hipFree(p->bws);
    delete p;
}

int set_shape(ztm_plan p, int64_t m) {
    const ztm_grid& g = p->prob.grid;
    p->shape.nx = g.nx;
    p->shape.ny = g.ny;
    p->shape.nz = g.nz;
    p->shape.n_rows = (int64_t)g.ny * g.nz;
    p->shape.n_cells = (int64_t)g.nx * g.ny * g.nz;
    if (m > (int64_t)1 << 30) return fail(ZTM_EINVAL, "too many stations (%lld)", (long long)m);
    p->shape.m = (int)m;
    p->shape.m_pad = (int)(((m + kStationPad - 1) / kStationPad) * kStationPad);
    if (p->shape.m_pad == 0) p->shape.m_pad = kStationPad;
    p->L = std::fabs(g.dx);
    p->sy = g.dy / p->L;
    p->sz = g.dz / p->L;
    // Rows whose FP32 term a station can clamp lie within |dy|,|dz| < 0.5 of
    // it: at most ceil(0.5/|s| + 0.5) row indices from its nearest one (+1
    // for rounding).
    p->jr = (int)std::ceil(0.5 / std::fabs(p->sy) + 0.5) + 1;
    p->kr = (int)std::ceil(0.5 / std::fabs(p->sz) + 0.5) + 1;
    p->rmy = std::max(kColMargin, p->jr + 1);
    p->rmz = std::max(kColMargin, p->kr + 1);
    // Correction radii per axis in index units, scaled by the largest cell edge
    // so that each zone covers the same distance on every axis.
    const double ex = 1.0, ey = std::fabs(p->sy), ez = std::fabs(p->sz);
    const double emax = std::max(ex, std::max(ey, ez));
    auto rad = [&](int k, double e) { return (int)std::ceil(k * emax / e - 1e-9); };
    p->nrx = rad(kNear, ex);
    p->nry = rad(kNear, ey);
    p->nrz = rad(kNear, ez);
    const bool prism = p->prob.formula == ZTM_PRISM;
    p->rx = prism ? rad(kMid, ex) : p->nrx;
    p->ry = prism ? rad(kMid, ey) : p->nry;
    p->rz = prism ? rad(kMid, ez) : p->nrz;
    p->cmx = std::max(kColMargin, p->rx + 1);
    p->cmy = std::max(kColMargin, p->ry + 1);
    p->clamp_margin = std::max(std::max(p->cmx, p->cmy), std::max(p->rz + 1, std::max(p->rmy, p->rmz)));
    return ZTM_OK;
}

// Counting sort of stations by key (key < 0: left out), stable in station
// order: start[nkeys + 1], list[start[nkeys]] (at least one entry).
void build_csr(const std::vector<int>& key, int nkeys, std::vector<int>& start,
               std::vector<int>& list) {
    start.assign((size_t)nkeys + 1, 0);
    for (int k : key)
        if (k >= 0) ++start[(size_t)k + 1];
    for (size_t c = 1; c < start.size(); ++c) start[c] += start[c - 1];
    list.assign(std::max(1, start.back()), 0);
    std::vector<int> fill(start.begin(), start.end() - 1);
    for (size_t m = 0; m < key.size(); ++m)
        if (key[m] >= 0) list[fill[key[m]]++] = (int)m;
}

template <typename T>
hipError_t upload(T** dst, const std::vector<T>& src) {
    hipError_t e = hipMalloc(dst, sizeof(T) * src.size());
    if (e == hipSuccess) e = hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice);
    return e;
}

NearArgs near_args(const ztm_plan p) {
    NearArgs a;
    const bool prism = p->prob.formula == ZTM_PRISM;
    a.mode = prism ? kCorrPrism : kCorrPmNear;
    a.rx = p->rx;
    a.ry = p->ry;
    a.rz = p->rz;
    a.nrx = p->nrx;
    a.nry = p->nry;
    a.nrz = p->nrz;
    a.cmx = p->cmx;
    a.cmy = p->cmy;
    a.st_f = p->st_f;
    a.st_d = p->st_d;
    a.st_c = p->st_c;
    a.col_start = p->col_start;
    a.col_st = p->col_st;
    a.nx = p->shape.nx;
    a.ny = p->shape.ny;
    a.nz = p->shape.nz;
    a.syf = (float)p->sy;
    a.szf = (float)p->sz;
    a.thr = (float)p->prob.threshold;
    a.syd = p->sy;
    a.szd = p->sz;
    a.vol = std::fabs(p->sy * p->sz);
    a.kmin = p->kmin;
    a.kmax = p->kmax;
    a.row_start = p->row_start;
    a.row_st = p->row_st;
    a.jr = p->jr;
    a.kr = p->kr;
    a.rmy = p->rmy;
    a.rmz = p->rmz;
    a.rkmin = p->rkmin;
    a.rkmax = p->rkmax;
    // Quadrupole term of the far field for non-cubic prism cells (half-widths
    // in cell units); it vanishes identically for cubes.
    const double hx = 0.5, hy = 0.5 * std::fabs(p->sy), hz = 0.5 * std::fabs(p->sz);
    const bool cubic = std::fabs(hy - hx) <= 1e-12 && std::fabs(hz - hx) <= 1e-12;
    a.quad = (prism && !cubic) ? 1 : 0;
    if (a.quad) {
        const double qxx = (2 * hx * hx - hy * hy - hz * hz) / 3.0;
        const double qyy = (2 * hy * hy - hx * hx - hz * hz) / 3.0;
        const double qzz = (2 * hz * hz - hx * hx - hy * hy) / 3.0;
        a.q.q2x = (float)(2.5 * qxx);
        a.q.q2y = (float)(2.5 * qyy);
        a.q.q2z = (float)(2.5 * qzz);
        a.q.qz = (float)qzz;
    }
    return a;
}

// The prism zones of the FP32 path for ZTM_FP64_ZONED: FP64 far field with
// the quadrupole term, FP64 corrections inside the correction box, and no
// clamped rows (FP64 row terms are never clamped).
NearArgs near_args64(const ztm_plan p) {
    NearArgs a = near_args(p);
    a.mode = kCorrPrism64;
    a.row_start = nullptr;
    a.row_st = nullptr;
    a.rkmin = 1;
    a.rkmax = 0;
    return a;
}

BatchedKernel batched_kernel_in_force(const ztm_plan p, int direction, int n_models) {
    BatchedKernel k = p->tuning.batched[direction][batched_class(n_models)];
    const int split = direction == 0 ? p->tuning.batched_split : batched_adjoint_split(p);
    if (split == 3) k.variant = 0;   // variant 1 is for the two-part splits
    if (k.variant == 0) k.waves = 4;
    return k;
}

int batched_adjoint_split(const ztm_plan p) {
    const int split = p->tuning.batched_split;
    return split == 4 && !p->fp16_adjoint ? 3 : split;
}

double pm_scale(const ztm_plan p) {
    // kappa * V / L^2 turns  sum rho dz'/r'^3  (cell units) into g.
    const ztm_grid& g = p->prob.grid;
    const double V = std::fabs(g.dx * g.dy * g.dz);
    return p->prob.kappa * V / (p->L * p->L);
}

}  // namespace ztm

PmTuning ztm::effective_tuning(ztm_plan plan, int precision) {
    PmTuning t = plan->tuning;
    if (precision == ZTM_FP64) {
        if (t.st_per_lane > 4) t.st_per_lane = 4;
        if (t.cells_per_lane > 8) t.cells_per_lane = 8;
    }
    return t;
}

extern "C" {

int ztm_version(void) { return ZTM_GFM_VERSION; }

const char* ztm_strerror(int code) {
    switch (code) {
        case ZTM_OK: return "ok";
        case ZTM_EINVAL: return "invalid argument";
        case ZTM_EHIP: return "HIP runtime error";
        case ZTM_ENOMEM: return "out of memory";
        case ZTM_ENODEV: return "no usable gfx950 device";
        case ZTM_EIO: return "I/O error";
        default: return "unknown error";
    }
}

const char* ztm_last_error(void) { return g_last_error.c_str(); }

int ztm_init(void) {
    int dev = 0;
    if (int rc = check_device(&dev)) return rc;
    return noalloc_setup(dev, nullptr);
}

namespace {
// Largest |dz| r^-3 over all station-cell pairs, in the kernels' units and
// from the FP32 station records they read. For one station the distance h
// to the nearest cell column (nearest index per axis) is the same in every
// layer, and |dz| (h^2 + dz^2)^-1.5 rises then falls in |dz| on either side
// of the station, peaking at |dz| = h / sqrt(2); so the largest value lies
// within a layer of Z -+ h / sqrt(2), and two layers either side of each are
// tried. The batched forward also generates the cells that pad the grid to
// a multiple of 16 (raster indices N .. N16 - 1, past the last layer), so
// they are included. bound[m] is station m's largest value (0 where its G
// row is 0 throughout); the result is the largest over all stations.
double g_bound(const std::vector<float4>& hf, int64_t n, int nx, int ny, int nz, double sy,
               double sz, std::vector<double>& bound) {
    double best = 0.0;
    bound.assign((size_t)n, 0.0);
    if (nx <= 0 || ny <= 0 || nz <= 0) return 0.0;
    auto nearest = [](double v, int cnt) {
        return std::max(0.0, std::min((double)(cnt - 1), std::nearbyint(v)));
    };
    for (int64_t m = 0; m < n; ++m) {
        const double X = hf[m].x, Y = hf[m].y, Z = hf[m].z;
        const double dx = X - nearest(X, nx), dy = Y - sy * nearest(Y / sy, ny);
        const double h2 = dx * dx + dy * dy, t = std::sqrt(h2 * 0.5);
        for (int side = -1; side <= 1; side += 2) {
            const double k0 = std::nearbyint((Z + side * t) / sz);
            for (int d = -2; d <= 2; ++d) {
                const double k = std::max(0.0, std::min((double)(nz - 1), k0 + d));
                const double dz = Z - k * sz, r2 = h2 + dz * dz;
                if (r2 > 0.0)
                    bound[m] = std::max(bound[m], std::fabs(dz) / (r2 * std::sqrt(r2)));
            }
        }
        const int64_t n_cells = (int64_t)nx * ny * nz;
        for (int64_t c = n_cells; c < (n_cells + 15) / 16 * 16; ++c) {
            const int64_t rw = c / nx;
            const double cx = X - (double)(c - rw * nx), cy = Y - sy * (double)(rw % ny),
                         cz = Z - sz * (double)(rw / ny);
            const double r2 = cx * cx + cy * cy + cz * cz;
            if (r2 > 0.0) bound[m] = std::max(bound[m], std::fabs(cz) / (r2 * std::sqrt(r2)));
        }
        best = std::max(best, bound[m]);
    }
    return best;
}
}  // namespace

int ztm_plan_create(ztm_plan* plan, const ztm_problem* problem, int64_t n_stations,
                    const double* stations_host) {
    if (!plan || !problem) return fail(ZTM_EINVAL, "null plan or problem");
    *plan = nullptr;
    if (int rc = grid_check(problem->grid)) return rc;
    if (problem->formula != ZTM_POINT_MASS && problem->formula != ZTM_PRISM)
        return fail(ZTM_EINVAL, "unknown formula %d", problem->formula);
    if (n_stations < 0) return fail(ZTM_EINVAL, "negative station count");
    if (n_stations > 0 && !stations_host) return fail(ZTM_EINVAL, "null stations");
    if (!std::isfinite(problem->kappa)) return fail(ZTM_EINVAL, "kappa must be finite");
    int dev = 0;
    if (int rc = check_device(&dev)) return rc;

    ztm_plan p = new ztm_plan_s;
    p->device = dev;
    p->prob = *problem;
    if (int rc = set_shape(p, n_stations)) { plan_free(p); return rc; }

    const ztm_grid& g = problem->grid;
    std::vector<float4> hf(p->shape.m_pad, make_float4(0.f, 0.f, 0.f, 0.f));
    std::vector<double4> hd(p->shape.m_pad, make_double4(0.0, 0.0, 0.0, 0.0));
    std::vector<int4> hc(p->shape.m_pad, make_int4(0, 0, 0, 0));
    for (int64_t m = 0; m < n_stations; ++m) {
        const double x = stations_host[3 * m], y = stations_host[3 * m + 1],
                     z = stations_host[3 * m + 2];
        if (!(std::isfinite(x) && std::isfinite(y) && std::isfinite(z))) {
            plan_free(p);
            return fail(ZTM_EINVAL, "station %lld has a non-finite coordinate", (long long)m);
        }
        const double X = (x - g.x0) / g.dx;
        const double Y = (y - g.y0) / p->L;
        const double Z = (z - g.z0) / p->L;
        hf[m] = make_float4((float)X, (float)Y, (float)Z, 0.f);
        hd[m] = make_double4(X, Y, Z, 0.0);
        // nearest cell index per axis, clamped far outside the grid so it fits int
        const double cm = p->clamp_margin;
        auto idx = [cm](double v, int n) {
            const double c = std::nearbyint(v);
            return (int)std::max(-cm - 1.0, std::min(c, (double)n + cm));
        };
        hc[m] = make_int4(idx(X, g.nx), idx(Y / p->sy, g.ny), idx(Z / p->sz, g.nz), 0);
    }
    std::vector<double> gb;
    p->gmax = g_bound(hf, n_stations, g.nx, g.ny, g.nz, p->sy, p->sz, gb);
    // The FP16 adjoint scales G by one power of two per plan (its stations are
    // the MFMA's K dimension), 2^14 / gmax. A term below the FP16 normal range
    // then keeps an absolute error of at most 2^-25 in scaled units, i.e.
    // 2^-39 gmax in G units; over M stations that is M 2^-39 gmax per unit of
    // residual, against |G^T r| ~ bmin max|r| at the cell under the largest
    // residual (bmin = the smallest per-station bound). So the FP16 split stays
    // within ~2^-20 of max|G^T r| while gmax / bmin <= 2^19 / M: stations at
    // similar heights over the grid. Plans with a wider spread (stations at
    // very different heights, or one almost on a cell centre) run the batched
    // adjoint in the BF16 three-way split instead, whose parts keep the FP32
    // exponent.
    {
        double bmin = 0.0;
        for (double b : gb)
            if (b > 0.0) bmin = bmin > 0.0 ? std::min(bmin, b) : b;
        p->fp16_adjoint = n_stations == 0 || p->gmax <= 0.0 ||
                          p->gmax <= bmin * std::ldexp(1.0, 19) / (double)n_stations;
    }
    p->tuning.batched_split = 4;
    // per-station FP16 G scales of the batched forward (padding records: 1)
    std::vector<float> hgs(p->shape.m_pad, 1.f);
    for (int64_t m = 0; m < n_stations; ++m) hgs[m] = half_g_scale(gb[m]);
    // Adjoint gather indexes (stable counting sorts by key):
    //  columns: stations by nearest (ci, cj) over the (cmx, cmy)-extended range;
    //  rows:    stations that may clamp some grid row, by nearest (cj, ck).
    const int wcols = g.nx + 2 * p->cmx, hcols = g.ny + 2 * p->cmy;
    const int wkey = g.ny + 2 * p->rmy, hkey = g.nz + 2 * p->rmz;
    std::vector<int> col_key((size_t)n_stations, -1), row_key((size_t)n_stations, -1);
    bool any_near = false, any_row = false;
    for (int64_t m = 0; m < n_stations; ++m) {
        const int ci = hc[m].x + p->cmx, cj = hc[m].y + p->cmy;
        if (ci >= 0 && ci < wcols && cj >= 0 && cj < hcols) {
            col_key[m] = cj * wcols + ci;
            p->kmin = any_near ? std::min(p->kmin, hc[m].z) : hc[m].z;
            p->kmax = any_near ? std::max(p->kmax, hc[m].z) : hc[m].z;
            any_near = true;
        }
        const int rj = hc[m].y + p->rmy, rk = hc[m].z + p->rmz;
        if (rj < 0 || rj >= wkey || rk < 0 || rk >= hkey) continue;
        // superset of the kernels' FP32 test dy^2 + dz^2 < 0.25 (cell units)
        bool clamps = false;
        const double Y = hd[m].y, Z = hd[m].z;
        for (int k = std::max(0, hc[m].z - p->kr); k <= std::min(g.nz - 1, hc[m].z + p->kr) && !clamps; ++k)
            for (int j = std::max(0, hc[m].y - p->jr); j <= std::min(g.ny - 1, hc[m].y + p->jr); ++j) {
                const double dy = Y - j * p->sy, dz = Z - k * p->sz;
                const double tol = 1e-4 + 1e-6 * (std::fabs(Y) + std::fabs(Z) + std::fabs(j * p->sy) +
                                                  std::fabs(k * p->sz));
                if (std::fabs(dy) < 0.5 + tol && std::fabs(dz) < 0.5 + tol) {
                    clamps = true;
                    break;
                }
            }
        if (!clamps) continue;
        row_key[m] = rk * wkey + rj;
        p->rkmin = any_row ? std::min(p->rkmin, hc[m].z) : hc[m].z;
        p->rkmax = any_row ? std::max(p->rkmax, hc[m].z) : hc[m].z;
        any_row = true;
    }
    if (!any_near) {   // no station is near any cell: the column gather is skipped
        p->kmin = 1 << 29;
        p->kmax = -(1 << 29);
    }
    std::vector<int> col_start, col_st, row_start, row_st;
    build_csr(col_key, wcols * hcols, col_start, col_st);
    build_csr(row_key, wkey * hkey, row_start, row_st);
    hipError_t e = upload(&p->st_f, hf);
    if (e == hipSuccess) e = upload(&p->st_d, hd);
    if (e == hipSuccess) e = upload(&p->st_c, hc);
    if (e == hipSuccess) e = upload(&p->st_gs, hgs);
    if (e == hipSuccess) e = upload(&p->col_start, col_start);
    if (e == hipSuccess) e = upload(&p->col_st, col_st);
    if (e == hipSuccess && any_row) e = upload(&p->row_start, row_start);
    if (e == hipSuccess && any_row) e = upload(&p->row_st, row_st);
    if (e != hipSuccess) {
        plan_free(p);
        return hip_fail(e, "plan station upload");
    }
    *plan = p;
    return ZTM_OK;
}

int ztm_plan_retain(ztm_plan plan) {
    if (!plan) return fail(ZTM_EINVAL, "null plan");
    plan->refs.fetch_add(1, std::memory_order_relaxed);
    return ZTM_OK;
}

int ztm_plan_destroy(ztm_plan plan) {
    if (plan && plan->refs.fetch_sub(1, std::memory_order_acq_rel) == 1) plan_free(plan);
    return ZTM_OK;
}

int64_t ztm_plan_num_cells(ztm_plan plan) { return plan ? plan->shape.n_cells : -1; }
int64_t ztm_plan_num_stations(ztm_plan plan) { return plan ? plan->shape.m : -1; }

int ztm_plan_get_problem(ztm_plan plan, ztm_problem* problem) {
    if (!plan || !problem) return fail(ZTM_EINVAL, "null plan or problem");
    *problem = plan->prob;
    return ZTM_OK;
}

int ztm_plan_set_tuning(ztm_plan plan, int st_per_lane, int rows_per_slice,
                        int cells_per_lane, int adj_splits) {
    if (!plan) return fail(ZTM_EINVAL, "null plan");
    if (st_per_lane && st_per_lane != 1 && st_per_lane != 2 && st_per_lane != 4 && st_per_lane != 8)
        return fail(ZTM_EINVAL, "st_per_lane must be 1, 2, 4 or 8");
    if (cells_per_lane && cells_per_lane != 4 && cells_per_lane != 8 && cells_per_lane != 16)
        return fail(ZTM_EINVAL, "cells_per_lane must be 4, 8 or 16");
    if (rows_per_slice < 0 || adj_splits < 0) return fail(ZTM_EINVAL, "negative tuning value");
    if (st_per_lane) plan->tuning.st_per_lane = st_per_lane;
    if (rows_per_slice) plan->tuning.rows_per_slice = rows_per_slice;
    if (cells_per_lane) plan->tuning.cells_per_lane = cells_per_lane;
    if (adj_splits) plan->tuning.adj_splits = adj_splits;
    return ZTM_OK;
}

int ztm_plan_get_tuning(ztm_plan plan, int* st_per_lane, int* rows_per_slice,
                        int* cells_per_lane, int* adj_splits) {
    if (!plan) return fail(ZTM_EINVAL, "null plan");
    if (st_per_lane) *st_per_lane = plan->tuning.st_per_lane;
    if (rows_per_slice) *rows_per_slice = plan->tuning.rows_per_slice;
    if (cells_per_lane) *cells_per_lane = plan->tuning.cells_per_lane;
    if (adj_splits) *adj_splits = plan->tuning.adj_splits;
    return ZTM_OK;
}


int ztm_plan_forward(ztm_plan plan, int precision, const void* rho_dev, double* g_dev,
                     void* stream) {
    if (!plan) return fail(ZTM_EINVAL, "null plan");
    if (precision != ZTM_FP32 && precision != ZTM_FP64 && precision != ZTM_FP64_ZONED)
        return fail(ZTM_EINVAL, "unknown precision %d", precision);
    if (plan->shape.m > 0 && (!rho_dev || !g_dev)) return fail(ZTM_EINVAL, "null rho or g");
    int dev = 0;
    ZTM_HIP(hipGetDevice(&dev));
    if (dev != plan->device)
        return fail(ZTM_EINVAL, "plan belongs to device %d, current device is %d", plan->device, dev);
    const bool zoned = precision == ZTM_FP64_ZONED;
    if (zoned) precision = ZTM_FP64;   // FP64 arithmetic throughout
    const PmTuning t = effective_tuning(plan, precision);
    hipStream_t s = (hipStream_t)stream;
    if (plan->prob.formula == ZTM_PRISM) {
        if (precision == ZTM_FP64 && !zoned) {
            if (int rc = ensure_workspace(plan, prism64_workspace(plan->shape, t, false))) return rc;
            return prism64_fwd_launch(plan->shape, t, plan->st_d, (const double*)rho_dev, plan->sy,
                                      plan->sz, plan->prob.threshold, plan->prob.kappa * plan->L,
                                      plan->part, g_dev, s);
        }
        NearArgs na = zoned ? near_args64(plan) : near_args(plan);
        const bool quad = na.quad != 0;
        const size_t ws = pm_fwd_workspace(plan->shape, t, precision, quad);
        if (int rc = ensure_workspace(plan, ws + plan->shape.m_pad)) return rc;
        na.corr = plan->part + ws;
        if (zoned) {
            if (int rc = prism_near_fwd64_launch(plan->shape, na, (const double*)rho_dev,
                                                 plan->prob.threshold, plan->part + ws, s))
                return rc;
        } else if (int rc = prism_near_fwd_launch(plan->shape, na, (const float*)rho_dev,
                                                  plan->part + ws, s)) {
            return rc;
        }
        const void* st = zoned ? (const void*)plan->st_d : (const void*)plan->st_f;
        return pm_fwd_launch(plan->shape, t, precision, st, rho_dev, plan->sy, plan->sz,
                             plan->prob.threshold, pm_scale(plan), plan->part, g_dev, nullptr,
                             nullptr, na, s);
    }
    if (int rc = ensure_workspace(plan, pm_fwd_workspace(plan->shape, t, precision, false)))
        return rc;
    const void* st = precision == ZTM_FP64 ? (const void*)plan->st_d : (const void*)plan->st_f;
    return pm_fwd_launch(plan->shape, t, precision, st, rho_dev, plan->sy, plan->sz,
                         plan->prob.threshold, pm_scale(plan), plan->part, g_dev, nullptr,
                         nullptr, near_args(plan), s);
}

int ztm_plan_adjoint(ztm_plan plan, int precision, const void* r_dev, void* out_dev,
                     void* stream) {
    if (!plan) return fail(ZTM_EINVAL, "null plan");
    if (precision != ZTM_FP32 && precision != ZTM_FP64 && precision != ZTM_FP64_ZONED)
        return fail(ZTM_EINVAL, "unknown precision %d", precision);
    if (!out_dev || (plan->shape.m > 0 && !r_dev)) return fail(ZTM_EINVAL, "null r or out");
    int dev = 0;
    ZTM_HIP(hipGetDevice(&dev));
    if (dev != plan->device)
        return fail(ZTM_EINVAL, "plan belongs to device %d, current device is %d", plan->device, dev);
    const bool zoned = precision == ZTM_FP64_ZONED;
    if (zoned) precision = ZTM_FP64;
    const PmTuning t = effective_tuning(plan, precision);
    hipStream_t s = (hipStream_t)stream;
    if (plan->shape.m == 0) {
        const size_t bytes = (size_t)plan->shape.n_cells *
                             (precision == ZTM_FP64 ? sizeof(double) : sizeof(float));
        return zero_fill(out_dev, bytes, s);
    }
    const bool prism = plan->prob.formula == ZTM_PRISM;
    if (prism && precision == ZTM_FP64 && !zoned) {
        if (int rc = ensure_workspace(plan, prism64_workspace(plan->shape, t, true))) return rc;
        return prism64_adj_launch(plan->shape, t, plan->st_d, (const double*)r_dev, plan->sy,
                                  plan->sz, plan->prob.kappa * plan->L, plan->part,
                                  (double*)out_dev, s);
    }
    const bool zoned_prism = zoned && prism;
    const NearArgs na = zoned_prism ? near_args64(plan) : near_args(plan);
    if (int rc = ensure_workspace(
            plan, std::max<size_t>(1, pm_adj_workspace(plan->shape, t, precision,
                                                       precision == ZTM_FP32 || zoned_prism,
                                                       na.quad != 0))))
        return rc;
    const void* st = precision == ZTM_FP64 ? (const void*)plan->st_d : (const void*)plan->st_f;
    return pm_adj_launch(plan->shape, t, precision, st, r_dev, plan->sy, plan->sz,
                         pm_scale(plan), plan->part, out_dev, na, s);
}

int64_t ztm_plan_store_bytes(ztm_plan plan) {
    if (!plan) return -1;
    return (int64_t)sizeof(float) * plan->shape.m * stored_ld(plan->shape);
}

int ztm_plan_store(ztm_plan plan, void* stream) {
    if (!plan) return fail(ZTM_EINVAL, "null plan");
    int dev = 0;
    ZTM_HIP(hipGetDevice(&dev));
    if (dev != plan->device)
        return fail(ZTM_EINVAL, "plan belongs to device %d, current device is %d", plan->device, dev);
    if (plan->G) return ZTM_OK;
    const size_t bytes = (size_t)ztm_plan_store_bytes(plan);
    if (bytes == 0) return ZTM_OK;
    hipError_t e = hipMalloc(&plan->G, bytes);
    if (e != hipSuccess) {
        plan->G = nullptr;
        (void)hipGetLastError();
        return fail(ZTM_ENOMEM, "stored matrix needs %.3g GB of device memory: %s", bytes / 1e9,
                    hipGetErrorString(e));
    }
    const bool prism = plan->prob.formula == ZTM_PRISM;
    const double scale = prism ? plan->prob.kappa * plan->L : pm_scale(plan);
    const int rc = stored_build_launch(plan->shape, prism, plan->st_d, plan->sy, plan->sz, scale,
                                       plan->G, (hipStream_t)stream);
    if (rc != ZTM_OK) {
        (void)hipFree(plan->G);
        plan->G = nullptr;
    }
    return rc;
}

int ztm_plan_store_release(ztm_plan plan) {
    if (!plan) return fail(ZTM_EINVAL, "null plan");
    if (plan->G) ZTM_HIP(hipFree(plan->G));
    plan->G = nullptr;
    return ZTM_OK;
}

int ztm_plan_forward_stored(ztm_plan plan, const float* rho_dev, double* g_dev, void* stream) {
    if (!plan) return fail(ZTM_EINVAL, "null plan");
    if (plan->shape.m == 0) return ZTM_OK;
    if (!plan->G) return fail(ZTM_EINVAL, "no stored matrix: call ztm_plan_store first");
    if (!rho_dev || !g_dev) return fail(ZTM_EINVAL, "null rho or g");
    return stored_fwd_launch(plan->shape, plan->G, rho_dev, plan->prob.threshold, g_dev,
                             (hipStream_t)stream);
}

int ztm_plan_adjoint_stored(ztm_plan plan, const float* r_dev, float* out_dev, void* stream) {
    if (!plan) return fail(ZTM_EINVAL, "null plan");
    if (!out_dev || (plan->shape.m > 0 && !r_dev)) return fail(ZTM_EINVAL, "null r or out");
    if (plan->shape.m > 0 && !plan->G)
        return fail(ZTM_EINVAL, "no stored matrix: call ztm_plan_store first");
    if (int rc = ensure_workspace(plan, (size_t)stored_adj_splits(plan->shape) * plan->shape.n_cells))
        return rc;
    return stored_adj_launch(plan->shape, plan->G, r_dev, plan->part, out_dev,
                             (hipStream_t)stream);
}

int ztm_plan_forward_batched(ztm_plan plan, int n_models, const float* rho_dev, double* g_dev,
                             void* stream) {
    if (!plan) return fail(ZTM_EINVAL, "null plan");
    if (n_models < 0 || n_models > 256)
        return fail(ZTM_EINVAL, "n_models must be in [0, 256] (got %d)", n_models);
    if (plan->prob.formula != ZTM_POINT_MASS)
        return fail(ZTM_EINVAL, "the batched forward supports the point-mass formula only");
    if (n_models == 0 || plan->shape.m == 0) return ZTM_OK;
    if (!rho_dev || !g_dev) return fail(ZTM_EINVAL, "null rho or g");
    int dev = 0;
    ZTM_HIP(hipGetDevice(&dev));
    if (dev != plan->device)
        return fail(ZTM_EINVAL, "plan belongs to device %d, current device is %d", plan->device, dev);
    // Up to 4 models: the multi-field FP32 forward (each pair's r^-3 formed
    // once, the plan's FP64 near-field correction per model), which costs
    // less than one 32-model MFMA tile's G generation.
    NearArgs na = near_args(plan);
    if (n_models <= 4 && pm_fwd_fields_supported(plan->shape, n_models) && !na.quad) {
        if (int rc = ensure_workspace(plan, pm_fwd_fields_workspace(plan->shape, n_models)))
            return rc;
        return pm_fwd_fields_launch(plan->shape, n_models, (const float4*)plan->st_f, rho_dev,
                                    plan->shape.n_cells, plan->sy, plan->sz,
                                    plan->prob.threshold, pm_scale(plan), plan->part, g_dev, na,
                                    (hipStream_t)stream);
    }
    const int split = plan->tuning.batched_split;
    const BatchedKernel k = batched_kernel_in_force(plan, 0, n_models);
    const int ks = k.slices > 0 ? k.slices : batched_default_slices(plan->shape, n_models);
    const size_t need = batched_workspace(plan->shape, n_models, ks, split);
    if (int rc = ensure_batched_workspace(plan, need)) return rc;
    return batched_fwd_launch(plan->shape, n_models, ks, split, (const float*)plan->st_f, rho_dev,
                              plan->prob.threshold, plan->sy, plan->sz, pm_scale(plan),
                              plan->st_gs, plan->bws, g_dev, (hipStream_t)stream, k.variant,
                              k.waves);
}

int ztm_plan_adjoint_batched(ztm_plan plan, int n_vectors, const float* r_dev, float* out_dev,
                             void* stream) {
    if (!plan) return fail(ZTM_EINVAL, "null plan");
    if (n_vectors < 0 || n_vectors > 256)
        return fail(ZTM_EINVAL, "n_vectors must be in [0, 256] (got %d)", n_vectors);
    if (plan->prob.formula != ZTM_POINT_MASS)
        return fail(ZTM_EINVAL, "the batched adjoint supports the point-mass formula only");
    if (n_vectors == 0 || plan->shape.n_cells == 0) return ZTM_OK;
    if (!out_dev || (plan->shape.m > 0 && !r_dev)) return fail(ZTM_EINVAL, "null r or out");
    int dev = 0;
    ZTM_HIP(hipGetDevice(&dev));
    if (dev != plan->device)
        return fail(ZTM_EINVAL, "plan belongs to device %d, current device is %d", plan->device, dev);
    // Up to 4 vectors: the FP32 adjoint per vector, the transpose of what the
    // batched forward runs for up to 4 models (one vector costs less than a
    // quarter of a 32-vector MFMA tile's G generation).
    if (n_vectors <= 4 && pm_fwd_fields_supported(plan->shape, n_vectors) &&
        !near_args(plan).quad) {
        for (int b = 0; b < n_vectors; ++b)
            if (int rc = ztm_plan_adjoint(plan, ZTM_FP32, r_dev + (size_t)b * plan->shape.m,
                                          out_dev + (size_t)b * plan->shape.n_cells, stream))
                return rc;
        return ZTM_OK;
    }
    const int split = batched_adjoint_split(plan);
    const BatchedKernel k = batched_kernel_in_force(plan, 1, n_vectors);
    const int ks = k.slices > 0 ? k.slices : batched_adj_default_slices(plan->shape, n_vectors);
    const size_t need = batched_adj_workspace(plan->shape, n_vectors, ks, split);
    if (int rc = ensure_batched_workspace(plan, need)) return rc;
    return batched_adj_launch(plan->shape, n_vectors, ks, split, (const float*)plan->st_f, r_dev,
                              plan->sy, plan->sz, pm_scale(plan), plan->gmax, plan->bws, out_dev,
                              (hipStream_t)stream, k.variant, k.waves);
}

// Shared checks of the FP64 ensemble calls: ZTM_OK, an error, or -1 when
// there is nothing to compute (no models, cells or stations).
static int batched64_check(ztm_plan plan, int n, const void* in, const void* out,
                           const char* what) {
    if (!plan) return fail(ZTM_EINVAL, "null plan");
    if (n < 0 || n > 256) return fail(ZTM_EINVAL, "%s must be in [0, 256] (got %d)", what, n);
    if (plan->prob.formula != ZTM_POINT_MASS)
        return fail(ZTM_EINVAL, "the FP64 ensemble supports the point-mass formula only");
    if (n == 0 || plan->shape.n_cells == 0) return -1;
    if (!out || (plan->shape.m > 0 && !in)) return fail(ZTM_EINVAL, "null input or output");
    int dev = 0;
    ZTM_HIP(hipGetDevice(&dev));
    if (dev != plan->device)
        return fail(ZTM_EINVAL, "plan belongs to device %d, current device is %d", plan->device, dev);
    return plan->shape.m == 0 ? -1 : ZTM_OK;
}

int ztm_plan_forward_batched64(ztm_plan plan, int n_models, const double* rho_dev,
                               double* g_dev, void* stream) {
    const int rc = batched64_check(plan, n_models, rho_dev, g_dev, "n_models");
    if (rc) return rc < 0 ? ZTM_OK : rc;
    if (int e = ensure_batched_workspace(plan, batched64_fwd_workspace(plan->shape, n_models)))
        return e;
    return batched64_fwd_launch(plan->shape, n_models, plan->st_d, rho_dev, plan->prob.threshold,
                                plan->sy, plan->sz, pm_scale(plan), (double*)plan->bws, g_dev,
                                (hipStream_t)stream);
}

int ztm_plan_adjoint_batched64(ztm_plan plan, int n_vectors, const double* r_dev,
                               double* out_dev, void* stream) {
    const int rc = batched64_check(plan, n_vectors, r_dev, out_dev, "n_vectors");
    if (rc > 0) return rc;
    if (rc < 0) {   // no stations: G^T r = 0 (nothing else to do)
        if (n_vectors == 0 || plan->shape.n_cells == 0) return ZTM_OK;
        return zero_fill(out_dev, sizeof(double) * (size_t)n_vectors * plan->shape.n_cells,
                         (hipStream_t)stream);
    }
    if (int e = ensure_batched_workspace(
            plan, std::max<size_t>(1, batched64_adj_workspace(plan->shape, n_vectors))))
        return e;
    return batched64_adj_launch(plan->shape, n_vectors, plan->st_d, r_dev, plan->sy, plan->sz,
                                pm_scale(plan), (double*)plan->bws, out_dev,
                                (hipStream_t)stream);
}

int64_t ztm_gn_workspace_bytes(int64_t rows, int n, int nd) {
    if (rows < 0 || n < 1 || nd < 1) return -1;
    return (int64_t)gn_workspace_bytes(rows, n, nd);
}

int ztm_gn_normal(int64_t rows, int n, const double* J_dev, const double* r_dev,
                  double* JJ_dev, double* jtr_dev, void* work_dev, void* stream) {
    if (rows < 0 || n < 1) return fail(ZTM_EINVAL, "need rows >= 0 and n >= 1");
    if (rows >= ((int64_t)1 << 40)) return fail(ZTM_EINVAL, "too many rows");
    if (!JJ_dev || (r_dev && !jtr_dev) || !work_dev || (rows > 0 && !J_dev))
        return fail(ZTM_EINVAL, "null J, JJ, jtr (with r) or workspace");
    hipStream_t s = (hipStream_t)stream;
    if (rows == 0) {   // no data: J^T J = 0, J^T r = 0 (r may be null: nothing to read)
        if (int rc = zero_fill(JJ_dev, sizeof(double) * (size_t)n * n, s)) return rc;
        return jtr_dev ? zero_fill(jtr_dev, sizeof(double) * (size_t)n, s) : ZTM_OK;
    }
    return gn_normal_launch(rows, n, J_dev, r_dev, JJ_dev, jtr_dev, work_dev, s);
}

int ztm_gn_damped_steps(int n, int nd, const double* JJ_dev, const double* jtr_dev,
                        const double* sig_dev, const double* dm_dev, const double* b_dev,
                        double* B_dev, void* work_dev, void* stream) {
    if (n < 1 || nd < 1) return fail(ZTM_EINVAL, "need n >= 1 and nd >= 1");
    if (!JJ_dev || !jtr_dev || !sig_dev || !dm_dev || !b_dev || !B_dev || !work_dev)
        return fail(ZTM_EINVAL, "null argument");
    return gn_steps_launch(n, nd, JJ_dev, jtr_dev, sig_dev, dm_dev, b_dev, B_dev, work_dev,
                           (hipStream_t)stream);
}

int ztm_gn_inverse_diag(int n, const double* JJ_dev, double d, double* out_dev, void* work_dev,
                        void* stream) {
    if (n < 1) return fail(ZTM_EINVAL, "need n >= 1");
    if (!JJ_dev || !out_dev || !work_dev) return fail(ZTM_EINVAL, "null argument");
    return gn_inverse_diag_launch(n, JJ_dev, d, out_dev, work_dev, (hipStream_t)stream);
}

int ztm_plan_get_batched_tuning(ztm_plan plan, int* slices, int* split, int* adjoint_split,
                                int* variant) {
    if (!plan) return fail(ZTM_EINVAL, "null plan");
    const BatchedKernel& k = plan->tuning.batched[0][3];
    if (slices) *slices = k.slices;
    if (split) *split = plan->tuning.batched_split;
    if (adjoint_split) *adjoint_split = batched_adjoint_split(plan);
    if (variant) *variant = k.variant;
    return ZTM_OK;
}

int ztm_plan_set_batched_variant(ztm_plan plan, int variant) {
    if (!plan) return fail(ZTM_EINVAL, "null plan");
    if (variant < 0 || variant > 2) return fail(ZTM_EINVAL, "variant must be 0, 1 or 2");
    for (auto& dir : plan->tuning.batched)
        for (BatchedKernel& k : dir) {
            k.variant = variant;
            if (variant == 0) k.waves = 4;
        }
    return ZTM_OK;
}

int ztm_plan_set_batched_kernel(ztm_plan plan, int direction, int n_models, int variant,
                                int waves, int slices) {
    if (!plan) return fail(ZTM_EINVAL, "null plan");
    if (direction != 0 && direction != 1) return fail(ZTM_EINVAL, "direction must be 0 or 1");
    if (n_models < 1 || n_models > 256) return fail(ZTM_EINVAL, "n_models must be in [1, 256]");
    if (variant < 0 || variant > 2) return fail(ZTM_EINVAL, "variant must be 0, 1 or 2");
    if (waves != 4 && waves != 8) return fail(ZTM_EINVAL, "waves must be 4 or 8");
    if (waves == 8 && variant == 0) return fail(ZTM_EINVAL, "8 waves per workgroup need variant 1 or 2");
    if (slices < 0) return fail(ZTM_EINVAL, "negative slice count");
    plan->tuning.batched[direction][batched_class(n_models)] = BatchedKernel{variant, waves, slices};
    return ZTM_OK;
}

int ztm_plan_get_batched_kernel(ztm_plan plan, int direction, int n_models, int* variant,
                                int* waves, int* slices) {
    if (!plan) return fail(ZTM_EINVAL, "null plan");
    if (direction != 0 && direction != 1) return fail(ZTM_EINVAL, "direction must be 0 or 1");
    if (n_models < 1 || n_models > 256) return fail(ZTM_EINVAL, "n_models must be in [1, 256]");
    const BatchedKernel k = batched_kernel_in_force(plan, direction, n_models);
    if (variant) *variant = k.variant;
    if (waves) *waves = k.waves;
    if (slices) *slices = k.slices;
    return ZTM_OK;
}

namespace {
// Pseudo-random values in [-1, 1) (an integer hash of the index): the MFMA
// kernels' clock, and so their times, depend on how much their operands'
// bits toggle -- on low-entropy data such as alternating 1 and -0.5 the same
// kernels ran 20-25% faster than on real models -- so the timed inputs must
// look like real ones.
__global__ void fill_random_kernel(float* __restrict__ x, int64_t n, unsigned seed) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        unsigned h = (unsigned)i * 0x9E3779B1u ^ seed;
        h ^= h >> 16;
        h *= 0x85EBCA6Bu;
        h ^= h >> 13;
        h *= 0xC2B2AE35u;
        h ^= h >> 16;
        x[i] = (float)(h >> 8) * (1.f / 8388608.f) - 1.f;
    }
}
}  // namespace

// Times every kernel setting of the batched forward and adjoint for n_models
// models on this plan (variant 0; variant 1 with 4 and 8 waves per
// workgroup; for the forward also variant 2, loads two chunks ahead; the
// automatic K-slice count and twice it, never fewer, which would lengthen the
// FP32 accumulation runs) and keeps the fastest per direction for the
// model-count class of n_models. The split is left as it
// is: tuning changes speed, not precision.
int ztm_plan_autotune_batched(ztm_plan plan, int n_models, int reps, void* stream) {
    if (!plan) return fail(ZTM_EINVAL, "null plan");
    if (n_models < 1 || n_models > 256) return fail(ZTM_EINVAL, "n_models must be in [1, 256]");
    if (reps < 1) return fail(ZTM_EINVAL, "reps must be positive");
    if (plan->prob.formula != ZTM_POINT_MASS)
        return fail(ZTM_EINVAL, "the batched calls support the point-mass formula only");
    const int64_t N = plan->shape.n_cells, M = plan->shape.m;
    if (N == 0 || M == 0) return ZTM_OK;
    if (n_models <= 4 && pm_fwd_fields_supported(plan->shape, n_models) && !near_args(plan).quad)
        return ZTM_OK;   // the FP32 point-mass path: no MFMA kernel to tune
    int dev = 0;
    ZTM_HIP(hipGetDevice(&dev));
    if (dev != plan->device)
        return fail(ZTM_EINVAL, "plan belongs to device %d, current device is %d", plan->device, dev);
    hipStream_t s = (hipStream_t)stream;
    const int cls = batched_class(n_models);
    float *P = nullptr, *R = nullptr, *X = nullptr;
    double* Y = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&P, sizeof(float) * N * n_models);
    if (e == hipSuccess) e = hipMalloc(&R, sizeof(float) * M * n_models);
    if (e == hipSuccess) e = hipMalloc(&X, sizeof(float) * N * n_models);
    if (e == hipSuccess) e = hipMalloc(&Y, sizeof(double) * M * n_models);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    int rc = e == hipSuccess ? ZTM_OK : hip_fail(e, "autotune buffers");
    if (rc == ZTM_OK) {
        hipLaunchKernelGGL(fill_random_kernel, dim3(4096), dim3(256), 0, s, P, N * n_models, 1u);
        hipLaunchKernelGGL(fill_random_kernel, dim3(1024), dim3(256), 0, s, R, M * n_models, 2u);
        if (hipGetLastError() != hipSuccess) rc = fail(ZTM_EHIP, "autotune fill launch failed");
    }
    for (int dir = 0; dir < 2 && rc == ZTM_OK; ++dir) {
        const int split = dir == 0 ? plan->tuning.batched_split : batched_adjoint_split(plan);
        const int auto_ks = dir == 0 ? batched_default_slices(plan->shape, n_models)
                                     : batched_adj_default_slices(plan->shape, n_models);
        const int64_t kdim = dir == 0 ? (N + 15) / 16 : (M + 15) / 16;
        std::vector<int> counts{auto_ks};
        const int twice = (int)std::min<int64_t>(2 * (int64_t)auto_ks, std::min<int64_t>(kdim, 65535));
        if (twice > auto_ks) counts.push_back(twice);
        std::vector<BatchedKernel> cand;
        for (int ks : counts) {
            cand.push_back(BatchedKernel{0, 4, ks});
            if (split != 3) {
                cand.push_back(BatchedKernel{1, 4, ks});
                cand.push_back(BatchedKernel{1, 8, ks});
                if (dir == 0) {   // the forward's two-ahead prefetch
                    cand.push_back(BatchedKernel{2, 4, ks});
                    cand.push_back(BatchedKernel{2, 8, ks});
                }
            }
        }
        const BatchedKernel keep = plan->tuning.batched[dir][cls];
        float best = INFINITY;
        BatchedKernel won = keep;
        auto call = [&]() {
            return dir == 0 ? ztm_plan_forward_batched(plan, n_models, P, Y, stream)
                            : ztm_plan_adjoint_batched(plan, n_models, R, X, stream);
        };
        for (const BatchedKernel& c : cand) {
            plan->tuning.batched[dir][cls] = c;
            if ((rc = call()) != ZTM_OK) break;   // warm-up (grows the workspace)
            if ((e = hipEventRecord(e0, s)) != hipSuccess) { rc = hip_fail(e, "event"); break; }
            for (int r = 0; r < reps && rc == ZTM_OK; ++r) rc = call();
            if (rc != ZTM_OK) break;
            if ((e = hipEventRecord(e1, s)) == hipSuccess) e = hipEventSynchronize(e1);
            float ms = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (e != hipSuccess) { rc = hip_fail(e, "autotune timing"); break; }
            if (ms < best) {
                best = ms;
                won = c;
            }
        }
        plan->tuning.batched[dir][cls] = rc == ZTM_OK ? won : keep;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    for (void* q : {(void*)P, (void*)R, (void*)X, (void*)Y})
        if (q) (void)hipFree(q);
    return rc;
}

int ztm_plan_set_batched_tuning(ztm_plan plan, int slices, int split) {
    if (!plan) return fail(ZTM_EINVAL, "null plan");
    if (slices < 0) return fail(ZTM_EINVAL, "negative slice count");
    if (split != 0 && split != 2 && split != 3 && split != 4)
        return fail(ZTM_EINVAL, "split must be 2, 3 or 4");
    if (split == 4 && !(plan->gmax >= 0.0 && (plan->st_gs || plan->shape.m == 0)))
        return fail(ZTM_EINVAL, "the FP16 split (4) needs a plan made by ztm_plan_create");
    for (auto& dir : plan->tuning.batched)
        for (BatchedKernel& k : dir) k.slices = slices;
    if (split) plan->tuning.batched_split = split;
    return ZTM_OK;
}

int ztm_plan_smoothness(ztm_plan plan, int precision, const void* rho_dev, void* out_dev,
                        void* stream) {
    if (!plan) return fail(ZTM_EINVAL, "null plan");
    if (precision != ZTM_FP32 && precision != ZTM_FP64)
        return fail(ZTM_EINVAL, "unknown precision %d", precision);
    if (!rho_dev || !out_dev) return fail(ZTM_EINVAL, "null rho or out");
    if (rho_dev == out_dev) return fail(ZTM_EINVAL, "rho and out must not alias");
    return ltl_launch(precision, rho_dev, out_dev, plan->shape.nx, plan->shape.ny,
                      plan->shape.nz, (hipStream_t)stream);
}

// ------------------------------------------------ reference-signature ABI
extern "C++" {
namespace ztm {
int ref_check(int bfr, int di, int dj, int dk, int fh, int fv) {
    if (di <= 0 || dj <= 0 || dk <= 0 || fh <= 0 || fv <= 0 || bfr < 0)
        return fail(ZTM_EINVAL, "bad reference geometry bfr=%d di=%d dj=%d dk=%d fh=%d fv=%d",
                    bfr, di, dj, dk, fh, fv);
    const double n = ((double)(di - 1) * fh + 1) * ((double)(dj - 1) * fh + 1) *
                     ((double)(dk - 1) * fv + 1);
    if (n > 2147483647.0) return fail(ZTM_EINVAL, "reference grid too large");
    return ZTM_OK;
}

// The reference geometry's problem and shape (index units, kappa 66.7, the
// .001f skip, interior stations) in p, without allocating anything.
int ref_plan_init(ztm_plan p, int dev, int bfr, int di, int dj, int dk, int fh, int fv) {
    p->device = dev;
    ztm_problem& pr = p->prob;
    pr.grid.nx = (di - 1) * fh + 1;
    pr.grid.ny = (dj - 1) * fh + 1;
    pr.grid.nz = (dk - 1) * fv + 1;
    pr.grid.x0 = pr.grid.y0 = pr.grid.z0 = 0.0;
    pr.grid.dx = pr.grid.dy = 1.0 / fh;
    pr.grid.dz = 1.0 / fv;
    pr.formula = ZTM_POINT_MASS;
    pr.kappa = ZTM_KAPPA_REF;
    pr.threshold = 0.001f;   // GFM_GPU.cu:19 compares in float against .001f
    const int ni = std::max(0, di - 2 * bfr), nj = std::max(0, dj - 2 * bfr);
    return set_shape(p, (int64_t)ni * nj);
}
}  // namespace ztm
}  // extern "C++"

namespace {
using RefKey = std::tuple<int, int, int, int, int, int, int>;   // dev,bfr,di,dj,dk,fh,fv
std::map<RefKey, ztm_plan> g_ref_plans;

// One cached plan per (device, geometry); station heights change per call and
// are re-prepared on the device by prep_ref_stations().
int ref_plan(int bfr, int di, int dj, int dk, int fh, int fv, ztm_plan* out) {
    int dev = 0;
    if (int rc = check_device(&dev)) return rc;
    std::lock_guard<std::mutex> lk(g_mu);
    const RefKey key{dev, bfr, di, dj, dk, fh, fv};
    auto it = g_ref_plans.find(key);
    if (it != g_ref_plans.end()) {
        *out = it->second;
        return ZTM_OK;
    }
    ztm_plan p = new ztm_plan_s;
    if (int rc = ref_plan_init(p, dev, bfr, di, dj, dk, fh, fv)) { plan_free(p); return rc; }
    hipError_t e = hipMalloc(&p->st_f, sizeof(float4) * p->shape.m_pad);
    if (e == hipSuccess) e = hipMalloc(&p->st_d, sizeof(double4) * p->shape.m_pad);
    if (e == hipSuccess) e = hipMalloc(&p->st_c, sizeof(int4) * p->shape.m_pad);
    if (e == hipSuccess) e = hipMalloc(&p->out_map, sizeof(int) * p->shape.m_pad);
    if (e != hipSuccess) {
        plan_free(p);
        return hip_fail(e, "reference plan allocation");
    }
    if (int rc = ensure_workspace(p, pm_fwd_workspace(p->shape, p->tuning, ZTM_FP32, false))) {
        plan_free(p);
        return rc;
    }
    g_ref_plans.emplace(key, p);
    *out = p;
    return ZTM_OK;
}

// Producer arguments: the interpolation stencils need at least 2 points per
// axis; RC[0..6] must be finite.
int tfm_check(int di, int dj, int dk, int fh, int fv, const double* rc) {
    if (di < 2 || dj < 2 || dk < 2 || fh <= 0 || fv <= 0)
        return fail(ZTM_EINVAL, "producer grid needs di, dj, dk >= 2 and fh, fv >= 1 "
                                "(di=%d dj=%d dk=%d fh=%d fv=%d)", di, dj, dk, fh, fv);
    if (int e = ref_check(0, di, dj, dk, fh, fv)) return e;
    if (!rc) return fail(ZTM_EINVAL, "null RC frame");
    for (int q = 0; q < 7; ++q)
        if (!std::isfinite(rc[q])) return fail(ZTM_EINVAL, "RC[%d] is not finite", q);
    return ZTM_OK;
}

// Device buffers of the in-process gravity FM, one set per (device,
// geometry), kept for the process like the reference-ABI plans.
struct TfmBuffers {
    double *sea = nullptr, *moho = nullptr, *vel = nullptr;
    float *s = nullptr, *drho = nullptr, *g = nullptr;
    int64_t n = 0;
};
std::map<RefKey, TfmBuffers> g_tfm_bufs;

int tfm_buffers(int dev, int di, int dj, int dk, int fh, int fv, TfmBuffers* out) {
    std::lock_guard<std::mutex> lk(g_mu);
    const RefKey key{dev, 0, di, dj, dk, fh, fv};
    auto it = g_tfm_bufs.find(key);
    if (it != g_tfm_bufs.end()) {
        *out = it->second;
        return ZTM_OK;
    }
    TfmBuffers b;
    const size_t M = (size_t)di * dj;
    b.n = ((int64_t)(di - 1) * fh + 1) * ((int64_t)(dj - 1) * fh + 1) * ((int64_t)(dk - 1) * fv + 1);
    hipError_t e = hipMalloc(&b.sea, sizeof(double) * M);
    if (e == hipSuccess) e = hipMalloc(&b.moho, sizeof(double) * M);
    if (e == hipSuccess) e = hipMalloc(&b.vel, sizeof(double) * M * dk);
    if (e == hipSuccess) e = hipMalloc(&b.s, sizeof(float) * M);
    if (e == hipSuccess) e = hipMalloc(&b.g, sizeof(float) * M);
    if (e == hipSuccess) e = hipMalloc(&b.drho, sizeof(float) * b.n);
    if (e != hipSuccess) {
        for (void* q : {(void*)b.sea, (void*)b.moho, (void*)b.vel, (void*)b.s, (void*)b.g,
                        (void*)b.drho})
            if (q) (void)hipFree(q);
        return hip_fail(e, "gravity FM buffers");
    }
    g_tfm_bufs.emplace(key, b);
    *out = b;
    return ZTM_OK;
}
}  // namespace

int ztm_gf_async(int bfr, int di, int dj, int dk, int fh, int fv, float* g_dev,
                 const float* rho_dev, const float* s_dev, void* stream) {
    if (int rc = ref_check(bfr, di, dj, dk, fh, fv)) return rc;
    if (!g_dev || !rho_dev || !s_dev) return fail(ZTM_EINVAL, "null g, rho or s");
    ztm_plan p = nullptr;
    if (int rc = ref_plan(bfr, di, dj, dk, fh, fv, &p)) return rc;
    if (p->shape.m == 0) return ZTM_OK;   // no interior stations
    hipStream_t s = (hipStream_t)stream;
    if (int rc = prep_ref_stations(s_dev, bfr, di, dj, fh, fv, p->st_f, p->st_d, p->st_c,
                                   p->out_map, p->shape.m, p->shape.m_pad, s))
        return rc;
    return pm_fwd_launch(p->shape, p->tuning, ZTM_FP32, p->st_f, rho_dev, p->sy, p->sz,
                         p->prob.threshold, pm_scale(p), p->part, nullptr, g_dev, p->out_map,
                         near_args(p), s);
}

int ztm_gf(int bfr, int di, int dj, int dk, int fh, int fv, float* g_dev,
           const float* rho_dev, const float* s_dev) {
    return ztm_gf_async(bfr, di, dj, dk, fh, fv, g_dev, rho_dev, s_dev, nullptr);
}

int ztm_gfm_host(int bfr, int di, int dj, int dk, int fh, int fv, float* g_host,
                 const float* rho_host, int64_t n_rho, const float* s_host) {
    if (int rc = ref_check(bfr, di, dj, dk, fh, fv)) return rc;
    if (!g_host || !rho_host || !s_host) return fail(ZTM_EINVAL, "null host buffer");
    const int64_t N = ((int64_t)(di - 1) * fh + 1) * ((int64_t)(dj - 1) * fh + 1) *
                      ((int64_t)(dk - 1) * fv + 1);
    if (n_rho != N)
        return fail(ZTM_EINVAL, "rho has %lld values, geometry needs ii*jj*kk = %lld",
                    (long long)n_rho, (long long)N);
    if (int rc = check_device(nullptr)) return rc;
    const size_t M = (size_t)di * dj;
    float *r = nullptr, *g = nullptr, *s = nullptr;
    int rc = ZTM_OK;
    hipError_t e = hipMalloc(&r, sizeof(float) * N);
    if (e == hipSuccess) e = hipMalloc(&g, sizeof(float) * M);
    if (e == hipSuccess) e = hipMalloc(&s, sizeof(float) * M);
    if (e == hipSuccess) e = hipMemcpy(r, rho_host, sizeof(float) * N, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s, s_host, sizeof(float) * M, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(g, 0, sizeof(float) * M);
    if (e != hipSuccess) rc = hip_fail(e, "GFM buffers");
    if (rc == ZTM_OK) rc = ztm_gf(bfr, di, dj, dk, fh, fv, g, r, s);
    if (rc == ZTM_OK) {
        e = hipMemcpy(g_host, g, sizeof(float) * M, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = hip_fail(e, "GFM result copy");
    }
    if (r) (void)hipFree(r);
    if (g) (void)hipFree(g);
    if (s) (void)hipFree(s);
    return rc;
}

// ------------------------------------- gravity producer (TFM.cpp:83-94)
int ztm_tfm_drho(int di, int dj, int dk, int fh, int fv, const double* rc, const double* sea_dev,
                 const double* moho_dev, const double* vel_dev, double vm, double ro, double dr,
                 float* drho_dev, void* stream) {
    if (int e = tfm_check(di, dj, dk, fh, fv, rc)) return e;
    if (!sea_dev || !moho_dev || !vel_dev || !drho_dev) return fail(ZTM_EINVAL, "null argument");
    if (int e = check_device(nullptr)) return e;
    return tfm_drho_launch(di, dj, dk, fh, fv, rc, sea_dev, moho_dev, vel_dev, vm, ro, dr,
                           drho_dev, (hipStream_t)stream);
}

int ztm_tfm_gravity_host(int bfr, int di, int dj, int dk, int fh, int fv, const double* rc,
                         const double* sea, const double* moho, const double* vel,
                         const double* surf, double vm, double ro, double dr, double* G,
                         float* drho_out) {
    if (int e = tfm_check(di, dj, dk, fh, fv, rc)) return e;
    if (int e = ref_check(bfr, di, dj, dk, fh, fv)) return e;
    if (!sea || !moho || !vel || !surf || !G) return fail(ZTM_EINVAL, "null host buffer");
    int dev = 0;
    if (int e = check_device(&dev)) return e;
    TfmBuffers b;
    if (int e = tfm_buffers(dev, di, dj, dk, fh, fv, &b)) return e;
    const size_t M = (size_t)di * dj, NV = M * dk;
    std::vector<float> s32(surf, surf + M), g32(M);   // S.dat narrowing, TFM.cpp:92
    hipError_t e = hipMemcpy(b.sea, sea, sizeof(double) * M, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b.moho, moho, sizeof(double) * M, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b.vel, vel, sizeof(double) * NV, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b.s, s32.data(), sizeof(float) * M, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(b.g, 0, sizeof(float) * M);
    if (e != hipSuccess) return hip_fail(e, "gravity FM inputs");
    if (int rc2 = tfm_drho_launch(di, dj, dk, fh, fv, rc, b.sea, b.moho, b.vel, vm, ro, dr,
                                  b.drho, nullptr))
        return rc2;
    if (int rc2 = ztm_gf_async(bfr, di, dj, dk, fh, fv, b.g, b.drho, b.s, nullptr)) return rc2;
    e = hipMemcpy(g32.data(), b.g, sizeof(float) * M, hipMemcpyDeviceToHost);
    if (e == hipSuccess && drho_out)
        e = hipMemcpy(drho_out, b.drho, sizeof(float) * b.n, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "gravity FM result copy");
    for (size_t m = 0; m < M; ++m) G[m] = g32[m];   // TFM.cpp:94
    return ZTM_OK;
}

namespace {
int tfm_fields_gravity(ztm_plan p, int bfr, int di, int dj, int fh, int fv, int F,
                       const float* fields, double thr, const float* s_dev, double* dg_dev,
                       hipStream_t s);

// Gravity at the interior stations of B float fields drho_dev[B][N] on the
// reference plan p (stations prepared from s_dev) with cell threshold thr,
// scattered to g_dev[B][di*dj] (0 outside): up to 16 fields through the
// FP32 multi-field forward in groups of at most 8 (tfm_fields_gravity), more
// on the split-BF16 ensemble forward. At the reference grid a group of 4
// costs ~0.19 ms and the ensemble call ~0.46 ms whatever B up to 32 (its
// 32-model tile is then mostly padding).
constexpr int kTfmFieldsMax = 16;
constexpr int kTfmFieldsGroup = 8;

int tfm_gravity_of_fields(ztm_plan p, int bfr, int di, int dj, int fh, int fv, int B,
                          const float* drho_dev, double thr, const float* s_dev, double* g_dev,
                          hipStream_t s) {
    if (B <= kTfmFieldsMax && pm_fwd_fields_supported(p->shape, std::min(B, kTfmFieldsGroup)) &&
        !near_args(p).quad) {
        const int64_t N = p->shape.n_cells;
        const int M = di * dj;
        for (int f0 = 0; f0 < B; f0 += kTfmFieldsGroup)
            if (int e = tfm_fields_gravity(p, bfr, di, dj, fh, fv,
                                           std::min(kTfmFieldsGroup, B - f0),
                                           drho_dev + f0 * N, thr, s_dev,
                                           g_dev + (size_t)f0 * M, s))
                return e;
        return ZTM_OK;
    }
    const int M = di * dj, m = p->shape.m;
    if (m == 0) return tfm_scatter_launch(nullptr, nullptr, 0, M, B, g_dev, s);
    if (int e = prep_ref_stations(s_dev, bfr, di, dj, fh, fv, p->st_f, p->st_d, p->st_c,
                                  p->out_map, m, p->shape.m_pad, s))
        return e;
    const int split = p->tuning.batched_split;
    const BatchedKernel k = batched_kernel_in_force(p, 0, B);
    const int ks = k.slices > 0 ? k.slices : batched_default_slices(p->shape, B);
    const size_t ws = (batched_workspace(p->shape, B, ks, split) + 255) / 256 * 256;
    const size_t need = ws + sizeof(double) * (size_t)B * m;
    if (int rc = ensure_batched_workspace(p, need)) return rc;
    double* Y = (double*)((char*)p->bws + ws);
    if (int e = batched_fwd_launch(p->shape, B, ks, split, (const float*)p->st_f, drho_dev, thr,
                                   p->sy, p->sz, pm_scale(p), nullptr, p->bws, Y, s, k.variant,
                                   k.waves))
        return e;
    return tfm_scatter_launch(Y, p->out_map, m, M, B, g_dev, s);
}
}  // namespace

namespace {
// Gravity at the interior stations of F fields fields[F][N] with cell
// threshold thr (0 for derivative fields, already masked by GF's skip at
// the parameter point): up to 8 fields through one FP32 GF forward that
// forms each pair's r^-3 once (each field with its FP64 near-field
// correction; one forward per field when F density tiles of a row do not
// fit LDS), more through tfm_gravity_of_fields; dg [F][di*dj].
int tfm_fields_gravity(ztm_plan p, int bfr, int di, int dj, int fh, int fv, int F,
                       const float* fields, double thr, const float* s_dev, double* dg_dev,
                       hipStream_t s) {
    if (F > kTfmFieldsGroup) return tfm_gravity_of_fields(p, bfr, di, dj, fh, fv, F, fields, thr, s_dev, dg_dev,
                                            s);
    const int M = di * dj, m = p->shape.m;
    if (m == 0) return tfm_scatter_launch(nullptr, nullptr, 0, M, F, dg_dev, s);
    if (int e = prep_ref_stations(s_dev, bfr, di, dj, fh, fv, p->st_f, p->st_d, p->st_c,
                                  p->out_map, m, p->shape.m_pad, s))
        return e;
    if (int rc = ensure_batched_workspace(p, sizeof(double) * F * (size_t)m)) return rc;
    double* Y = (double*)p->bws;
    NearArgs na = near_args(p);
    na.thr = (float)thr;
    const int64_t N = p->shape.n_cells;
    if (pm_fwd_fields_supported(p->shape, F) && !na.quad) {
        if (int rc = ensure_workspace(p, pm_fwd_fields_workspace(p->shape, F))) return rc;
        if (int e = pm_fwd_fields_launch(p->shape, F, (const float4*)p->st_f, fields, N, p->sy,
                                         p->sz, thr, pm_scale(p), p->part, Y, na, s))
            return e;
    } else {
        for (int f = 0; f < F; ++f)
            if (int e = pm_fwd_launch(p->shape, p->tuning, ZTM_FP32, p->st_f, fields + f * N,
                                      p->sy, p->sz, thr, pm_scale(p), p->part,
                                      Y + (size_t)f * m, nullptr, nullptr, na, s))
                return e;
    }
    return tfm_scatter_launch(Y, p->out_map, m, M, F, dg_dev, s);
}
}  // namespace

int ztm_tfm_gravity_batched(int bfr, int di, int dj, int dk, int fh, int fv, int n_models,
                            const double* rc, const double* sea_dev, const double* moho_dev,
                            int64_t moho_stride, const double* vel_dev, int64_t vel_stride,
                            const double* params_dev, const float* s_dev, float* drho_dev,
                            double* g_dev, void* stream) {
    if (int e = tfm_check(di, dj, dk, fh, fv, rc)) return e;
    if (int e = ref_check(bfr, di, dj, dk, fh, fv)) return e;
    if (n_models < 1 || n_models > 256)
        return fail(ZTM_EINVAL, "n_models must be in [1, 256] (got %d)", n_models);
    if (!sea_dev || !moho_dev || !vel_dev || !params_dev || !s_dev || !drho_dev || !g_dev)
        return fail(ZTM_EINVAL, "null argument");
    if (moho_stride < 0 || vel_stride < 0) return fail(ZTM_EINVAL, "negative stride");
    ztm_plan p = nullptr;
    if (int e = ref_plan(bfr, di, dj, dk, fh, fv, &p)) return e;
    hipStream_t s = (hipStream_t)stream;
    if (int e = tfm_drho_batched_launch(di, dj, dk, fh, fv, n_models, rc, sea_dev, moho_dev,
                                        moho_stride, vel_dev, vel_stride, params_dev, drho_dev, s))
        return e;
    return tfm_gravity_of_fields(p, bfr, di, dj, fh, fv, n_models, drho_dev, p->prob.threshold,
                                 s_dev, g_dev, s);
}

int ztm_tfm_drho_jvp(int di, int dj, int dk, int fh, int fv, const double* rc,
                     const double* sea_dev, const double* moho_dev, const double* vel_dev,
                     double vm, double ro, double dr, float* ddrho_dev, void* stream) {
    if (int e = tfm_check(di, dj, dk, fh, fv, rc)) return e;
    if (!sea_dev || !moho_dev || !vel_dev || !ddrho_dev) return fail(ZTM_EINVAL, "null argument");
    if (!std::isfinite(vm) || !std::isfinite(ro) || !std::isfinite(dr))
        return fail(ZTM_EINVAL, "vm, ro and dr must be finite");
    if (int e = check_device(nullptr)) return e;
    return tfm_drho_jvp_launch(di, dj, dk, fh, fv, rc, sea_dev, moho_dev, vel_dev, vm, ro, dr,
                               ddrho_dev, (hipStream_t)stream);
}

int ztm_tfm_gravity_jvp(int bfr, int di, int dj, int dk, int fh, int fv, const double* rc,
                        const double* sea_dev, const double* moho_dev, const double* vel_dev,
                        double vm, double ro, double dr, const float* s_dev, float* ddrho_dev,
                        double* dg_dev, void* stream) {
    if (int e = tfm_check(di, dj, dk, fh, fv, rc)) return e;
    if (int e = ref_check(bfr, di, dj, dk, fh, fv)) return e;
    if (!sea_dev || !moho_dev || !vel_dev || !s_dev || !ddrho_dev || !dg_dev)
        return fail(ZTM_EINVAL, "null argument");
    if (!std::isfinite(vm) || !std::isfinite(ro) || !std::isfinite(dr))
        return fail(ZTM_EINVAL, "vm, ro and dr must be finite");
    ztm_plan p = nullptr;
    if (int e = ref_plan(bfr, di, dj, dk, fh, fv, &p)) return e;
    hipStream_t s = (hipStream_t)stream;
    if (int e = tfm_drho_jvp_launch(di, dj, dk, fh, fv, rc, sea_dev, moho_dev, vel_dev, vm, ro,
                                    dr, ddrho_dev, s))
        return e;
    // The fields are already masked by GF's skip at the parameter point, so
    // neither the main kernel nor the correction thresholds them again.
    return tfm_fields_gravity(p, bfr, di, dj, fh, fv, 3, ddrho_dev, 0.0, s_dev, dg_dev, s);
}

int ztm_tfm_gravity_jvp_grids(int bfr, int di, int dj, int dk, int fh, int fv, const double* rc,
                              const double* sea_dev, const double* moho_dev,
                              const double* vel_dev, double vm, double ro, double dr,
                              const float* s_dev, int n_dirs, const double* dmoho_dev,
                              const double* dvel_dev, float* ddrho_dev, double* dg_dev,
                              void* stream) {
    if (int e = tfm_check(di, dj, dk, fh, fv, rc)) return e;
    if (int e = ref_check(bfr, di, dj, dk, fh, fv)) return e;
    if (!sea_dev || !moho_dev || !vel_dev || !s_dev || !ddrho_dev || !dg_dev)
        return fail(ZTM_EINVAL, "null argument");
    if (!dmoho_dev && !dvel_dev) return fail(ZTM_EINVAL, "no tangent grid (dmoho and dvel null)");
    if (n_dirs < 1 || n_dirs > 256)
        return fail(ZTM_EINVAL, "n_dirs must be in [1, 256] (got %d)", n_dirs);
    if (!std::isfinite(vm) || !std::isfinite(ro) || !std::isfinite(dr))
        return fail(ZTM_EINVAL, "vm, ro and dr must be finite");
    ztm_plan p = nullptr;
    if (int e = ref_plan(bfr, di, dj, dk, fh, fv, &p)) return e;
    hipStream_t s = (hipStream_t)stream;
    if (int e = tfm_drho_jvp_grids_launch(di, dj, dk, fh, fv, rc, sea_dev, moho_dev, vel_dev, vm,
                                          ro, dr, n_dirs, dmoho_dev, dvel_dev, ddrho_dev, s))
        return e;
    return tfm_fields_gravity(p, bfr, di, dj, fh, fv, n_dirs, ddrho_dev, 0.0, s_dev, dg_dev, s);
}

int ztm_tfm_gravity_vjp_grids(int bfr, int di, int dj, int dk, int fh, int fv, const double* rc,
                              const double* sea_dev, const double* moho_dev,
                              const double* vel_dev, double vm, double ro, double dr,
                              const float* s_dev, const double* r_dev, double* gmoho_dev,
                              double* gvel_dev, void* stream) {
    if (int e = tfm_check(di, dj, dk, fh, fv, rc)) return e;
    if (int e = ref_check(bfr, di, dj, dk, fh, fv)) return e;
    if (!sea_dev || !moho_dev || !vel_dev || !s_dev || !r_dev)
        return fail(ZTM_EINVAL, "null argument");
    if (!gmoho_dev && !gvel_dev) return fail(ZTM_EINVAL, "no output (gmoho and gvel null)");
    if (!std::isfinite(vm) || !std::isfinite(ro) || !std::isfinite(dr))
        return fail(ZTM_EINVAL, "vm, ro and dr must be finite");
    ztm_plan p = nullptr;
    if (int e = ref_plan(bfr, di, dj, dk, fh, fv, &p)) return e;
    hipStream_t s = (hipStream_t)stream;
    const int64_t N = p->shape.n_cells;
    const int m = p->shape.m;
    const size_t M = (size_t)di * dj;
    if (m == 0) {   // no interior stations: r . G = 0
        if (gmoho_dev)
            if (int e = zero_fill(gmoho_dev, sizeof(double) * M, s)) return e;
        if (gvel_dev)
            if (int e = zero_fill(gvel_dev, sizeof(double) * M * dk, s)) return e;
        return ZTM_OK;
    }
    if (int e = prep_ref_stations(s_dev, bfr, di, dj, fh, fv, p->st_f, p->st_d, p->st_c,
                                  p->out_map, m, p->shape.m_pad, s))
        return e;
    // a = G^T r (FP32 adjoint with its near-field pass), then the transpose
    // of the tangent-field map onto the grid nodes, in FP64
    const size_t a_bytes = ((size_t)(N + m) * sizeof(float) + 255) / 256 * 256;
    const size_t need = a_bytes + tfm_grid_vjp_workspace(di, dj, dk, fh, fv);
    if (int rc = ensure_batched_workspace(p, need)) return rc;
    float* a = (float*)p->bws;
    float* r_int = a + N;
    double* ws = (double*)((char*)p->bws + a_bytes);
    if (int e = ensure_workspace(p, std::max(pm_fwd_workspace(p->shape, p->tuning, ZTM_FP32, false),
                                             pm_adj_workspace(p->shape, p->tuning, ZTM_FP32, true,
                                                              false))))
        return e;
    if (int e = tfm_gather_launch(r_dev, p->out_map, m, r_int, s)) return e;
    if (int e = pm_adj_launch(p->shape, p->tuning, ZTM_FP32, p->st_f, r_int, p->sy, p->sz,
                              pm_scale(p), p->part, a, near_args(p), s))
        return e;
    return tfm_grid_vjp_launch(di, dj, dk, fh, fv, rc, sea_dev, moho_dev, vel_dev, vm, ro, dr, a,
                               ws, gmoho_dev, gvel_dev, s);
}

int ztm_tfm_gravity_vjp(int bfr, int di, int dj, int dk, int fh, int fv, const double* rc,
                        const double* sea_dev, const double* moho_dev, const double* vel_dev,
                        double vm, double ro, double dr, const float* s_dev, const double* r_dev,
                        float* ddrho_dev, double* out_dev, void* stream) {
    if (int e = tfm_check(di, dj, dk, fh, fv, rc)) return e;
    if (int e = ref_check(bfr, di, dj, dk, fh, fv)) return e;
    if (!sea_dev || !moho_dev || !vel_dev || !s_dev || !r_dev || !ddrho_dev || !out_dev)
        return fail(ZTM_EINVAL, "null argument");
    if (!std::isfinite(vm) || !std::isfinite(ro) || !std::isfinite(dr))
        return fail(ZTM_EINVAL, "vm, ro and dr must be finite");
    ztm_plan p = nullptr;
    if (int e = ref_plan(bfr, di, dj, dk, fh, fv, &p)) return e;
    hipStream_t s = (hipStream_t)stream;
    if (int e = tfm_drho_jvp_launch(di, dj, dk, fh, fv, rc, sea_dev, moho_dev, vel_dev, vm, ro,
                                    dr, ddrho_dev, s))
        return e;
    const int64_t N = p->shape.n_cells;
    const int m = p->shape.m;
    if (m == 0) return tfm_scatter_launch(nullptr, nullptr, 0, 3, 1, out_dev, s);   // zeros
    if (int e = prep_ref_stations(s_dev, bfr, di, dj, fh, fv, p->st_f, p->st_d, p->st_c,
                                  p->out_map, m, p->shape.m_pad, s))
        return e;
    // a = G^T r on the kept-cell raster (FP32 adjoint; the producer's depth
    // filter keeps cells at least a coarse cell below every station, where
    // the FP32 pair kernel needs no near-field correction), then the three
    // FP64 dot products with the derivative fields
    const size_t need = tfm_vjp_workspace(N, m);
    if (int rc = ensure_batched_workspace(p, need)) return rc;
    float* a = (float*)p->bws;
    float* r_int = a + N;
    double* part = (double*)((char*)p->bws + ((size_t)(N + m) * sizeof(float) + 255) / 256 * 256);
    if (int e = ensure_workspace(p, std::max(pm_fwd_workspace(p->shape, p->tuning, ZTM_FP32, false),
                                             pm_adj_workspace(p->shape, p->tuning, ZTM_FP32, true,
                                                              false))))
        return e;
    if (int e = tfm_gather_launch(r_dev, p->out_map, m, r_int, s)) return e;
    if (int e = pm_adj_launch(p->shape, p->tuning, ZTM_FP32, p->st_f, r_int, p->sy, p->sz,
                              pm_scale(p), p->part, a, near_args(p), s))
        return e;
    return tfm_dot3_launch(a, ddrho_dev, N, part, out_dev, s);
}

// ----------------------------------------------------------- profiling
int ztm_profile_enable(int on) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_prof_on = on != 0;
    return ZTM_OK;
}

int ztm_profile_read(const char* kernel, double* total_ms, int64_t* count) {
    if (!kernel || !total_ms || !count) return fail(ZTM_EINVAL, "null argument");
    std::vector<ProfRec> recs;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        recs = g_prof;
    }
    double tot = 0.0;
    int64_t n = 0;
    for (auto& r : recs) {
        if (r.name != kernel) continue;
        ZTM_HIP(hipEventSynchronize(r.b));
        float ms = 0.f;
        ZTM_HIP(hipEventElapsedTime(&ms, r.a, r.b));
        tot += ms;
        ++n;
    }
    *total_ms = tot;
    *count = n;
    return ZTM_OK;
}

int ztm_profile_reset(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto& r : g_prof) {
        (void)hipEventSynchronize(r.b);
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    g_prof.clear();
    return ZTM_OK;
}

}  // extern "C"
