// Uneven batch of independent point-mass problems (FP32 path): every forward
// or adjoint pass over the whole batch is ONE persistent launch (plus its
// reduce) taking tickets from a device counter, and a regularised
// least-squares solve of every problem at once,
//     minimise 0.5 ||G_s x_s - d_s||^2 + 0.5 lambda_s ||grad x_s||^2,
// runs conjugate gradients on the normal equations
//     (G_s^T G_s + lambda_s L^T L) x_s = G_s^T d_s,   x_s(0) = 0,
// with per-problem FP64 scalars, captured once into a HIP graph and replayed.
// Tiles (stations per lane, row slices, cells per lane, station splits) are
// tuned per size class by timing the grouped kernels on the class alone.
//
// No reference counterpart: ZTM's inversion runs one forward model per
// finite-difference Jacobian column on the CPU (TFM.cpp:124-135) and has no
// batched or iterative solver.
#include "gfm_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

namespace ztm {

constexpr int kChunk = 4096;       // cells per CG workgroup (256 lanes x 16)
constexpr int kStAlign = 256;      // station offsets in the concatenated vectors
constexpr int kClasses = 6;        // size classes: 3 station bands x 2 cell bands

struct CgProb {
    int64_t off, n;    // cell offset and count
    int nx, ny, nz;
    int c0, nch;       // first chunk, chunks
};

// ----------------------------------------------------------- CG kernels
// Sum over a 256-lane workgroup in a fixed order (the same bits in every
// workgroup that sums the same values); every lane gets the result.
__device__ __forceinline__ double block_sum256(double v, double* sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    const double t = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    __syncthreads();
    return t;
}

// Sum of a problem's per-chunk partials, fixed order.
__device__ __forceinline__ double problem_sum(const double* __restrict__ part, const CgProb& q,
                                              double* sh) {
    double a = 0.0;
    for (int c = threadIdx.x; c < q.nch; c += 256) a += part[q.c0 + c];
    return block_sum256(a, sh);
}

// (L^T L v)_n with L = forward differences inside the grid (no differences
// across the boundary): deg(n) v_n - sum of the axis neighbours.
__device__ __forceinline__ float ltl_at(const float* __restrict__ v, int64_t n, int nx, int ny,
                                        int nz) {
    const int i = (int)(n % nx);
    const int64_t rw = n / nx;
    const int j = (int)(rw % ny);
    const int k = (int)(rw / ny);
    const int64_t sy = nx, sz = (int64_t)nx * ny;
    const float c = v[n];
    float acc = 0.0f;
    if (i > 0) acc += c - v[n - 1];
    if (i < nx - 1) acc += c - v[n + 1];
    if (j > 0) acc += c - v[n - sy];
    if (j < ny - 1) acc += c - v[n + sy];
    if (k > 0) acc += c - v[n - sz];
    if (k < nz - 1) acc += c - v[n + sz];
    return acc;
}

// x = 0, p = r (r holds G^T d), partial sums of r.r per chunk.
__global__ __launch_bounds__(256) void cg_init_kernel(float* __restrict__ x,
                                                      const float* __restrict__ r,
                                                      float* __restrict__ p,
                                                      const CgProb* __restrict__ pr,
                                                      const int* __restrict__ chunk_prob,
                                                      double* __restrict__ part) {
    __shared__ double sh[4];
    const int c = blockIdx.x;
    const CgProb q = pr[chunk_prob[c]];
    const int64_t base = (int64_t)(c - q.c0) * kChunk;
    double acc = 0.0;
    for (int e = 0; e < kChunk / 256; ++e) {
        const int64_t n = base + e * 256 + threadIdx.x;
        if (n < q.n) {
            const float v = r[q.off + n];
            x[q.off + n] = 0.0f;
            p[q.off + n] = v;
            acc += (double)v * v;
        }
    }
    const double t = block_sum256(acc, sh);
    if (threadIdx.x == 0) part[c] = t;
}

// rr[s] = sum of problem s's partials.
__global__ __launch_bounds__(256) void cg_sum_kernel(const double* __restrict__ part,
                                                     const CgProb* __restrict__ pr,
                                                     double* __restrict__ rr) {
    __shared__ double sh[4];
    const CgProb q = pr[blockIdx.x];
    const double t = problem_sum(part, q, sh);
    if (threadIdx.x == 0) rr[blockIdx.x] = t;
}

// Ap += lambda L^T L p (Ap holds G^T G p), partial sums of p.Ap per chunk.
__global__ __launch_bounds__(256) void cg_ltl_dot_kernel(const float* __restrict__ p,
                                                         float* __restrict__ ap,
                                                         const CgProb* __restrict__ pr,
                                                         const int* __restrict__ chunk_prob,
                                                         const double* __restrict__ lam,
                                                         double* __restrict__ part) {
    __shared__ double sh[4];
    const int c = blockIdx.x;
    const int s = chunk_prob[c];
    const CgProb q = pr[s];
    const double l = lam[s];
    const float* pp = p + q.off;
    float* aq = ap + q.off;
    const int64_t base = (int64_t)(c - q.c0) * kChunk;
    double acc = 0.0;
    for (int e = 0; e < kChunk / 256; ++e) {
        const int64_t n = base + e * 256 + threadIdx.x;
        if (n < q.n) {
            const float v = (float)((double)aq[n] + l * (double)ltl_at(pp, n, q.nx, q.ny, q.nz));
            aq[n] = v;
            acc += (double)pp[n] * v;
        }
    }
    const double t = block_sum256(acc, sh);
    if (threadIdx.x == 0) part[c] = t;
}

// alpha = rr_k / p.Ap;  x += alpha p;  r -= alpha Ap;  partial sums of r.r.
__global__ __launch_bounds__(256) void cg_update_xr_kernel(
    float* __restrict__ x, float* __restrict__ r, const float* __restrict__ p,
    const float* __restrict__ ap, const CgProb* __restrict__ pr,
    const int* __restrict__ chunk_prob, const double* __restrict__ rr_k,
    const double* __restrict__ part_pap, double* __restrict__ part_rr) {
    __shared__ double sh[4];
    const int c = blockIdx.x;
    const int s = chunk_prob[c];
    const CgProb q = pr[s];
    const double pap = problem_sum(part_pap, q, sh);
    const double rk = rr_k[s];
    const float alpha = (pap > 0.0 && rk > 0.0) ? (float)(rk / pap) : 0.0f;
    const int64_t base = (int64_t)(c - q.c0) * kChunk;
    double acc = 0.0;
    for (int e = 0; e < kChunk / 256; ++e) {
        const int64_t n = q.off + base + e * 256 + threadIdx.x;
        if (n < q.off + q.n) {
            x[n] += alpha * p[n];
            const float v = r[n] - alpha * ap[n];
            r[n] = v;
            acc += (double)v * v;
        }
    }
    const double t = block_sum256(acc, sh);
    if (threadIdx.x == 0) part_rr[c] = t;
}

// rr_{k+1} = sum of the r.r partials;  beta = rr_{k+1} / rr_k;  p = r + beta p.
__global__ __launch_bounds__(256) void cg_update_p_kernel(
    const float* __restrict__ r, float* __restrict__ p, const CgProb* __restrict__ pr,
    const int* __restrict__ chunk_prob, const double* __restrict__ rr_k,
    double* __restrict__ rr_k1, const double* __restrict__ part_rr) {
    __shared__ double sh[4];
    const int c = blockIdx.x;
    const int s = chunk_prob[c];
    const CgProb q = pr[s];
    const double rn = problem_sum(part_rr, q, sh);
    const double rk = rr_k[s];
    const float beta = rk > 0.0 ? (float)(rn / rk) : 0.0f;
    const int64_t base = (int64_t)(c - q.c0) * kChunk;
    for (int e = 0; e < kChunk / 256; ++e) {
        const int64_t n = q.off + base + e * 256 + threadIdx.x;
        if (n < q.off + q.n) p[n] = r[n] + beta * p[n];
    }
    if (c == q.c0 && threadIdx.x == 0) rr_k1[s] = rn;
}

namespace {

// Device copy of one grouped pass (descriptors in ticket order + prefixes).
struct PassBuf {
    GroupedPass gp;
    void* desc = nullptr;
    int64_t* starts = nullptr;   // blk_start[np+1], item_start[np+1], near_start[np+1]
    void release() {
        if (desc) (void)hipFree(desc);
        if (starts) (void)hipFree(starts);
        desc = nullptr;
        starts = nullptr;
        gp = GroupedPass{};
    }
};

int size_class(int64_t m, int64_t n) {
    const int sb = m < 4096 ? 0 : (m < 16384 ? 1 : 2);
    return sb * 2 + (n < 200000 ? 0 : 1);
}

template <typename T>
int upload_vec(T** dst, const std::vector<T>& v) {
    if (*dst) (void)hipFree(*dst);
    *dst = nullptr;
    if (v.empty()) return ZTM_OK;
    ZTM_HIP(hipMalloc(dst, v.size() * sizeof(T)));
    ZTM_HIP(hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return ZTM_OK;
}

}  // namespace
}  // namespace ztm

struct ztm_batch_s {
    int device = 0;
    int np = 0;
    std::vector<ztm_plan> plans;
    std::vector<int64_t> cell_off, st_off;   // np + 1 each (aligned starts)
    std::vector<int> cls;
    ztm::PassBuf fwd, adj;
    unsigned* ticket = nullptr;              // [0] forward, [1] adjoint
    // CG state
    ztm::CgProb* cgp = nullptr;
    int* chunk_prob = nullptr;
    int n_chunks = 0;
    float *r = nullptr, *p = nullptr, *ap = nullptr, *qf = nullptr;
    double *part_a = nullptr, *part_b = nullptr, *lam = nullptr, *rr = nullptr;
    double* lam_host = nullptr;          // pinned staging of lambda
    hipEvent_t lam_copied = nullptr;     // the last staging copy has run
    bool lam_recorded = false;
    int rr_cap = 0;   // entries
    hipStream_t cap = nullptr;
    hipGraphExec_t exec = nullptr;
    const void* key[4] = {nullptr, nullptr, nullptr, nullptr};
    int key_iters = -1;
};

namespace ztm {
namespace {

void drop_graph(ztm_batch b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
    b->key_iters = -1;
}

void batch_free(ztm_batch b) {
    if (!b) return;
    drop_graph(b);
    b->fwd.release();
    b->adj.release();
    for (void* q : {(void*)b->ticket, (void*)b->cgp, (void*)b->chunk_prob, (void*)b->r,
                    (void*)b->p, (void*)b->ap, (void*)b->qf, (void*)b->part_a, (void*)b->part_b,
                    (void*)b->lam, (void*)b->rr})
        if (q) (void)hipFree(q);
    if (b->cap) (void)hipStreamDestroy(b->cap);
    if (b->lam_copied) (void)hipEventDestroy(b->lam_copied);
    if (b->lam_host) (void)hipHostFree(b->lam_host);
    for (ztm_plan p : b->plans) plan_free(p);
    delete b;
}

// Build the forward (fwd = true) or adjoint pass over `members`, ordered by
// decreasing per-block cost.
int build_pass(ztm_batch b, const std::vector<int>& members, bool fwd, PassBuf& pb) {
    pb.release();
    const int np = (int)members.size();
    std::vector<GFwdDesc> fd(np);
    std::vector<GAdjDesc> ad(np);
    std::vector<int64_t> blocks(np), items(np), nears(np, 0);
    std::vector<double> cost(np);
    size_t lds = 0;
    int maxv = 0;
    for (int a = 0; a < np; ++a) {
        const int s = members[a];
        ztm_plan p = b->plans[s];
        if (int rc = pm_grouped_workspace(p)) return rc;
        size_t l = 0;
        if (fwd) {
            if (int rc = pm_fwd_desc(p, b->cell_off[s], b->st_off[s], &fd[a], &l, &blocks[a]))
                return rc;
            items[a] = p->shape.m;
            cost[a] = (double)fd[a].rps * fd[a].nx * 64.0 * fd[a].S;
            maxv = std::max(maxv, fd[a].S);
        } else {
            if (int rc = pm_adj_desc(p, b->cell_off[s], b->st_off[s], &ad[a], &l, &blocks[a]))
                return rc;
            items[a] = p->shape.n_cells;
            nears[a] = ad[a].n_near;
            cost[a] = 64.0 * ad[a].C * ad[a].st_per_split;
            maxv = std::max(maxv, ad[a].C);
        }
        lds = std::max(lds, l);
    }
    std::vector<int> order(np);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[x] > cost[y]; });
    std::vector<int64_t> starts(3 * (np + 1), 0);
    std::vector<GFwdDesc> fo;
    std::vector<GAdjDesc> ao;
    for (int a = 0; a < np; ++a) {
        const int o = order[a];
        starts[a + 1] = starts[a] + blocks[o];
        starts[np + 1 + a + 1] = starts[np + 1 + a] + items[o];
        starts[2 * (np + 1) + a + 1] = starts[2 * (np + 1) + a] + nears[o];
        if (fwd) fo.push_back(fd[o]); else ao.push_back(ad[o]);
    }
    const size_t desc_bytes = fwd ? np * sizeof(GFwdDesc) : np * sizeof(GAdjDesc);
    if (np > 0) {
        ZTM_HIP(hipMalloc(&pb.desc, desc_bytes));
        ZTM_HIP(hipMemcpy(pb.desc, fwd ? (const void*)fo.data() : (const void*)ao.data(),
                          desc_bytes, hipMemcpyHostToDevice));
    }
    if (int rc = upload_vec(&pb.starts, starts)) return rc;
    pb.gp.desc = pb.desc;
    pb.gp.blk_start = pb.starts;
    pb.gp.item_start = pb.starts + (np + 1);
    pb.gp.near_start = pb.starts + 2 * (np + 1);
    pb.gp.ticket = b->ticket + (fwd ? 0 : 1);
    pb.gp.np = np;
    pb.gp.max_variant = maxv;
    pb.gp.total_blocks = starts[np];
    pb.gp.total_items = starts[2 * np + 1];
    pb.gp.total_near = starts[3 * np + 2];
    pb.gp.lds = lds;
    if (lds > 65536) return fail(ZTM_EINVAL, "grouped pass needs %zu bytes of LDS", lds);
    return ZTM_OK;
}

int rebuild(ztm_batch b) {
    drop_graph(b);
    std::vector<int> all(b->np);
    std::iota(all.begin(), all.end(), 0);
    if (int rc = build_pass(b, all, true, b->fwd)) return rc;
    return build_pass(b, all, false, b->adj);
}

int enqueue_solve(ztm_batch b, const float* d, int iters, float* x, double* rr,
                  hipStream_t s) {
    const int np = b->np;
    const dim3 cg(b->n_chunks), blk(256);
    if (int rc = pm_adj_grouped_launch(b->adj.gp, d, b->r, s)) return rc;   // r = G^T d
    hipLaunchKernelGGL(cg_init_kernel, cg, blk, 0, s, x, b->r, b->p, b->cgp, b->chunk_prob,
                       b->part_b);
    hipLaunchKernelGGL(cg_sum_kernel, dim3(np), blk, 0, s, b->part_b, b->cgp, rr);
    ZTM_HIP(hipGetLastError());
    for (int k = 0; k < iters; ++k) {
        double* rk = rr + (size_t)k * np;
        if (int rc = pm_fwd_grouped_launch(b->fwd.gp, b->p, nullptr, b->qf, s)) return rc;
        if (int rc = pm_adj_grouped_launch(b->adj.gp, b->qf, b->ap, s)) return rc;
        hipLaunchKernelGGL(cg_ltl_dot_kernel, cg, blk, 0, s, b->p, b->ap, b->cgp, b->chunk_prob,
                           b->lam, b->part_a);
        hipLaunchKernelGGL(cg_update_xr_kernel, cg, blk, 0, s, x, b->r, b->p, b->ap, b->cgp,
                           b->chunk_prob, rk, b->part_a, b->part_b);
        hipLaunchKernelGGL(cg_update_p_kernel, cg, blk, 0, s, b->r, b->p, b->cgp, b->chunk_prob,
                           rk, rk + np, b->part_b);
        ZTM_HIP(hipGetLastError());
    }
    return ZTM_OK;
}

// Time `reps` forward (or adjoint) grouped passes over one class alone.
int time_pass(ztm_batch b, const std::vector<int>& members, bool fwd, int reps, hipStream_t s,
              float* ms) {
    PassBuf pb;
    int rc = build_pass(b, members, fwd, pb);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (rc == ZTM_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = fail(ZTM_EHIP, "hipEventCreate failed");
    for (int i = 0; rc == ZTM_OK && i <= reps; ++i) {
        if (i == 1) (void)hipEventRecord(e0, s);
        rc = fwd ? pm_fwd_grouped_launch(pb.gp, b->p, nullptr, b->qf, s)
                 : pm_adj_grouped_launch(pb.gp, b->qf, b->ap, s);
    }
    if (rc == ZTM_OK) {
        (void)hipEventRecord(e1, s);
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(ms, e0, e1) != hipSuccess)
            rc = fail(ZTM_EHIP, "timing events failed");
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    pb.release();
    return rc;
}

}  // namespace
}  // namespace ztm

using namespace ztm;

extern "C" {

int ztm_batch_create(ztm_batch* batch, int n_problems, const ztm_problem* problems,
                     const int64_t* n_stations, const double* const* stations_host) {
    if (!batch) return fail(ZTM_EINVAL, "null batch pointer");
    *batch = nullptr;
    if (n_problems <= 0 || !problems || !n_stations || !stations_host)
        return fail(ZTM_EINVAL, "batch needs n_problems > 0 and problem, station arrays");
    for (int s = 0; s < n_problems; ++s) {
        if (problems[s].formula != ZTM_POINT_MASS)
            return fail(ZTM_EINVAL, "problem %d: batches support the point-mass formula only", s);
        if (n_stations[s] <= 0) return fail(ZTM_EINVAL, "problem %d has no stations", s);
    }
    ztm_batch b = new ztm_batch_s;
    b->np = n_problems;
    int rc = hipGetDevice(&b->device) == hipSuccess ? ZTM_OK : fail(ZTM_EHIP, "hipGetDevice");
    b->cell_off.assign(n_problems + 1, 0);
    b->st_off.assign(n_problems + 1, 0);
    std::vector<CgProb> cgp(n_problems);
    std::vector<int> chunk_prob;
    for (int s = 0; rc == ZTM_OK && s < n_problems; ++s) {
        ztm_plan p = nullptr;
        rc = ztm_plan_create(&p, &problems[s], n_stations[s], stations_host[s]);
        if (rc) {
            rc = fail(rc, "problem %d: %s", s, ztm_last_error());
            break;
        }
        b->plans.push_back(p);
        const int64_t n = p->shape.n_cells, m = p->shape.m;
        b->cls.push_back(size_class(m, n));
        const int64_t nch = (n + kChunk - 1) / kChunk;
        b->cell_off[s + 1] = b->cell_off[s] + nch * kChunk;
        b->st_off[s + 1] = b->st_off[s] + (m + kStAlign - 1) / kStAlign * kStAlign;
        cgp[s] = CgProb{b->cell_off[s], n, p->shape.nx, p->shape.ny, p->shape.nz,
                        (int)(b->cell_off[s] / kChunk), (int)nch};
        chunk_prob.insert(chunk_prob.end(), (size_t)nch, s);
    }
    if (rc == ZTM_OK && b->cell_off[n_problems] / kChunk > INT32_MAX)
        rc = fail(ZTM_EINVAL, "batch too large");
    const size_t nc = rc == ZTM_OK ? (size_t)b->cell_off[n_problems] : 0;
    const size_t ns = rc == ZTM_OK ? (size_t)b->st_off[n_problems] : 0;
    b->n_chunks = (int)chunk_prob.size();
    auto dmalloc = [&](auto** ptr, size_t count) {
        if (rc == ZTM_OK) {
            const hipError_t e = hipMalloc(ptr, count * sizeof(**ptr));
            if (e != hipSuccess) rc = hip_fail(e, "hipMalloc");
        }
    };
    dmalloc(&b->ticket, 2);
    dmalloc(&b->r, nc);
    dmalloc(&b->p, nc);
    dmalloc(&b->ap, nc);
    dmalloc(&b->qf, ns);
    dmalloc(&b->part_a, (size_t)b->n_chunks);
    dmalloc(&b->part_b, (size_t)b->n_chunks);
    dmalloc(&b->lam, (size_t)n_problems);
    if (rc == ZTM_OK && hipMemset(b->ticket, 0, 2 * sizeof(unsigned)) != hipSuccess)
        rc = fail(ZTM_EHIP, "hipMemset failed");
    if (rc == ZTM_OK && hipMemset(b->p, 0, nc * sizeof(float)) != hipSuccess)
        rc = fail(ZTM_EHIP, "hipMemset failed");
    if (rc == ZTM_OK && hipMemset(b->qf, 0, ns * sizeof(float)) != hipSuccess)
        rc = fail(ZTM_EHIP, "hipMemset failed");
    if (rc == ZTM_OK) rc = upload_vec(&b->cgp, cgp);
    if (rc == ZTM_OK) rc = upload_vec(&b->chunk_prob, chunk_prob);
    if (rc == ZTM_OK && hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking) != hipSuccess)
        rc = fail(ZTM_EHIP, "hipStreamCreate failed");
    if (rc == ZTM_OK && hipHostMalloc(&b->lam_host, sizeof(double) * n_problems) != hipSuccess)
        rc = fail(ZTM_ENOMEM, "hipHostMalloc failed");
    if (rc == ZTM_OK &&
        hipEventCreateWithFlags(&b->lam_copied, hipEventDisableTiming) != hipSuccess)
        rc = fail(ZTM_EHIP, "hipEventCreate failed");
    if (rc == ZTM_OK) rc = rebuild(b);
    if (rc) {
        batch_free(b);
        return rc;
    }
    *batch = b;
    return ZTM_OK;
}

int ztm_batch_destroy(ztm_batch batch) {
    batch_free(batch);
    return ZTM_OK;
}

int ztm_batch_num_problems(ztm_batch batch) { return batch ? batch->np : 0; }

int ztm_batch_layout(ztm_batch batch, int64_t* cell_offsets, int64_t* station_offsets) {
    if (!batch) return fail(ZTM_EINVAL, "null batch");
    if (cell_offsets) std::memcpy(cell_offsets, batch->cell_off.data(), (batch->np + 1) * 8);
    if (station_offsets) std::memcpy(station_offsets, batch->st_off.data(), (batch->np + 1) * 8);
    return ZTM_OK;
}

int ztm_batch_forward(ztm_batch batch, const float* rho_dev, double* g_dev, void* stream) {
    if (!batch || !rho_dev || !g_dev) return fail(ZTM_EINVAL, "null batch, rho or g");
    return pm_fwd_grouped_launch(batch->fwd.gp, rho_dev, g_dev, nullptr, (hipStream_t)stream);
}

int ztm_batch_adjoint(ztm_batch batch, const float* r_dev, float* out_dev, void* stream) {
    if (!batch || !r_dev || !out_dev) return fail(ZTM_EINVAL, "null batch, r or out");
    if ((const void*)r_dev == (const void*)out_dev) return fail(ZTM_EINVAL, "r and out alias");
    return pm_adj_grouped_launch(batch->adj.gp, r_dev, out_dev, (hipStream_t)stream);
}

int ztm_batch_solve(ztm_batch batch, const float* d_dev, const double* lambda, int iters,
                    float* x_dev, double* rr_dev, int use_graph, void* stream) {
    if (!batch || !d_dev || !lambda || !x_dev)
        return fail(ZTM_EINVAL, "null batch, data, lambda or x");
    if (iters < 0 || iters > 10000) return fail(ZTM_EINVAL, "iters=%d out of range", iters);
    ztm_batch b = batch;
    for (int s = 0; s < b->np; ++s)
        if (!(lambda[s] >= 0.0) || !std::isfinite(lambda[s]))
            return fail(ZTM_EINVAL, "lambda[%d] must be finite and >= 0", s);
    hipStream_t s = (hipStream_t)stream;
    double* rr = rr_dev;
    const int need = (iters + 1) * b->np;
    if (!rr) {
        if (b->rr_cap < need) {
            drop_graph(b);
            if (b->rr) (void)hipFree(b->rr);
            b->rr = nullptr;
            b->rr_cap = 0;
            ZTM_HIP(hipMalloc(&b->rr, (size_t)need * sizeof(double)));
            b->rr_cap = need;
        }
        rr = b->rr;
    }
    // Stage lambda in pinned memory (the caller's array may change as soon as
    // this returns); the previous staging copy must have run before reuse.
    if (b->lam_recorded) ZTM_HIP(hipEventSynchronize(b->lam_copied));
    std::memcpy(b->lam_host, lambda, b->np * sizeof(double));
    ZTM_HIP(hipMemcpyAsync(b->lam, b->lam_host, b->np * sizeof(double), hipMemcpyHostToDevice, s));
    ZTM_HIP(hipEventRecord(b->lam_copied, s));
    b->lam_recorded = true;
    if (!use_graph) return enqueue_solve(b, d_dev, iters, x_dev, rr, s);
    const void* key[4] = {d_dev, x_dev, rr, nullptr};
    if (!b->exec || b->key_iters != iters || std::memcmp(key, b->key, sizeof key) != 0) {
        drop_graph(b);
        ZTM_HIP(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_solve(b, d_dev, iters, x_dev, rr, b->cap);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b->cap, &g);
        if (rc == ZTM_OK && e != hipSuccess) rc = hip_fail(e, "hipStreamEndCapture");
        if (rc == ZTM_OK) {
            const hipError_t ei = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
            if (ei != hipSuccess) {
                b->exec = nullptr;
                rc = hip_fail(ei, "hipGraphInstantiate");
            }
        }
        if (g) (void)hipGraphDestroy(g);
        if (rc) return rc;
        std::memcpy(b->key, key, sizeof key);
        b->key_iters = iters;
    }
    ZTM_HIP(hipGraphLaunch(b->exec, s));
    return ZTM_OK;
}

int ztm_batch_set_tuning(ztm_batch batch, int problem, int st_per_lane, int rows_per_slice,
                         int cells_per_lane, int adj_splits) {
    if (!batch || problem < -1 || problem >= batch->np)
        return fail(ZTM_EINVAL, "bad batch or problem index");
    if (st_per_lane != 2 && st_per_lane != 4 && st_per_lane != 8)
        return fail(ZTM_EINVAL, "st_per_lane must be 2, 4 or 8");
    if (cells_per_lane != 8 && cells_per_lane != 16)
        return fail(ZTM_EINVAL, "cells_per_lane must be 8 or 16");
    if (rows_per_slice < 0 || adj_splits < 0) return fail(ZTM_EINVAL, "negative tuning value");
    for (int s = 0; s < batch->np; ++s) {
        if (problem >= 0 && s != problem) continue;
        PmTuning& t = batch->plans[s]->tuning;
        t.st_per_lane = st_per_lane;
        t.rows_per_slice = rows_per_slice;
        t.cells_per_lane = cells_per_lane;
        t.adj_splits = adj_splits;
    }
    return rebuild(batch);
}

int ztm_batch_get_tuning(ztm_batch batch, int problem, int* st_per_lane, int* rows_per_slice,
                         int* cells_per_lane, int* adj_splits, int* size_class_out) {
    if (!batch || problem < 0 || problem >= batch->np)
        return fail(ZTM_EINVAL, "bad batch or problem index");
    ztm_plan p = batch->plans[problem];
    const PmTuning t = effective_tuning(p, ZTM_FP32);
    if (st_per_lane)   // the stations per lane the forward runs with (wide rows: 8)
        *st_per_lane = fwd_adjust(p->shape, t, sizeof(float), false).st_per_lane;
    if (rows_per_slice) *rows_per_slice = fwd_rows_per_slice(p->shape, t, sizeof(float));
    if (cells_per_lane) *cells_per_lane = t.cells_per_lane;
    if (adj_splits) *adj_splits = ztm::adj_splits(p->shape, t, sizeof(float), false);
    if (size_class_out) *size_class_out = batch->cls[problem];
    return ZTM_OK;
}

// Per size class, time the grouped forward over stations-per-lane x row-slice
// candidates and the grouped adjoint over cells-per-lane x split candidates,
// on the class's problems alone; keep the fastest of each.
int ztm_batch_autotune(ztm_batch batch, int reps, void* stream) {
    if (!batch || reps <= 0) return fail(ZTM_EINVAL, "bad batch or reps");
    ztm_batch b = batch;
    hipStream_t s = (hipStream_t)stream;
    drop_graph(b);
    // time on non-zero inputs (1.0f everywhere, real cells and padding)
    ZTM_HIP(hipMemsetD32Async((hipDeviceptr_t)b->p, 0x3f800000, (size_t)b->cell_off[b->np], s));
    ZTM_HIP(hipMemsetD32Async((hipDeviceptr_t)b->qf, 0x3f800000, (size_t)b->st_off[b->np], s));
    for (int c = 0; c < kClasses; ++c) {
        std::vector<int> mem;
        for (int i = 0; i < b->np; ++i)
            if (b->cls[i] == c) mem.push_back(i);
        if (mem.empty()) continue;
        // forward: rows_per_slice as a multiple of each problem's automatic choice
        float best = INFINITY;
        int best_s = 4;
        double best_f = 1.0;
        for (int S : {2, 4, 8}) {
            for (double f : {0.5, 1.0, 2.0}) {
                for (int i : mem) {
                    PmTuning& t = b->plans[i]->tuning;
                    t.st_per_lane = S;
                    t.rows_per_slice = 0;
                    const int auto_rps = fwd_rows_per_slice(b->plans[i]->shape, t, sizeof(float));
                    t.rows_per_slice = std::max(1, (int)std::lround(auto_rps * f));
                }
                float ms = 0.0f;
                if (int rc = time_pass(b, mem, true, reps, s, &ms)) return rc;
                if (ms < best) {
                    best = ms;
                    best_s = S;
                    best_f = f;
                }
            }
        }
        for (int i : mem) {
            PmTuning& t = b->plans[i]->tuning;
            t.st_per_lane = best_s;
            t.rows_per_slice = 0;
            const int auto_rps = fwd_rows_per_slice(b->plans[i]->shape, t, sizeof(float));
            t.rows_per_slice = std::max(1, (int)std::lround(auto_rps * best_f));
        }
        // adjoint: station splits as a multiple of the automatic choice
        best = INFINITY;
        int best_c = 16;
        best_f = 1.0;
        for (int C : {8, 16}) {
            for (double f : {0.5, 1.0, 2.0}) {
                for (int i : mem) {
                    PmTuning& t = b->plans[i]->tuning;
                    t.cells_per_lane = C;
                    t.adj_splits = 0;
                    const int auto_sp = ztm::adj_splits(b->plans[i]->shape, t, sizeof(float), false);
                    t.adj_splits = std::max(1, (int)std::lround(auto_sp * f));
                }
                float ms = 0.0f;
                if (int rc = time_pass(b, mem, false, reps, s, &ms)) return rc;
                if (ms < best) {
                    best = ms;
                    best_c = C;
                    best_f = f;
                }
            }
        }
        for (int i : mem) {
            PmTuning& t = b->plans[i]->tuning;
            t.cells_per_lane = best_c;
            t.adj_splits = 0;
            const int auto_sp = ztm::adj_splits(b->plans[i]->shape, t, sizeof(float), false);
            t.adj_splits = std::max(1, (int)std::lround(auto_sp * best_f));
        }
    }
    return rebuild(b);
}

}  // extern "C"
