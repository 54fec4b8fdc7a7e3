// ztm_gf_noalloc: the reference-kernel call (GF, GFM_GPU.cu:16-22; the same
// kernels and results as ztm_gf_async) for callers that cannot allocate: code
// under HIP graph capture, tight loops, harnesses that forbid hipMalloc.
//
// ztm_gf_async keeps a cached plan per geometry whose station records and
// split-K partials are hipMalloc'd on the first call. Here the same buffers
// are carved out of `scratch`, a fixed device array in this library's code
// object (one copy per GPU, held until the process ends), and the plan's
// shape lives on the caller's stack, so no call allocates device or host
// memory. Every call on a GPU shares that GPU's array: calls that overlap in
// time on one GPU (two streams, two threads) race on it and must be
// serialised by the caller. Nothing in the array needs clearing: each call
// writes every station record and every partial it reads before reading it.
#include "gfm_internal.h"

#include <mutex>

namespace ztm {

namespace {
constexpr size_t kScratchBytes = (size_t)16 << 20;   // 16 MiB per GPU
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
char* g_scratch[kMaxDevices];   // device address of `scratch` per device, once set up

size_t align256(size_t b) { return (b + 255) / 256 * 256; }
}  // namespace

// Looks up this device's copy of `scratch` once, with `dev` current. The
// symbol's name in host code is only a host-side handle: kernels get the
// address returned here.
int noalloc_setup(int dev, char** out) {
    if (dev < 0 || dev >= kMaxDevices)
        return fail(ZTM_ENODEV, "device %d is beyond the %d devices the no-allocation path tracks",
                    dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!ptr) return fail(ZTM_EHIP, "hipGetSymbolAddress(scratch) returned null");
        g_scratch[dev] = (char*)ptr;
    }
    if (out) *out = g_scratch[dev];
    return ZTM_OK;
}

}  // namespace ztm

using namespace ztm;

extern "C" int ztm_gf_noalloc(int bfr, int di, int dj, int dk, int fh, int fv, float* g_dev,
                              const float* rho_dev, const float* s_dev, void* stream) {
    if (int rc = ref_check(bfr, di, dj, dk, fh, fv)) return rc;
    if (!g_dev || !rho_dev || !s_dev) return fail(ZTM_EINVAL, "null g, rho or s");
    int dev = 0;
    ZTM_HIP(hipGetDevice(&dev));
    char* base = nullptr;
    if (int rc = noalloc_setup(dev, &base)) return rc;
    ztm_plan_s p;   // shape, tuning and near-field radii only; no buffers
    if (int rc = ref_plan_init(&p, dev, bfr, di, dj, dk, fh, fv)) return rc;
    if (p.shape.m == 0) return ZTM_OK;   // no interior stations
    const size_t mp = (size_t)p.shape.m_pad;
    const size_t part_b = align256(sizeof(double) * pm_fwd_workspace(p.shape, p.tuning, ZTM_FP32,
                                                                     false));
    const size_t stf_b = align256(sizeof(float4) * mp), std_b = align256(sizeof(double4) * mp);
    const size_t stc_b = align256(sizeof(int4) * mp), map_b = align256(sizeof(int) * mp);
    const size_t need = part_b + stf_b + std_b + stc_b + map_b;
    if (need > kScratchBytes)
        return fail(ZTM_EINVAL,
                    "ztm_gf_noalloc: %d interior stations over %lld cells need %zu bytes of "
                    "scratch; the fixed maximum is %zu bytes (use ztm_gf_async)",
                    p.shape.m, (long long)p.shape.n_cells, need, kScratchBytes);
    p.part = (double*)base;
    p.st_f = (float4*)(base + part_b);
    p.st_d = (double4*)(base + part_b + stf_b);
    p.st_c = (int4*)(base + part_b + stf_b + std_b);
    p.out_map = (int*)(base + part_b + stf_b + std_b + stc_b);
    ProfTimingOff no_events;   // timing events are runtime objects: none here
    hipStream_t s = (hipStream_t)stream;
    if (int rc = prep_ref_stations(s_dev, bfr, di, dj, fh, fv, p.st_f, p.st_d, p.st_c,
                                   p.out_map, p.shape.m, p.shape.m_pad, s))
        return rc;
    return pm_fwd_launch(p.shape, p.tuning, ZTM_FP32, p.st_f, rho_dev, p.sy, p.sz,
                         p.prob.threshold, pm_scale(&p), p.part, nullptr, g_dev, p.out_map,
                         near_args(&p), s);
}
