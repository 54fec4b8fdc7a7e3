// MFMA pipe calibration: the rate this chip sustains for one MFMA shape when
// nothing else runs -- each wave streams MFMAs on register operands into
// independent accumulators, no VALU, LDS or memory traffic in the loop --
// with operands of real entropy (uniform random mantissas) or of low entropy
// (values 1 and -0.5). Under MI355X's power limit the sustained rate depends
// on the operands, so this, and not the published peak alone, is the ceiling
// a data-dependent MFMA kernel is measured against. Set-up/measurement call:
// it allocates and frees its own buffers and synchronises the stream.
#include "gfm_internal.h"

#include <cstring>
#include <type_traits>
#include <vector>

namespace ztm {
namespace {

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kOps = 4096;   // operand slots (16 bytes each) shared by all lanes

// op 0: v_mfma_f32_32x32x16_f16, 8 accumulators; op 1: _bf16, 8; op 2:
// v_mfma_f32_16x16x32_f16, 32 accumulators (the same work per iteration as 8
// of the 32x32 shape).
template <int OP>
__global__ __launch_bounds__(256) void calib_kernel(const uint4* __restrict__ ops,
                                                    float* __restrict__ out, int iters) {
    const int gid = blockIdx.x * 256 + threadIdx.x;
    const uint4 r0 = ops[(gid * 4 + 0) % kOps], r1 = ops[(gid * 4 + 1) % kOps];
    const uint4 r2 = ops[(gid * 4 + 2) % kOps], r3 = ops[(gid * 4 + 3) % kOps];
    float s = 0.f;
    if constexpr (OP == 0 || OP == 1) {
        using V = std::conditional_t<OP == 0, f16x8, bf16x8>;
        const V a0 = __builtin_bit_cast(V, r0), a1 = __builtin_bit_cast(V, r1);
        const V b0 = __builtin_bit_cast(V, r2), b1 = __builtin_bit_cast(V, r3);
        f32x16 acc[8];
        for (int t = 0; t < 8; ++t)
            for (int i = 0; i < 16; ++i) acc[t][i] = 0.f;
        for (int it = 0; it < iters; ++it) {
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                if constexpr (OP == 0)
                    acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_f16((t & 1) ? a1 : a0,
                                                                    (t & 2) ? b1 : b0, acc[t], 0,
                                                                    0, 0);
                else
                    acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16((t & 1) ? a1 : a0,
                                                                     (t & 2) ? b1 : b0, acc[t],
                                                                     0, 0, 0);
            }
        }
        for (int t = 0; t < 8; ++t)
            for (int i = 0; i < 16; ++i) s += acc[t][i];
    } else {   // OP == 2
        const f16x8 a0 = __builtin_bit_cast(f16x8, r0), a1 = __builtin_bit_cast(f16x8, r1);
        const f16x8 b0 = __builtin_bit_cast(f16x8, r2), b1 = __builtin_bit_cast(f16x8, r3);
        f32x4 acc[32];
        for (int t = 0; t < 32; ++t)
            for (int i = 0; i < 4; ++i) acc[t][i] = 0.f;
        for (int it = 0; it < iters; ++it) {
#pragma unroll
            for (int t = 0; t < 32; ++t)
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16((t & 1) ? a1 : a0,
                                                                (t & 2) ? b1 : b0, acc[t], 0, 0,
                                                                0);
        }
        for (int t = 0; t < 32; ++t)
            for (int i = 0; i < 4; ++i) s += acc[t][i];
    }
    out[gid] = s;   // keeps the accumulators live
}

// 32-bit words of two 16-bit values each, FP16 or (op 1) BF16: uniform
// random in [-0.5, 0.5), or (low) each 1 or -0.5.
void calib_operands(int op, int low, std::vector<uint32_t>& w) {
    uint32_t x = 0x9E3779B9u;
    auto next = [&]() { return x = x * 1664525u + 1013904223u; };
    for (auto& u : w) {
        uint16_t h[2];
        for (auto& hh : h) {
            const float v = low ? ((next() >> 20) & 1 ? 1.0f : -0.5f)
                                : (float)((next() >> 8) & 0xFFFF) / 65536.f - 0.5f;
            if (op == 1) {
                const __bf16 b = (__bf16)v;
                memcpy(&hh, &b, 2);
            } else {
                const _Float16 f = (_Float16)v;
                memcpy(&hh, &f, 2);
            }
        }
        u = (uint32_t)h[0] | ((uint32_t)h[1] << 16);
    }
}

}  // namespace
}  // namespace ztm

extern "C" int ztm_calib_mfma(int op, int low_entropy, int iters, double* tflops, void* stream) {
    using namespace ztm;
    if (op < 0 || op > 2)
        return fail(ZTM_EINVAL, "op must be 0 (32x32x16 f16), 1 (32x32x16 bf16) or 2 (16x16x32 f16)");
    if (iters < 1 || iters > 1000000) return fail(ZTM_EINVAL, "iters must be in [1, 1e6]");
    if (!tflops) return fail(ZTM_EINVAL, "null tflops");
    hipStream_t s = (hipStream_t)stream;
    const int blocks = 8 * 256;   // 8 workgroups of 4 waves per CU on 256 CUs
    const int threads = blocks * 256;
    std::vector<uint32_t> host(kOps * 4);
    calib_operands(op, low_entropy, host);
    uint4* ops = nullptr;
    float* out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&ops, sizeof(uint32_t) * host.size());
    if (e == hipSuccess) e = hipMalloc(&out, sizeof(float) * threads);
    if (e == hipSuccess)
        e = hipMemcpyAsync(ops, host.data(), sizeof(uint32_t) * host.size(),
                           hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    float ms = 0.f;
    for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) {   // rep 0 warms up
        e = hipEventRecord(e0, s);
        if (e != hipSuccess) break;
        switch (op) {
            case 0: hipLaunchKernelGGL(calib_kernel<0>, dim3(blocks), dim3(256), 0, s, ops, out, iters); break;
            case 1: hipLaunchKernelGGL(calib_kernel<1>, dim3(blocks), dim3(256), 0, s, ops, out, iters); break;
            default: hipLaunchKernelGGL(calib_kernel<2>, dim3(blocks), dim3(256), 0, s, ops, out, iters); break;
        }
        e = hipGetLastError();
        if (e == hipSuccess) e = hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    }
    const int rc = e == hipSuccess ? ZTM_OK : hip_fail(e, "MFMA calibration");
    if (rc == ZTM_OK) {
        // FLOP per wave and iteration: 8 x 32x32x16 or 32 x 16x16x32
        const double per_iter = op == 2 ? 32.0 * 16384.0 : 8.0 * 32768.0;
        *tflops = (double)threads / 64.0 * iters * per_iter / ((double)ms * 1e-3) / 1e12;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (ops) (void)hipFree(ops);
    if (out) (void)hipFree(out);
    return rc;
}
