// Issue-rate probe (measurement only; no counterpart in the reference).
//
// The network variant of the fused step spends its time issuing VALU
// instructions, transcendentals and MFMAs. This probe times each of those
// instruction kinds alone and VALU mixed with MFMA, on registers only (no
// memory traffic besides one store per thread), so that the cost of each
// kind and whether MFMA and VALU work overlap can be read off:
//   kind 0: 64 FMAs per iteration (8 independent chains)
//   kind 1: 64 exp2
//   kind 2: 64 (add, rcp) pairs
//   kind 3: 16 SiLU activations as the network evaluates them: exp2, add,
//           rcp, mul
//   kind 4: 16 v_mfma_f32_16x16x32_bf16 (4 independent accumulators)
//   kind 5: kind 4 and kind 0 in the same loop
//   kind 6: 16 v_mfma_f32_16x16x4_f32
//   kind 7: kind 6 and kind 0 in the same loop
// Each launch runs IP_ITER iterations in every thread of 4096 blocks of 256
// threads. Values stay bounded (contractions towards fixed points), so no
// Inf/NaN path is timed.
#include <hip/hip_runtime.h>

#include "../../include/ippi_lbm.h"

namespace {

constexpr int IP_ITER = 256;
constexpr int IP_BLOCKS = 4096;

typedef __bf16 ip_bf16x8 __attribute__((ext_vector_type(8)));
typedef float ip_f32x4 __attribute__((ext_vector_type(4)));

template <int KIND>
__global__ void __launch_bounds__(256) k_issue(float* __restrict__ out, float seed) {
    float x[8];
#pragma unroll
    for (int i = 0; i < 8; i++) x[i] = seed * (float)(threadIdx.x + 1) * 1e-3f + (float)i * 1e-2f;
    ip_f32x4 acc[4] = {};
    ip_bf16x8 a, b;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        a[i] = (__bf16)(x[i] * 0.5f);
        b[i] = (__bf16)(x[i] * 0.25f);
    }
    const float af = x[0], bf = x[1];
    for (int it = 0; it < IP_ITER; it++) {
        if (KIND == 1) {
#pragma unroll
            for (int r = 0; r < 8; r++)
#pragma unroll
                for (int i = 0; i < 8; i++) x[i] = __builtin_amdgcn_exp2f(-x[i]);
        } else if (KIND == 2) {
#pragma unroll
            for (int r = 0; r < 8; r++)
#pragma unroll
                for (int i = 0; i < 8; i++) x[i] = __builtin_amdgcn_rcpf(x[i] + 1.0f);
        } else if (KIND == 3) {
#pragma unroll
            for (int r = 0; r < 2; r++)
#pragma unroll
                for (int i = 0; i < 8; i++) x[i] = x[i] * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-x[i]));
        } else if (KIND == 4 || KIND == 5) {
#pragma unroll
            for (int r = 0; r < 4; r++)
#pragma unroll
                for (int c = 0; c < 4; c++) acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[c], 0, 0, 0);
        } else if (KIND == 6 || KIND == 7) {
#pragma unroll
            for (int r = 0; r < 4; r++)
#pragma unroll
                for (int c = 0; c < 4; c++) acc[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(af, bf, acc[c], 0, 0, 0);
        }
        if (KIND == 0 || KIND == 5 || KIND == 7) {
#pragma unroll
            for (int r = 0; r < 8; r++)
#pragma unroll
                for (int i = 0; i < 8; i++) x[i] = __builtin_fmaf(x[i], 0.999f, 0.001f);
        }
    }
    float s = 0.0f;
#pragma unroll
    for (int i = 0; i < 8; i++) s += x[i];
#pragma unroll
    for (int c = 0; c < 4; c++) s += acc[c][0] + acc[c][1] + acc[c][2] + acc[c][3];
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

}  // namespace

int ippi_issue_probe(int device, int kind, int reps, double* us_per_launch) {
    if (kind < 0 || kind > 7 || reps < 1 || !us_per_launch) return IPPI_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return IPPI_ERR_HIP;
    float* out = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = IPPI_OK;
    auto chk = [&](hipError_t e, int code) {
        if (e != hipSuccess && rc == IPPI_OK) rc = code;
        return e == hipSuccess;
    };
    if (chk(hipMalloc(&out, sizeof(float) * 256 * IP_BLOCKS), IPPI_ERR_NOMEM) &&
        chk(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), IPPI_ERR_HIP) &&
        chk(hipEventCreate(&e0), IPPI_ERR_HIP) && chk(hipEventCreate(&e1), IPPI_ERR_HIP)) {
        auto launch = [&]() {
            const dim3 g(IP_BLOCKS), b(256);
            switch (kind) {
                case 0: hipLaunchKernelGGL(k_issue<0>, g, b, 0, st, out, 1.0f); break;
                case 1: hipLaunchKernelGGL(k_issue<1>, g, b, 0, st, out, 1.0f); break;
                case 2: hipLaunchKernelGGL(k_issue<2>, g, b, 0, st, out, 1.0f); break;
                case 3: hipLaunchKernelGGL(k_issue<3>, g, b, 0, st, out, 1.0f); break;
                case 4: hipLaunchKernelGGL(k_issue<4>, g, b, 0, st, out, 1.0f); break;
                case 5: hipLaunchKernelGGL(k_issue<5>, g, b, 0, st, out, 1.0f); break;
                case 6: hipLaunchKernelGGL(k_issue<6>, g, b, 0, st, out, 1.0f); break;
                default: hipLaunchKernelGGL(k_issue<7>, g, b, 0, st, out, 1.0f); break;
            }
        };
        launch();  // warm-up
        chk(hipGetLastError(), IPPI_ERR_HIP);
        chk(hipEventRecord(e0, st), IPPI_ERR_HIP);
        for (int i = 0; i < reps; i++) launch();
        chk(hipEventRecord(e1, st), IPPI_ERR_HIP);
        float ms = 0.0f;
        if (chk(hipGetLastError(), IPPI_ERR_HIP) && chk(hipEventSynchronize(e1), IPPI_ERR_HIP) &&
            chk(hipEventElapsedTime(&ms, e0, e1), IPPI_ERR_HIP))
            *us_per_launch = (double)ms * 1e3 / reps;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (out) (void)hipFree(out);
    return rc;
}
