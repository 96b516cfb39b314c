// Relaxation-time label search on the GPU: the candidate loop of
// Mesh::FindAndWriteTau (IPPI-LBM/mesh.cpp:2667-2704), which the reference
// runs on one CPU core per node (20,001 candidates, 18 divisions each).
//
// For a node, population k streams in from a source node s_k with
// post-streaming populations fnew, equilibrium feq and relaxation time tau.
// For each candidate dtau = m * step the source populations are collided
// with tau + dtau,
//     fcol_k = (1 - 1/(tau_k + dtau)) fnew_k + feq_k / (tau_k + dtau),
// their sums rho, sum e_x fcol, sum e_y fcol are compared with the target
// (rho*, u*, v*) of the high-fidelity data, score
//     |rho - rho*| + |jx - u*| + |jy - v*|,
// and the first m with the smallest score is kept. Every operation is FP64
// in the reference's order (IEEE division, no contraction), so the chosen
// candidate and its score equal the CPU search bit for bit.
//
// Layout: one wave per node. Lane l scores m = m_lo + l, m_lo + l + 64, ...
// in increasing order and keeps its first minimum; the wave then reduces the
// lanes' (score, m) pairs to the smallest score, ties to the smallest m,
// which is the sequential loop's first minimum.
//
// Divisions. The two divisions of a population share their denominator
// x = tau_k + dtau. gfx950 divides in FP64 by scaling (v_div_scale), a
// reciprocal refined by two Newton steps, the quotient and one
// correction (v_div_fmas), and a fix-up of special cases (v_div_fixup).
// When every x of the node's candidate range and every numerator lie within
// 2^-200 .. 2^200 in magnitude (numerators may also be 0), the scaling steps
// are identities and the fix-up returns the corrected quotient unchanged
// (up to the sign of a zero quotient, which the score's absolute values
// cannot see), so the same steps without them give the same bits. That
// path computes the refined reciprocal once for both divisions: 14 FP64
// instructions per population instead of ~27. Nodes outside those bounds
// take plain divisions.
#include <hip/hip_runtime.h>

#include <climits>
#include <mutex>
#include <string>

#include "../../include/ippi_lbm.h"

// Staging of ippi_tau_search_noalloc: one fixed device array per GPU that
// the library holds until the process ends, carved per call into the
// inputs fnew, feq, tau [n][9], target [n][3] and the outputs best_dtau,
// best_err [n] (256 bytes per node). Host code reaches it only through
// the address hipGetSymbolAddress returns for the current device.
constexpr size_t TS_BYTES_PER_NODE = 8 * (9 + 9 + 9 + 3 + 1 + 1);
constexpr size_t TS_SCRATCH_BYTES = TS_BYTES_PER_NODE * IPPI_TAU_SEARCH_NOALLOC_MAX;
static __device__ __align__(256) unsigned char scratch[TS_SCRATCH_BYTES];

namespace {

constexpr int TS_NODES = 4;  // nodes (waves) per block

constexpr double TS_TINY = 0x1p-200, TS_HUGE = 0x1p200;
__constant__ double TS_EX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
__constant__ double TS_EY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};

__global__ void __launch_bounds__(64 * TS_NODES)
    k_tau_search(int n, const double* __restrict__ fnew, const double* __restrict__ feq,
                 const double* __restrict__ tau, const double* __restrict__ target, int m_lo, int m_hi, double step,
                 double* __restrict__ best_dtau, double* __restrict__ best_err) {
#pragma clang fp contract(off)
    const int node = blockIdx.x * TS_NODES + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (node >= n) return;  // whole waves
    double fn[9], fq[9], tk[9];
#pragma unroll
    for (int k = 0; k < 9; k++) {
        fn[k] = fnew[9 * (size_t)node + k];
        fq[k] = feq[9 * (size_t)node + k];
        tk[k] = tau[9 * (size_t)node + k];
    }
    const double rt = target[3 * (size_t)node], ut = target[3 * (size_t)node + 1], vt = target[3 * (size_t)node + 2];
    double beste = 1e10;
    int bestm = INT_MAX;
    auto score = [&](int m, const double* fc) {
        double r = 0, u = 0, v = 0;
#pragma unroll
        for (int k = 0; k < 9; k++) {
            r += fc[k];
            u += fc[k] * TS_EX[k];
            v += fc[k] * TS_EY[k];
        }
        const double err = fabs(r - rt) + fabs(u - ut) + fabs(v - vt);
        if (err < beste) {
            beste = err;
            bestm = m;
        }
    };
    // bounds of the shared-reciprocal path (the same for every lane of the wave)
    const double d0 = (double)m_lo * step, d1 = (double)m_hi * step;
    const double dlo = fmin(d0, d1), dhi = fmax(d0, d1);
    bool shared = true;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const double lo = tk[k] + dlo, hi = tk[k] + dhi, a = fabs(fq[k]);
        shared = shared && (lo > TS_TINY || hi < -TS_TINY) && fabs(lo) < TS_HUGE && fabs(hi) < TS_HUGE &&
                 (fq[k] == 0.0 || (a > TS_TINY && a < TS_HUGE));
    }
    if (shared) {
        for (int m = m_lo + lane; m <= m_hi; m += 64) {
            const double dt = (double)m * step;
            double fc[9];
#pragma unroll
            for (int k = 0; k < 9; k++) {
                const double x = tk[k] + dt;
                const double y0 = __builtin_amdgcn_rcp(x);
                const double y1 = fma(y0, fma(-x, y0, 1.0), y0);
                const double y2 = fma(y1, fma(-x, y1, 1.0), y1);
                const double qa = fma(fma(-x, y2, 1.0), y2, y2);  // 1 / x
                const double qb0 = fq[k] * y2;
                const double qb = fma(fma(-x, qb0, fq[k]), y2, qb0);  // fq / x
                fc[k] = (1.0 - qa) * fn[k] + qb;
            }
            score(m, fc);
        }
    } else {
        for (int m = m_lo + lane; m <= m_hi; m += 64) {
            const double dt = (double)m * step;
            double fc[9];
#pragma unroll
            for (int k = 0; k < 9; k++) fc[k] = (1.0 - 1.0 / (tk[k] + dt)) * fn[k] + fq[k] / (tk[k] + dt);
            score(m, fc);
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double oe = __shfl_xor(beste, off, 64);
        const int om = __shfl_xor(bestm, off, 64);
        if (oe < beste || (oe == beste && om < bestm)) {
            beste = oe;
            bestm = om;
        }
    }
    if (lane == 0) {
        best_err[node] = beste;
        best_dtau[node] = bestm == INT_MAX ? 1e10 : (double)bestm * step;
    }
}

}  // namespace

int ippi_tau_search(int device, int n, const double* fnew, const double* feq, const double* tau,
                    const double* target, int m_lo, int m_hi, double step, double* best_dtau, double* best_err) {
    if (n < 0 || (n > 0 && (!fnew || !feq || !tau || !target || !best_dtau || !best_err)) || m_hi < m_lo ||
        (long long)m_hi - (long long)m_lo > (1LL << 30) || m_hi > INT_MAX - 64)
        return IPPI_ERR_ARG;
    if (n == 0) return IPPI_OK;
    if (hipSetDevice(device) != hipSuccess) return IPPI_ERR_HIP;
    const size_t s9 = sizeof(double) * 9 * (size_t)n, s3 = sizeof(double) * 3 * (size_t)n, s1 = sizeof(double) * n;
    double *d_fn = nullptr, *d_fq = nullptr, *d_tk = nullptr, *d_tg = nullptr, *d_bd = nullptr, *d_be = nullptr;
    int rc = IPPI_OK;
    auto ok = [&](hipError_t e, int code) {
        if (e != hipSuccess && rc == IPPI_OK) rc = code;
        return e == hipSuccess;
    };
    if (ok(hipMalloc(&d_fn, s9), IPPI_ERR_NOMEM) && ok(hipMalloc(&d_fq, s9), IPPI_ERR_NOMEM) &&
        ok(hipMalloc(&d_tk, s9), IPPI_ERR_NOMEM) && ok(hipMalloc(&d_tg, s3), IPPI_ERR_NOMEM) &&
        ok(hipMalloc(&d_bd, s1), IPPI_ERR_NOMEM) && ok(hipMalloc(&d_be, s1), IPPI_ERR_NOMEM) &&
        ok(hipMemcpy(d_fn, fnew, s9, hipMemcpyHostToDevice), IPPI_ERR_HIP) &&
        ok(hipMemcpy(d_fq, feq, s9, hipMemcpyHostToDevice), IPPI_ERR_HIP) &&
        ok(hipMemcpy(d_tk, tau, s9, hipMemcpyHostToDevice), IPPI_ERR_HIP) &&
        ok(hipMemcpy(d_tg, target, s3, hipMemcpyHostToDevice), IPPI_ERR_HIP)) {
        const dim3 g((unsigned)((n + TS_NODES - 1) / TS_NODES));
        hipLaunchKernelGGL(k_tau_search, g, dim3(64 * TS_NODES), 0, 0, n, d_fn, d_fq, d_tk, d_tg, m_lo, m_hi, step,
                           d_bd, d_be);
        if (ok(hipGetLastError(), IPPI_ERR_HIP) && ok(hipDeviceSynchronize(), IPPI_ERR_HIP)) {
            ok(hipMemcpy(best_dtau, d_bd, s1, hipMemcpyDeviceToHost), IPPI_ERR_HIP);
            ok(hipMemcpy(best_err, d_be, s1, hipMemcpyDeviceToHost), IPPI_ERR_HIP);
        }
    }
    for (double* p : {d_fn, d_fq, d_tk, d_tg, d_bd, d_be})
        if (p) (void)hipFree(p);
    return rc;
}

namespace {

constexpr int TS_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[TS_MAX_DEVICES] = {};

// Device address of this GPU's staging array (device must be current),
// looked up once per device.
int scratch_ptr(int device, unsigned char** out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return IPPI_ERR_HIP;
        g_scratch_ptr[device] = p;
    }
    *out = static_cast<unsigned char*>(g_scratch_ptr[device]);
    return IPPI_OK;
}

}  // namespace

int ippi_tau_search_noalloc(int device, int n, const double* fnew, const double* feq, const double* tau,
                            const double* target, int m_lo, int m_hi, double step, double* best_dtau,
                            double* best_err) {
    if (n < 0 || (n > 0 && (!fnew || !feq || !tau || !target || !best_dtau || !best_err)) || m_hi < m_lo ||
        (long long)m_hi - (long long)m_lo > (1LL << 30) || m_hi > INT_MAX - 64)
        return IPPI_ERR_ARG;
    if (n > IPPI_TAU_SEARCH_NOALLOC_MAX) return IPPI_ERR_CAPACITY;
    if (n == 0) return IPPI_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess) return IPPI_ERR_HIP;
    if (device < 0 || device >= ndev || device >= TS_MAX_DEVICES) return IPPI_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return IPPI_ERR_HIP;
    unsigned char* base = nullptr;
    int rc = scratch_ptr(device, &base);
    if (rc) return rc;
    const size_t s9 = sizeof(double) * 9 * (size_t)n, s3 = sizeof(double) * 3 * (size_t)n, s1 = sizeof(double) * n;
    double* d_fn = reinterpret_cast<double*>(base);
    double* d_fq = d_fn + 9 * (size_t)n;
    double* d_tk = d_fq + 9 * (size_t)n;
    double* d_tg = d_tk + 9 * (size_t)n;
    double* d_bd = d_tg + 3 * (size_t)n;
    double* d_be = d_bd + (size_t)n;
    hipStream_t st = nullptr;  // the device's null stream: ordered after earlier calls of this function
    auto ok = [&](hipError_t e) {
        if (e != hipSuccess && rc == IPPI_OK) rc = IPPI_ERR_HIP;
        return e == hipSuccess;
    };
    if (ok(hipMemcpyAsync(d_fn, fnew, s9, hipMemcpyHostToDevice, st)) &&
        ok(hipMemcpyAsync(d_fq, feq, s9, hipMemcpyHostToDevice, st)) &&
        ok(hipMemcpyAsync(d_tk, tau, s9, hipMemcpyHostToDevice, st)) &&
        ok(hipMemcpyAsync(d_tg, target, s3, hipMemcpyHostToDevice, st))) {
        const dim3 g((unsigned)((n + TS_NODES - 1) / TS_NODES));
        hipLaunchKernelGGL(k_tau_search, g, dim3(64 * TS_NODES), 0, st, n, d_fn, d_fq, d_tk, d_tg, m_lo, m_hi, step,
                           d_bd, d_be);
        if (ok(hipGetLastError()) && ok(hipMemcpyAsync(best_dtau, d_bd, s1, hipMemcpyDeviceToHost, st)) &&
            ok(hipMemcpyAsync(best_err, d_be, s1, hipMemcpyDeviceToHost, st)))
            ok(hipStreamSynchronize(st));
    }
    return rc;
}
