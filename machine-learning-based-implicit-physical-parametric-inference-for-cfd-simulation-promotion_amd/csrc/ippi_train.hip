// Training of the relaxation-time MLP on the GPU (part of libippi_lbm.so).
//
// Reference counterpart: the reference ships pre-trained TorchScript models
// (NNmodel/AutoT0*.pt) whose labels come from FindAndWriteTau
// (IPPI-LBM/mesh.cpp:2625-2715); here the network is trained by the
// project's own kernels on pairs the solver generates (ippi_coarse_pairs).
//
// Loss: L = sum_i w_i (1/tau(x_i) - y_i)^2 / (2 sum_i w_i) with targets y
// the relaxation rates omega = 1/tau, FP32 arithmetic.
// Gradient: one block per chunk of 128 samples; each lane runs its sample's
// forward and backward pass, the per-sample factors go to LDS and every
// parameter's partial sum over the chunk is formed by one lane in sample
// order; chunk partials are added in chunk order (deterministic, no atomics).
// Adam then updates FP32 parameters. Feature normalisation (mu, isig) is part
// of the model but not trained.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ippi_lbm.h"
#include "lbm_node.hpp"
#include "mlp.hpp"

using namespace lbm;

namespace {

constexpr int TS = 128;  // samples per gradient block
constexpr int NP = MLP_NPARAM;

struct TrainArgs {
    const float* x;   // [n][6] raw features
    const float* y;   // [n] target tau
    const float* w;   // [n] sample weight
    const int* idx;   // [n_batch] sample indices (null: identity)
    int n;            // samples in this pass
    const float* p;   // [NP] parameters
    float mu[6], isig[6];
    float* part;      // [nchunk][NP + 2]: gradient partials, loss partial, weight partial
};

__device__ __forceinline__ float sig(float z) { return 1.0f / (1.0f + expf(-z)); }

__global__ void __launch_bounds__(TS) k_grad_partial(TrainArgs a) {
    __shared__ float X[TS][8];
    __shared__ float H1[TS][MLP_NH];
    __shared__ float D1[TS][MLP_NH];  // dL/dz1
    __shared__ float D2[TS][MLP_NH];  // dL/dz2, then (do * h2) for the output layer
    __shared__ float DO[TS];
    __shared__ float LS[TS], WS[TS];
    const int t = threadIdx.x;
    const int s = blockIdx.x * TS + t;
    const float* p = a.p;
    float x[6], z1[MLP_NH], h1[MLP_NH];
    float wgt = 0.f, yv = 0.f;
    if (s < a.n) {
        const int i = a.idx ? a.idx[s] : s;
#pragma unroll
        for (int k = 0; k < 6; k++) x[k] = (a.x[6 * i + k] - a.mu[k]) * a.isig[k];
        wgt = a.w[i];
        yv = a.y[i];
    } else {
#pragma unroll
        for (int k = 0; k < 6; k++) x[k] = 0.f;
    }
    // forward
#pragma unroll 4
    for (int j = 0; j < MLP_NH; j++) {
        float z = p[OFF_B1 + j];
#pragma unroll
        for (int k = 0; k < 6; k++) z += p[OFF_W1 + j * 6 + k] * x[k];
        z1[j] = z;
        h1[j] = z * sig(z);
    }
    float o = p[OFF_B3];
    float d2[MLP_NH];
    for (int j = 0; j < MLP_NH; j++) {
        float z = p[OFF_B2 + j];
#pragma unroll
        for (int k = 0; k < MLP_NH; k++) z += p[OFF_W2 + j * MLP_NH + k] * h1[k];
        const float sg = sig(z);
        const float h2 = z * sg;
        o += p[OFF_W3 + j] * h2;
        d2[j] = sg * (1.0f + z * (1.0f - sg));  // silu'(z2), scaled below
        D2[t][j] = h2;                            // keep h2 for the output-layer gradient
    }
    const float tau = (o > 15.0f ? o : log1pf(expf(o))) + 0.5f;
    const float r = 1.0f / tau - yv;
    const float dout = -wgt * r * sig(o) / (tau * tau);  // dL/do (times sum w, applied at the end)
    LS[t] = wgt * r * r;
    WS[t] = wgt;
    DO[t] = dout;
    // backward through layer 2
    float dh1[MLP_NH];
#pragma unroll
    for (int k = 0; k < MLP_NH; k++) dh1[k] = 0.f;
    for (int j = 0; j < MLP_NH; j++) {
        const float dz2 = dout * p[OFF_W3 + j] * d2[j];
        d2[j] = dz2;
#pragma unroll
        for (int k = 0; k < MLP_NH; k++) dh1[k] += dz2 * p[OFF_W2 + j * MLP_NH + k];
    }
#pragma unroll
    for (int k = 0; k < 8; k++) X[t][k] = k < 6 ? x[k] : 0.f;
#pragma unroll
    for (int j = 0; j < MLP_NH; j++) {
        H1[t][j] = h1[j];
        const float sg = sig(z1[j]);
        D1[t][j] = dh1[j] * sg * (1.0f + z1[j] * (1.0f - sg));
    }
    __syncthreads();
    // output layer: dW3[j] = sum do*h2[j], db3 = sum do
    float* out = a.part + (size_t)blockIdx.x * (NP + 2);
    for (int q = t; q < MLP_NH + 1; q += TS) {
        float acc = 0.f;
        for (int u = 0; u < TS; u++) acc += DO[u] * (q < MLP_NH ? D2[u][q] : 1.0f);
        out[q < MLP_NH ? OFF_W3 + q : OFF_B3] = acc;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < MLP_NH; j++) D2[t][j] = d2[j];
    __syncthreads();
    // hidden layers
    for (int q = t; q < MLP_NH * MLP_NH + MLP_NH; q += TS) {
        float acc = 0.f;
        if (q < MLP_NH * MLP_NH) {
            const int j = q / MLP_NH, k = q % MLP_NH;
            for (int u = 0; u < TS; u++) acc += D2[u][j] * H1[u][k];
            out[OFF_W2 + q] = acc;
        } else {
            const int j = q - MLP_NH * MLP_NH;
            for (int u = 0; u < TS; u++) acc += D2[u][j];
            out[OFF_B2 + j] = acc;
        }
    }
    for (int q = t; q < MLP_NH * 6 + MLP_NH; q += TS) {
        float acc = 0.f;
        if (q < MLP_NH * 6) {
            const int j = q / 6, k = q % 6;
            for (int u = 0; u < TS; u++) acc += D1[u][j] * X[u][k];
            out[OFF_W1 + q] = acc;
        } else {
            const int j = q - MLP_NH * 6;
            for (int u = 0; u < TS; u++) acc += D1[u][j];
            out[OFF_B1 + j] = acc;
        }
    }
    if (t == 0) {
        float l = 0.f, w = 0.f;
        for (int u = 0; u < TS; u++) {
            l += LS[u];
            w += WS[u];
        }
        out[NP] = l;
        out[NP + 1] = w;
    }
}

// Sum the chunk partials in chunk order; grad = sum / sum w, loss = sum / (2 sum w).
// With adam != 0, also take one Adam step on p (m, v moments, step count t).
__global__ void k_grad_final(const float* __restrict__ part, int nchunk, float* __restrict__ grad,
                             float* __restrict__ loss, float* __restrict__ p, float* __restrict__ m,
                             float* __restrict__ v, int adam, float lr, float b1, float b2, float eps, int step) {
    __shared__ float sw;
    if (threadIdx.x == 0) {
        float w = 0.f, l = 0.f;
        for (int c = 0; c < nchunk; c++) {
            w += part[(size_t)c * (NP + 2) + NP + 1];
            l += part[(size_t)c * (NP + 2) + NP];
        }
        sw = w > 0.f ? w : 1.f;
        if (loss) *loss = l / (2.0f * sw);
    }
    __syncthreads();
    for (int q = threadIdx.x; q < NP; q += blockDim.x) {
        float g = 0.f;
        for (int c = 0; c < nchunk; c++) g += part[(size_t)c * (NP + 2) + q];
        g /= sw;
        if (grad) grad[q] = g;
        if (adam) {
            const float mq = b1 * m[q] + (1.0f - b1) * g;
            const float vq = b2 * v[q] + (1.0f - b2) * g * g;
            m[q] = mq;
            v[q] = vq;
            const float mh = mq / (1.0f - powf(b1, (float)step));
            const float vh = vq / (1.0f - powf(b2, (float)step));
            p[q] -= lr * mh / (sqrtf(vh) + eps);
        }
    }
}

// Minibatch indices: splitmix64 keyed by (seed, iteration, slot), modulo n.
__global__ void k_batch(int* __restrict__ idx, int nb, int n, unsigned long long seed, int it) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    unsigned long long z = seed + ((unsigned long long)it * 0x100000001ull + (unsigned long long)i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    idx[i] = (int)(z % (unsigned long long)n);
}

}  // namespace

struct ippi_trainer {
    int cap = 0, n = 0;
    float *x = nullptr, *y = nullptr, *w = nullptr, *p = nullptr, *m = nullptr, *v = nullptr;
    float *part = nullptr, *grad = nullptr, *loss = nullptr;
    int* idx = nullptr;
    int step = 0;
    double mu[6] = {0}, isig[6] = {0};
    hipStream_t s = nullptr;
    std::string err;
};

namespace {
int tfail(ippi_trainer* t, int code, const std::string& msg) {
    if (t) t->err = msg;
    return code;
}
#define TCK(t, expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) return tfail(t, IPPI_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

int run_grad(ippi_trainer* t, int n, const int* idx, bool adam, float lr) {
    if (n <= 0) return tfail(t, IPPI_ERR_ARG, "no samples");
    const int nchunk = (n + TS - 1) / TS;
    TrainArgs a{};
    a.x = t->x;
    a.y = t->y;
    a.w = t->w;
    a.idx = idx;
    a.n = n;
    a.p = t->p;
    for (int i = 0; i < 6; i++) {
        a.mu[i] = (float)t->mu[i];
        a.isig[i] = (float)t->isig[i];
    }
    a.part = t->part;
    hipLaunchKernelGGL(k_grad_partial, dim3(nchunk), dim3(TS), 0, t->s, a);
    TCK(t, hipGetLastError());
    if (adam) t->step++;
    hipLaunchKernelGGL(k_grad_final, dim3(1), dim3(1024), 0, t->s, (const float*)t->part, nchunk, t->grad, t->loss,
                       t->p, t->m, t->v, adam ? 1 : 0, lr, 0.9f, 0.999f, 1e-8f, t->step);
    TCK(t, hipGetLastError());
    return IPPI_OK;
}
}  // namespace

extern "C" {

const char* ippi_trainer_last_error(const ippi_trainer* t) { return t ? t->err.c_str() : "null trainer"; }

int ippi_trainer_create(int capacity, ippi_trainer** out) {
    if (!out || capacity <= 0) return IPPI_ERR_ARG;
    ippi_trainer* t = new ippi_trainer();
    t->cap = capacity;
    const int nchunk = (capacity + TS - 1) / TS;
    bool ok = hipStreamCreateWithFlags(&t->s, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipMalloc(&t->x, sizeof(float) * 6 * capacity) == hipSuccess;
    ok = ok && hipMalloc(&t->y, sizeof(float) * capacity) == hipSuccess;
    ok = ok && hipMalloc(&t->w, sizeof(float) * capacity) == hipSuccess;
    ok = ok && hipMalloc(&t->idx, sizeof(int) * capacity) == hipSuccess;
    ok = ok && hipMalloc(&t->p, sizeof(float) * NP) == hipSuccess;
    ok = ok && hipMalloc(&t->m, sizeof(float) * NP) == hipSuccess;
    ok = ok && hipMalloc(&t->v, sizeof(float) * NP) == hipSuccess;
    ok = ok && hipMalloc(&t->grad, sizeof(float) * NP) == hipSuccess;
    ok = ok && hipMalloc(&t->loss, sizeof(float)) == hipSuccess;
    ok = ok && hipMalloc(&t->part, sizeof(float) * (size_t)nchunk * (NP + 2)) == hipSuccess;
    if (!ok) {
        ippi_trainer_destroy(t);
        return IPPI_ERR_NOMEM;
    }
    *out = t;
    return IPPI_OK;
}

int ippi_trainer_destroy(ippi_trainer* t) {
    if (!t) return IPPI_OK;
    if (t->s) (void)hipStreamSynchronize(t->s);
    for (void* q : {(void*)t->x, (void*)t->y, (void*)t->w, (void*)t->idx, (void*)t->p, (void*)t->m, (void*)t->v,
                    (void*)t->grad, (void*)t->loss, (void*)t->part})
        if (q) (void)hipFree(q);
    if (t->s) (void)hipStreamDestroy(t->s);
    delete t;
    return IPPI_OK;
}

int ippi_trainer_set_data(ippi_trainer* t, const float* x, const float* y, const float* w, int n) {
    if (!t || !x || !y || n <= 0) return IPPI_ERR_ARG;
    if (n > t->cap) return tfail(t, IPPI_ERR_ARG, "more samples than the trainer's capacity");
    std::vector<float> ones;
    if (!w) {
        ones.assign(n, 1.0f);
        w = ones.data();
    }
    TCK(t, hipMemcpy(t->x, x, sizeof(float) * 6 * n, hipMemcpyHostToDevice));
    TCK(t, hipMemcpy(t->y, y, sizeof(float) * n, hipMemcpyHostToDevice));
    TCK(t, hipMemcpy(t->w, w, sizeof(float) * n, hipMemcpyHostToDevice));
    t->n = n;
    return IPPI_OK;
}

int ippi_trainer_set_params(ippi_trainer* t, const double* blob, int n) {
    if (!t || !blob || n != MLP_BLOB) return IPPI_ERR_ARG;
    std::vector<float> p(NP);
    for (int i = 0; i < NP; i++) p[i] = (float)blob[i];
    for (int i = 0; i < 6; i++) {
        t->mu[i] = blob[OFF_MU + i];
        t->isig[i] = blob[OFF_IS + i];
    }
    TCK(t, hipMemcpy(t->p, p.data(), sizeof(float) * NP, hipMemcpyHostToDevice));
    TCK(t, hipMemset(t->m, 0, sizeof(float) * NP));
    TCK(t, hipMemset(t->v, 0, sizeof(float) * NP));
    t->step = 0;
    return IPPI_OK;
}

int ippi_trainer_get_params(ippi_trainer* t, double* blob, int n) {
    if (!t || !blob || n != MLP_BLOB) return IPPI_ERR_ARG;
    std::vector<float> p(NP);
    TCK(t, hipStreamSynchronize(t->s));
    TCK(t, hipMemcpy(p.data(), t->p, sizeof(float) * NP, hipMemcpyDeviceToHost));
    for (int i = 0; i < NP; i++) blob[i] = p[i];
    for (int i = 0; i < 6; i++) {
        blob[OFF_MU + i] = t->mu[i];
        blob[OFF_IS + i] = t->isig[i];
    }
    return IPPI_OK;
}

int ippi_trainer_loss_grad(ippi_trainer* t, double* loss, double* grad) {
    if (!t) return IPPI_ERR_ARG;
    int r = run_grad(t, t->n, nullptr, false, 0.f);
    if (r) return r;
    std::vector<float> g(NP);
    float l = 0.f;
    TCK(t, hipStreamSynchronize(t->s));
    TCK(t, hipMemcpy(g.data(), t->grad, sizeof(float) * NP, hipMemcpyDeviceToHost));
    TCK(t, hipMemcpy(&l, t->loss, sizeof(float), hipMemcpyDeviceToHost));
    if (grad)
        for (int i = 0; i < NP; i++) grad[i] = g[i];
    if (loss) *loss = l;
    return IPPI_OK;
}

int ippi_trainer_adam(ippi_trainer* t, int iters, int batch, double lr, uint64_t seed, double* loss) {
    if (!t || iters < 0 || batch <= 0) return IPPI_ERR_ARG;
    if (t->n == 0) return tfail(t, IPPI_ERR_STATE, "no training data");
    if (batch > t->cap) return tfail(t, IPPI_ERR_ARG, "batch larger than the trainer's capacity");
    for (int it = 0; it < iters; it++) {
        const bool full = batch >= t->n;
        if (!full) {
            hipLaunchKernelGGL(k_batch, dim3((batch + 255) / 256), dim3(256), 0, t->s, t->idx, batch, t->n,
                               (unsigned long long)seed, it + t->step);
            TCK(t, hipGetLastError());
        }
        int r = run_grad(t, full ? t->n : batch, full ? nullptr : t->idx, true, (float)lr);
        if (r) return r;
    }
    float l = 0.f;
    TCK(t, hipStreamSynchronize(t->s));
    TCK(t, hipMemcpy(&l, t->loss, sizeof(float), hipMemcpyDeviceToHost));
    if (loss) *loss = l;
    return IPPI_OK;
}

}  // extern "C"
