// Relaxation-time MLP 6 -> 32 -> 32 -> 1 (SiLU, tau = softplus(o) + 0.5),
// evaluated inside the collision kernels.
//
// Reference counterpart: the per-node libtorch inference of
// Mesh::NNmodel_Type00..07 (IPPI-LBM/mesh.cpp:2767-3260) with its guard
// tau = NN if NN >= 0.501 else the LES value (mesh.cpp:2821-2827, :393).
//
// Parameter blob (host, double): W1[32][6] b1[32] W2[32][32] b2[32] W3[32] b3
// mu[6] isig[6]; features are normalised as (f - mu) * isig.
//
// FP32 path: one wave (64 nodes) evaluates the two hidden layers with bf16
// MFMA tiles (v_mfma_f32_16x16x16_bf16 for layer 1, K = 6 padded to 16;
// v_mfma_f32_16x16x32_bf16 for layer 2) with FP32 accumulation. Each bf16
// operand is split into hi + lo parts and the product is taken as
// hi*hi + hi*lo + lo*hi, which carries ~16 mantissa bits. Tiles are laid
// out as D[o][node]: A = weights (16 outputs x K), B = activations
// (K x 16 nodes). Lane l of an MFMA supplies A[l%16][4(l/16)..+4] (or
// 8(l/16)..+8 for K = 32), B[4(l/16)..+4][l%16] and receives
// D[4(l/16)+r][l%16], r = 0..3. The layer-1 result tiles feed the layer-2 B
// operand without a transpose: the host permutes W2's columns to the
// order in which the lanes hold the layer-1 outputs.
//
// SPLIT = false (IPPI_TAU_NN_BF16): single bf16 products (the hi parts
// only), FP32 accumulation: a third of the MFMAs and no lo parts, about 8
// mantissa bits per operand.
//
// Exact weights (mlp_tau_wave_exact): when every entry of W1 and W2 is a
// bf16 value, the products need no lo part of the weights: a product is
// W x_hi + W x_lo, exact in FP32 up to the 16-bit representation of x.
// Layer 1 then takes one 16x16x16 MFMA per 16 outputs, its K holding
// [x_hi(6), 1, 1, x_lo(6), 1, 0] against [W(6), p0, p1, W(6), p2, 0], where
// p0 + p1 + p2 are the bf16 parts of the scaled layer-1 bias; layer 2
// two 16x16x32 MFMAs (h_hi, h_lo). 24 MFMAs per wave instead of 48. The
// -log2(e) of the activations is folded into the feature scaling (isig)
// and the biases, so W1 stays exact.
//
// FP64 path: the same network on VALU in FP64 with the oracle's loop order.
#pragma once
#include <hip/hip_runtime.h>

namespace lbm {

constexpr int MLP_NIN = 6, MLP_NH = 32;
constexpr int MLP_NPARAM = MLP_NIN * MLP_NH + MLP_NH + MLP_NH * MLP_NH + MLP_NH + MLP_NH + 1;  // 1313
constexpr int MLP_BLOB = MLP_NPARAM + 2 * MLP_NIN;                                              // 1325
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + MLP_NH * MLP_NIN, OFF_W2 = OFF_B1 + MLP_NH,
              OFF_B2 = OFF_W2 + MLP_NH * MLP_NH, OFF_W3 = OFF_B2 + MLP_NH, OFF_B3 = OFF_W3 + MLP_NH,
              OFF_MU = OFF_B3 + 1, OFF_IS = OFF_MU + MLP_NIN;

typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));

// Per-lane fragments of the FP32/MFMA path (built on the host from the blob).
// Constants are folded so that every hidden unit costs exp2 + add + rcp +
// mul on VALU: with L = log2(e), layer 1 computes t1 = -L z1 (W1, b1 scaled
// by -L) and carries h1' = t1 sigma = -L h1; layer 2 computes
// t2 = W2 h1' + b2' = -L z2 (b2 scaled by -L, W2 unchanged); the output
// weights are scaled by -ln(2) so that w3' t2 sigma(t2) = w3 h2, where
// sigma = 1 / (1 + 2^t). The biases enter as the MFMA accumulator input.
struct MlpFrag {
    bf16x4 a1h[2], a1l[2];  // layer-1 A tiles (two 16-output blocks), hi / lo
    bf16x8 a2h[2], a2l[2];  // layer-2 A tiles, columns in lane order
    float b1[8], b2[8], w3[8];  // scaled bias / output weight of this lane's 8 outputs
    bf16x4 a1e[2];          // exact weights: layer-1 A tiles [W, b1 parts, W, b1 part, 0] (unscaled W1)
};
// Exact-weight network: the per-lane fragment (112 bytes) is read from LDS,
// where each workgroup copies it once (global layout chunk-major
// [MLP_FRAGE_CHUNKS][64 lanes] of 16-byte chunks, so that a wave's reads of
// one chunk are contiguous): chunk 0 = a1e[0..1] (W1 with the layer-1 bias
// parts), 1..2 = a2h[0..1] (exact W2), 3..4 = b2, 5..6 = w3, with the
// foldings of MlpFrag.
constexpr int MLP_FRAGE_CHUNKS = 7;
struct MlpDev {
    const MlpFrag* frag;  // [64 lanes]
    float mu[6], isig[6], b3;
    float isig_e[6];      // exact weights: -log2(e) isig
    const u32x4* frage;   // exact weights: [MLP_FRAGE_CHUNKS][64]
    const double* blob;   // FP64 path
};

// Copy of the exact-weight fragments into the workgroup's LDS (every thread
// of the block takes part; the caller synchronises before the first use).
template <int NT>
__device__ __forceinline__ void mlp_frage_to_lds(const MlpDev& m, u32x4* lf) {
    for (int i = threadIdx.x; i < MLP_FRAGE_CHUNKS * 64; i += NT) lf[i] = m.frage[i];
}

// 1 / (1 + 2^t) = sigma(z) for t = -log2(e) z
__device__ __forceinline__ float sig2(float t) {
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(t));
}
// softplus: the output is tau - 1/2, often ~1e-3, and log(1 + e^z) alone
// would lose its digits to the rounding of w = 1 + e^z. With u = e^z,
// log1p(u) = log(w) - ((w - 1) - u) / w to first order in the rounding
// error of w, where (w - 1) - u is that error exactly (Sterbenz). The term
// is kept for w < 2 only, without the division: it is at most 2^-24 there,
// and dividing it by w would change it by at most 2^-24 u / w, about half
// an ulp of log1p(u). For w >= 2 the whole term is below 2^-24 / w, under
// an ulp of log(w) >= 0.69. (Max relative error against log1p of the same
// u: ~2e-7.)
__device__ __forceinline__ float softplus_f(float z) {
    if (z > 15.0f) return z;
    const float u = __expf(z), w = 1.0f + u;
    const float err = w < 2.0f ? (w - 1.0f) - u : 0.0f;
    return __builtin_amdgcn_logf(w) * 0.69314718055994531f - err;
}

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// a, b -> hi = (bf16(a), bf16(b)), lo = (bf16(a - hi_a), bf16(b - hi_b)),
// each pair packed in one dword (a in the low half): one conversion per
// pair and per part, the hi values read back from the packed bits
__device__ __forceinline__ void split2(float a, float b, unsigned& hi, unsigned& lo) {
    hi = __builtin_bit_cast(unsigned, __builtin_convertvector((f32x2){a, b}, bf16x2));
    const float ah = __uint_as_float(hi << 16), bh = __uint_as_float(hi & 0xffff0000u);
    lo = __builtin_bit_cast(unsigned, __builtin_convertvector((f32x2){a - ah, b - bh}, bf16x2));
}

// tau - 1/2 of this lane's node from its six raw features; every lane of the wave
// must call it (the MFMAs are wave-wide). lds: 64 x 8 bf16 x 2 (hi, lo) per
// wave. The 64 nodes are handled as four blocks of 16; each block runs
// layer 1, the activation, layer 2 and the output sum before the next, so
// only one block's tiles are live at a time.
template <bool SPLIT>
__device__ __forceinline__ float mlp_tau_wave(const float* feat, const MlpDev& m, __bf16* lds) {
    // contraction lets the output sum acc += (w3 t2) sigma(t2) use FMAs; the
    // feature normalisation below keeps its subtraction (exact for features
    // close to their mean) before the scaling
#pragma clang fp contract(fast)
    const int l = threadIdx.x & 63;
    const int g = l >> 4;  // lane group
    const MlpFrag& fr = m.frag[l];
    // normalised features, split hi / lo, to LDS as [node][k] (k 0..7)
    unsigned vh[4], vl[4];
#pragma unroll
    for (int k = 0; k < 6; k += 2) {
        const float a = (feat[k] - m.mu[k]) * m.isig[k], b = (feat[k + 1] - m.mu[k + 1]) * m.isig[k + 1];
        if (SPLIT)
            split2(a, b, vh[k / 2], vl[k / 2]);
        else
            vh[k / 2] = __builtin_bit_cast(unsigned, __builtin_convertvector((f32x2){a, b}, bf16x2));
    }
    *reinterpret_cast<u32x4*>(lds + 16 * l) = (u32x4){vh[0], vh[1], vh[2], 0u};
    if (SPLIT) *reinterpret_cast<u32x4*>(lds + 16 * l + 8) = (u32x4){vl[0], vl[1], vl[2], 0u};
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    float part[4];
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
        const int node = nb * 16 + (l & 15);
        bf16x4 bh, bl;
        if (g < 2) {
            bh = *reinterpret_cast<const bf16x4*>(lds + 16 * node + 4 * g);
            if (SPLIT) bl = *reinterpret_cast<const bf16x4*>(lds + 16 * node + 8 + 4 * g);
        } else {
#pragma unroll
            for (int j = 0; j < 4; j++) {
                bh[j] = (__bf16)0.0f;
                bl[j] = (__bf16)0.0f;
            }
        }
        // layer 1 + activation -> layer-2 B operand (8 values per lane)
        unsigned ph[4], pl[4] = {0u, 0u, 0u, 0u};
#pragma unroll
        for (int mb = 0; mb < 2; mb++) {
            f32x4 c = {fr.b1[4 * mb], fr.b1[4 * mb + 1], fr.b1[4 * mb + 2], fr.b1[4 * mb + 3]};
            c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(fr.a1h[mb], bh, c, 0, 0, 0);
            if (SPLIT) {
                c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(fr.a1h[mb], bl, c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(fr.a1l[mb], bh, c, 0, 0, 0);
            }
#pragma unroll
            for (int r = 0; r < 4; r += 2) {
                const float h0 = c[r] * sig2(c[r]), h1 = c[r + 1] * sig2(c[r + 1]);
                if (SPLIT)
                    split2(h0, h1, ph[2 * mb + r / 2], pl[2 * mb + r / 2]);
                else
                    ph[2 * mb + r / 2] = __builtin_bit_cast(unsigned, __builtin_convertvector((f32x2){h0, h1}, bf16x2));
            }
        }
        const bf16x8 hh = __builtin_bit_cast(bf16x8, (u32x4){ph[0], ph[1], ph[2], ph[3]});
        const bf16x8 hl = __builtin_bit_cast(bf16x8, (u32x4){pl[0], pl[1], pl[2], pl[3]});
        // layer 2 + activation + output weights
        float acc = 0.0f;
#pragma unroll
        for (int mb = 0; mb < 2; mb++) {
            f32x4 c = {fr.b2[4 * mb], fr.b2[4 * mb + 1], fr.b2[4 * mb + 2], fr.b2[4 * mb + 3]};
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fr.a2h[mb], hh, c, 0, 0, 0);
            if (SPLIT) {
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fr.a2h[mb], hl, c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fr.a2l[mb], hh, c, 0, 0, 0);
            }
#pragma unroll
            for (int r = 0; r < 4; r++) acc += (fr.w3[4 * mb + r] * c[r]) * sig2(c[r]);
        }
        part[nb] = acc;
    }
    // layer 3: sum the four lane groups holding the same node
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
        part[nb] += __shfl_xor(part[nb], 16);
        part[nb] += __shfl_xor(part[nb], 32);
    }
    const float o = (g == 0 ? part[0] : g == 1 ? part[1] : g == 2 ? part[2] : part[3]) + m.b3;
    return softplus_f(o);
}

// Network with bf16-exact W1 and W2 (see the header): same layouts and
// folding as mlp_tau_wave, two MFMAs per product in layer 2 and one in
// layer 1. The layer-1 bias enters the MFMA as three bf16 parts of the
// weights against 1.0 entries in K slots 6, 7 and 14 of the activations
// (so the accumulator starts at zero and b1 takes no registers).
__device__ __forceinline__ float mlp_tau_wave_exact(const float* feat, const MlpDev& m, __bf16* lds,
                                                    const u32x4* lf) {
#pragma clang fp contract(fast)
    const int l = threadIdx.x & 63;
    const int g = l >> 4;
    struct {
        bf16x4 a1e[2];
        bf16x8 a2h[2];
        float b2[8], w3[8];
    } fr;
    static_assert(sizeof(fr) == 16 * MLP_FRAGE_CHUNKS, "fragment layout");
    u32x4* fc = reinterpret_cast<u32x4*>(&fr);
#pragma unroll
    for (int c = 0; c < MLP_FRAGE_CHUNKS; c++) fc[c] = lf[c * 64 + l];
    unsigned vh[3], vl[3];
#pragma unroll
    for (int k = 0; k < 6; k += 2) {
        const float a = (feat[k] - m.mu[k]) * m.isig_e[k], b = (feat[k + 1] - m.mu[k + 1]) * m.isig_e[k + 1];
        split2(a, b, vh[k / 2], vl[k / 2]);
    }
    // the node's K row: [x_hi(6), 1, 1, x_lo(6), 1, 0] (bf16 1.0 = 0x3f80)
    *reinterpret_cast<u32x4*>(lds + 16 * l) = (u32x4){vh[0], vh[1], vh[2], 0x3f803f80u};
    *reinterpret_cast<u32x4*>(lds + 16 * l + 8) = (u32x4){vl[0], vl[1], vl[2], 0x00003f80u};
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    float part[4];
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
        const int node = nb * 16 + (l & 15);
        // K rows 4g..4g+3 of the node's row as written above
        const bf16x4 bx = *reinterpret_cast<const bf16x4*>(lds + 16 * node + 4 * g);
        unsigned ph[4], pl[4];
#pragma unroll
        for (int mb = 0; mb < 2; mb++) {
            f32x4 c = {0.f, 0.f, 0.f, 0.f};
            c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(fr.a1e[mb], bx, c, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 4; r += 2) {
                const float h0 = c[r] * sig2(c[r]), h1 = c[r + 1] * sig2(c[r + 1]);
                split2(h0, h1, ph[2 * mb + r / 2], pl[2 * mb + r / 2]);
            }
        }
        const bf16x8 hh = __builtin_bit_cast(bf16x8, (u32x4){ph[0], ph[1], ph[2], ph[3]});
        const bf16x8 hl = __builtin_bit_cast(bf16x8, (u32x4){pl[0], pl[1], pl[2], pl[3]});
        float acc = 0.0f;
#pragma unroll
        for (int mb = 0; mb < 2; mb++) {
            f32x4 c = {fr.b2[4 * mb], fr.b2[4 * mb + 1], fr.b2[4 * mb + 2], fr.b2[4 * mb + 3]};
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fr.a2h[mb], hh, c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fr.a2h[mb], hl, c, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 4; r++) acc += (fr.w3[4 * mb + r] * c[r]) * sig2(c[r]);
        }
        part[nb] = acc;
    }
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
        part[nb] += __shfl_xor(part[nb], 16);
        part[nb] += __shfl_xor(part[nb], 32);
    }
    const float o = (g == 0 ? part[0] : g == 1 ? part[1] : g == 2 ? part[2] : part[3]) + m.b3;
    return softplus_f(o);
}

// Network mode of the FP32 kernels: 1 compensated bf16 products, 2 plain
// bf16 products, 3 bf16-exact weights.
template <int NN>
__device__ __forceinline__ float mlp_tau_nn(const float* feat, const MlpDev& m, __bf16* lds, const u32x4* lf) {
    if constexpr (NN == 3)
        return mlp_tau_wave_exact(feat, m, lds, lf);
    else
        return mlp_tau_wave<NN == 1>(feat, m, lds);
}

// FP64 network on VALU, the oracle's loop order (weights read wave-uniformly).
__device__ __forceinline__ double mlp_tau_f64(const double* in, const double* __restrict__ p) {
#pragma clang fp contract(off)
    double x[MLP_NIN], h1[MLP_NH];
#pragma unroll
    for (int i = 0; i < MLP_NIN; i++) x[i] = (in[i] - p[OFF_MU + i]) * p[OFF_IS + i];
#pragma unroll 4
    for (int j = 0; j < MLP_NH; j++) {
        double s = p[OFF_B1 + j];
#pragma unroll
        for (int i = 0; i < MLP_NIN; i++) s += p[OFF_W1 + j * MLP_NIN + i] * x[i];
        h1[j] = s / (1.0 + exp(-s));
    }
    double o = p[OFF_B3];
    for (int j = 0; j < MLP_NH; j++) {
        double s = p[OFF_B2 + j];
#pragma unroll
        for (int i = 0; i < MLP_NH; i++) s += p[OFF_W2 + j * MLP_NH + i] * h1[i];
        o += p[OFF_W3 + j] * (s / (1.0 + exp(-s)));
    }
    const double sp = o > 30.0 ? o : log1p(exp(o));
    return sp + 0.5;
}

}  // namespace lbm
