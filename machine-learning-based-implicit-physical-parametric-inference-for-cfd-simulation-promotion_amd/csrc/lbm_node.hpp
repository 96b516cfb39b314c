// Per-node D2Q9 arithmetic for the gfx950 stream-collide kernels.
//
// Two families:
//   ref::   FP64 on full populations f, in the evaluation order of the
//           reference's Lattice methods so the FP64 GPU path reproduces the
//           reference arithmetic bit for bit (the translation unit turns FP
//           contraction off for these functions).
//   sh::    FP32 on shifted populations g = f - w (w = lattice weights), the
//           bandwidth path. Same physics, FMA-friendly forms.
//
// Velocity order (IPPI-LBM/NodeAndLattice.cpp:3-13):
//   k : 0  1  2  3  4  5  6  7  8
//   ex: 0  1  0 -1  0  1 -1 -1  1
//   ey: 0  0  1  0 -1  1  1 -1 -1
#pragma once
#include <hip/hip_runtime.h>

#define IPPI_HD __host__ __device__ __forceinline__

namespace lbm {

constexpr int EX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
constexpr int EY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
constexpr int OPP[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
constexpr double WD[9] = {4. / 9., 1. / 9., 1. / 9., 1. / 9., 1. / 9., 1. / 36., 1. / 36., 1. / 36., 1. / 36.};
constexpr float WF[9] = {4.f / 9.f, 1.f / 9.f, 1.f / 9.f, 1.f / 9.f, 1.f / 9.f,
                         1.f / 36.f, 1.f / 36.f, 1.f / 36.f, 1.f / 36.f};
// MRT transform (NodeAndLattice.cpp:18-27) and 36 x its inverse (:28-29)
constexpr double TM[9][9] = {
    {1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0},
    {-4.0, -1.0, -1.0, -1.0, -1.0, 2.0, 2.0, 2.0, 2.0},
    {4.0, -2.0, -2.0, -2.0, -2.0, 1.0, 1.0, 1.0, 1.0},
    {0.0, 1.0, 0.0, -1.0, 0.0, 1.0, -1.0, -1.0, 1.0},
    {0.0, -2.0, 0.0, 2.0, 0.0, 1.0, -1.0, -1.0, 1.0},
    {0.0, 0.0, 1.0, 0.0, -1.0, 1.0, 1.0, -1.0, -1.0},
    {0.0, 0.0, -2.0, 0.0, 2.0, 1.0, 1.0, -1.0, -1.0},
    {0.0, 1.0, -1.0, 1.0, -1.0, 0.0, 0.0, 0.0, 0.0},
    {0.0, 0.0, 0.0, 0.0, 0.0, 1.0, -1.0, 1.0, -1.0}};
constexpr int TMINV_K[9][9] = {
    {4, -4, 4, 0, 0, 0, 0, 0, 0},    {4, -1, -2, 6, -6, 0, 0, 9, 0},
    {4, -1, -2, 0, 0, 6, -6, -9, 0}, {4, -1, -2, -6, 6, 0, 0, 9, 0},
    {4, -1, -2, 0, 0, -6, 6, -9, 0}, {4, 2, 1, 6, 3, 6, 3, 0, 9},
    {4, 2, 1, -6, -3, 6, 3, 0, -9},  {4, 2, 1, -6, -3, -6, -3, 0, 9},
    {4, 2, 1, 6, 3, -6, -3, 0, -9}};
// fixed MRT rates of the non-hydrodynamic moments (NodeAndLattice.cpp:306);
// entries 7 and 8 are 1/tau at run time, 0, 3, 5 are conserved.
constexpr double SM_FIXED[9] = {0.0, 1.2, 1.1, 0.0, 1.2, 0.0, 1.2, -1.0, -1.0};

// Peskin 4-point kernel, mesh.cpp:1354-1385 (host: marker weights; device: adjoint)
IPPI_HD double delta4(double x) {
#pragma clang fp contract(off)
    double ax = x < 0 ? -x : x;
    if (ax < 1) {
        double R = 1.0 + 4.0 * ax - 4 * (x * x);
        return (1.0 / 8.0) * (3.0 - 2.0 * ax + sqrt(R));
    } else if (ax < 2) {
        double R = -7.0 + 12.0 * ax - 4 * (x * x);
        return (1.0 / 8.0) * (5.0 - 2.0 * ax - sqrt(R));
    }
    return 0;
}

IPPI_HD double noise(unsigned long long seed, long long node, int comp) {
    unsigned long long z = seed + (unsigned long long)(2 * node + comp + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    double r = (double)(z >> 11) * (1.0 / 9007199254740992.0);
    return 2.0 * r - 1.0;
}

// ------------------------------------------------------------------ FP64
namespace ref {

// Compute_feq, NodeAndLattice.cpp:124-149
IPPI_HD void feq(double rho, double ux, double uy, double* q) {
#pragma clang fp contract(off)
    double w0 = rho * (4.0 / 9.0);
    double w1 = rho / 9.0;
    double w2 = rho / 36.0;
    double uu = 1 - 1.5 * (ux * ux + uy * uy);
    double s = ux + uy, d = ux - uy;
    q[0] = w0 * (uu);
    q[1] = w1 * (uu + 3.0 * (ux) + 4.5 * (ux) * (ux));
    q[2] = w1 * (uu + 3.0 * (uy) + 4.5 * (uy) * (uy));
    q[3] = w1 * (uu - 3.0 * (ux) + 4.5 * (ux) * (ux));
    q[4] = w1 * (uu - 3.0 * (uy) + 4.5 * (uy) * (uy));
    q[5] = w2 * (uu + 3.0 * s + 4.5 * s * s);
    q[6] = w2 * (uu - 3.0 * d + 4.5 * d * d);
    q[7] = w2 * (uu - 3.0 * s + 4.5 * s * s);
    q[8] = w2 * (uu + 3.0 * d + 4.5 * d * d);
}

// InitialLatticeFeq, NodeAndLattice.cpp:104-119 (w0 = rho*4.0/9.0)
IPPI_HD void feq_init(double rho, double ux, double uy, double* q) {
#pragma clang fp contract(off)
    double w0 = rho * 4.0 / 9.0;
    double w1 = rho / 9.0;
    double w2 = rho / 36.0;
    double uu = 1 - 1.5 * (ux * ux + uy * uy);
    double s = ux + uy, d = ux - uy;
    q[0] = w0 * (uu);
    q[1] = w1 * (uu + 3.0 * (ux) + 4.5 * (ux) * (ux));
    q[2] = w1 * (uu + 3.0 * (uy) + 4.5 * (uy) * (uy));
    q[3] = w1 * (uu - 3.0 * (ux) + 4.5 * (ux) * (ux));
    q[4] = w1 * (uu - 3.0 * (uy) + 4.5 * (uy) * (uy));
    q[5] = w2 * (uu + 3.0 * s + 4.5 * s * s);
    q[6] = w2 * (uu - 3.0 * d + 4.5 * d * d);
    q[7] = w2 * (uu - 3.0 * s + 4.5 * s * s);
    q[8] = w2 * (uu + 3.0 * d + 4.5 * d * d);
}

// CalMacro, NodeAndLattice.cpp:196-226. The reference loops over all nine
// e components; terms with e = 0 add a signed zero, so the sums below are
// the same values in the same order.
IPPI_HD void macro(const double* f, double& rho, double& ux, double& uy) {
#pragma clang fp contract(off)
    double r = 0.0;
#pragma unroll
    for (int k = 0; k < 9; k++) r += f[k];
    double tx = ((((f[1] - f[3]) + f[5]) - f[6]) - f[7]) + f[8];
    double ty = ((((f[2] - f[4]) + f[5]) + f[6]) - f[7]) - f[8];
    if (r < 1.e-6) {
        tx = 0.0;
        ty = 0.0;
    } else {
        tx /= r;
        ty /= r;
    }
    rho = r;
    ux = tx;
    uy = ty;
}

// ComputeBGKCollision, NodeAndLattice.cpp:150-161, given q = feq(rho, u)
IPPI_HD void bgk_q(const double* f, const double* q, double tau, double* o) {
#pragma clang fp contract(off)
    double a = (1.0 - 1.0 / tau);
#pragma unroll
    for (int k = 0; k < 9; k++) o[k] = a * f[k] + q[k] / tau;
}
IPPI_HD void bgk(const double* f, double rho, double ux, double uy, double tau, double* o) {
    double q[9];
    feq(rho, ux, uy, q);
    bgk_q(f, q, tau, o);
}

// ComputeMRTCollision, NodeAndLattice.cpp:290-360. Zero coefficients are
// skipped: each skipped term is a signed zero added to a running sum.
IPPI_HD void mrt(const double* f, double rho, double ux, double uy, double tau, double* o) {
#pragma clang fp contract(off)
    const double a1 = 1.0 / 36.0;
    double s78 = 1. / tau;
    double t1 = ux * ux + uy * uy;
    double t2 = ux * ux - uy * uy;
    double meq[9];
    meq[0] = rho;
    meq[1] = rho * (-2.0 + 3.0 * t1);
    meq[2] = rho * (1.0 - 3.0 * t1);
    meq[3] = rho * ux;
    meq[4] = -rho * ux;
    meq[5] = rho * uy;
    meq[6] = -rho * uy;
    meq[7] = rho * t2;
    meq[8] = rho * ux * uy;
    double dm[9];
#pragma unroll
    for (int k = 0; k < 9; k++) {
        double s = 0.0;
        bool first = true;
#pragma unroll
        for (int l = 0; l < 9; l++) {
            if (TM[k][l] != 0.0) {
                double t = TM[k][l] * f[l];
                s = first ? t : s + t;
                first = false;
            }
        }
        dm[k] = s - meq[k];
    }
#pragma unroll
    for (int k = 0; k < 9; k++) {
        double s = 0.;
        bool first = true;
#pragma unroll
        for (int l = 0; l < 9; l++) {
            if (TMINV_K[k][l] != 0 && SM_FIXED[l] != 0.0) {
                double sl = SM_FIXED[l] < 0 ? s78 : SM_FIXED[l];
                double t = (((double)TMINV_K[k][l] * a1) * sl) * dm[l];
                s = first ? t : s + t;
                first = false;
            }
        }
        o[k] = f[k] - s;
    }
}

// ComputeRRCollision, NodeAndLattice.cpp:228-286, given q = feq(rho, u)
IPPI_HD void rr_q(const double* f, const double* q, double tau, double* o) {
#pragma clang fp contract(off)
    // second moments of f and feq, reference loop order with e = 0 terms skipped
    double xx = ((((f[1] + f[3]) + f[5]) + f[6]) + f[7]) + f[8];
    double txx = ((((q[1] + q[3]) + q[5]) + q[6]) + q[7]) + q[8];
    double xy = ((f[5] - f[6]) + f[7]) - f[8];
    double txy = ((q[5] - q[6]) + q[7]) - q[8];
    double yy = ((((f[2] + f[4]) + f[5]) + f[6]) + f[7]) + f[8];
    double tyy = ((((q[2] + q[4]) + q[5]) + q[6]) + q[7]) + q[8];
    double nxx = xx - txx, nxy = xy - txy, nyy = yy - tyy;
    double a = (1.0 - 1.0 / tau);
    const double third = 1. / 3;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        double q0 = EX[k] * EX[k] - third;
        double q3 = EY[k] * EY[k] - third;
        double s = q0 * nxx;
        if (EX[k] * EY[k] != 0) {
            double q1 = EX[k] * EY[k];
            s = s + q1 * nxy;
            s = s + q1 * nxy;
        }
        s = s + q3 * nyy;
        double fone = 4.5 * WD[k] * s;
        o[k] = q[k] + a * fone;
    }
}

IPPI_HD void rr(const double* f, double rho, double ux, double uy, double tau, double* o) {
    double q[9];
    feq(rho, ux, uy, q);
    rr_q(f, q, tau, o);
}

// Compute_Disforce (Guo term), NodeAndLattice.cpp:162-195
IPPI_HD void guo(double ux, double uy, double fx, double fy, double tau_begin, double* o) {
#pragma clang fp contract(off)
    double omega = 1.0 - 0.5 * (1.0 / tau_begin);
#pragma unroll
    for (int k = 0; k < 9; k++) {
        double ex = EX[k], ey = EY[k];
        double eu = (ex * ux + ey * uy);
        double gx = WD[k] * omega * (3.0 * (ex - ux) + 9.0 * eu * ex) * fx;
        double gy = WD[k] * omega * (3.0 * (ey - uy) + 9.0 * eu * ey) * fy;
        o[k] = gx + gy;
    }
}

// CaculateLESTau live body, mesh.cpp:877-910, given q = feq(rho, u)
IPPI_HD double les_tau_q(const double* f, const double* q, double rho, double tau0, double cs) {
#pragma clang fp contract(off)
    double d[9];
#pragma unroll
    for (int k = 0; k < 9; k++) d[k] = f[k] - q[k];
    double Sxx = ((((d[1] + d[3]) + d[5]) + d[6]) + d[7]) + d[8];
    double Sxy = ((d[5] - d[6]) + d[7]) - d[8];
    double Syy = ((((d[2] + d[4]) + d[5]) + d[6]) + d[7]) + d[8];
    double c = -3.0 / (2 * rho * tau0);
    Sxx *= c;
    Sxy *= c;
    double Syx = Sxy;
    Syy *= c;
    double S2 = Sxx * Sxx + Sxy * Sxy + Syx * Syx + Syy * Syy;
    double Q = sqrt(2.0 * S2);
    double g1 = tau0 * tau0 + 18.0 * cs * cs * Q / (rho);
    return (sqrt(g1) + tau0) / 2.0;
}
IPPI_HD double les_tau(const double* f, double rho, double ux, double uy, double tau0, double cs) {
    double q[9];
    feq(rho, ux, uy, q);
    return les_tau_q(f, q, rho, tau0, cs);
}

}  // namespace ref

// ------------------------------------------------------------------ FP32
// Stored: s_k = f_k - c_k with c = feq(1, (U, 0)), the equilibrium of the
// free stream (U = the configured inflow velocity, 0 when there is none, in
// which case c = w). In a flow that is mostly free stream the stored values
// are small deviations, so their FP32 rounding is small in absolute terms:
// the storage error that each step re-injects shrinks by the ratio of
// |f - c| to |f - w|. Every FP32 formula works on deviations from that
// reference state: with drho = sum s and dj = sum s e,
//     rho = 1 + drho,   u = (U + dj_x, dj_y) / rho,
//     du = u - (U, 0) = ((dj_x - U drho) / rho, dj_y / rho),
// and the equilibrium, its moments and the non-equilibrium stress are
// written as differences from those of c, whose terms are all small
// (sum c = 1, sum c e = (U, 0), sum c e e = (U, 0)(U, 0) + I/3).
//
// Relaxation times are carried as tx = tau - 1/2, the quantity the
// viscosity is proportional to (nu = tx / 3). Near tau = 1/2 (high
// Reynolds numbers) tau and omega = 1/tau hold only a few significant
// digits of tx in FP32: rounding tau0 = 0.50192 to FP32 alone lowers the
// viscosity by 8e-6 relative, a bias that the flow integrates step after
// step. Every FP32 formula below takes tx and uses
//     dl = 2 - omega = 4 tx / (1 + 2 tx)   and   1 - omega / 2 = dl / 2,
// both with full relative precision.
namespace sh {

// hardware reciprocal / square root (1 ulp) on the device, exact on the host
IPPI_HD float rcp(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}
IPPI_HD float fsqrt(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_sqrtf(x);
#else
    return sqrtf(x);
#endif
}

// Equilibrium deviation q_k = feq_k(rho, u) - c_k for rho = 1 + drho and
// u = (U + dux, duy). With a = e.(U, 0), b = e.du and
// Q = |u|^2 - U^2 = dux (2 U + dux) + duy^2:
//     q_k = c_k drho + w_k rho (b (3 + 4.5 (2 a + b)) - 1.5 Q),
// every term small for a flow near the reference state. cf = c in FP32.
IPPI_HD void geq_dev(float drho, float rho, float dux, float duy, float U, const float* cf, float* q) {
#pragma clang fp contract(fast)
    const float Q = dux * (2.0f * U + dux) + duy * duy;
    const float hq = 1.5f * Q;
    const float bs = dux + duy, bd = dux - duy;
    auto one = [&](int k, float a, float b) {
        q[k] = cf[k] * drho + (WF[k] * rho) * (b * (3.0f + 4.5f * (2.0f * a + b)) - hq);
    };
    one(0, 0.0f, 0.0f);
    one(1, U, dux);
    one(3, -U, -dux);
    one(2, 0.0f, duy);
    one(4, 0.0f, -duy);
    one(5, U, bs);
    one(7, -U, -bs);
    one(8, U, bd);
    one(6, -U, -bd);
}

// geq_dev at drho = 0, rho = 1: the same values (c_k 0 + x = x, w_k 1 = w_k)
IPPI_HD void geq_dev0(float dux, float duy, float U, float* q) {
#pragma clang fp contract(fast)
    const float Q = dux * (2.0f * U + dux) + duy * duy;
    const float hq = 1.5f * Q;
    const float bs = dux + duy, bd = dux - duy;
    auto one = [&](int k, float a, float b) { q[k] = WF[k] * (b * (3.0f + 4.5f * (2.0f * a + b)) - hq); };
    one(0, 0.0f, 0.0f);
    one(1, U, dux);
    one(3, -U, -dux);
    one(2, 0.0f, duy);
    one(4, 0.0f, -duy);
    one(5, U, bs);
    one(7, -U, -bs);
    one(8, U, bd);
    one(6, -U, -bd);
}

// 2 - omega = 4 tx / (1 + 2 tx) for tau = 1/2 + tx: the hardware reciprocal
// and one Newton correction of the quotient (within an ulp of the correctly
// rounded division, in 4 instructions instead of 10)
IPPI_HD float dl_of(float tx) {
    const float a = 4.0f * tx, b = 1.0f + 2.0f * tx;
    const float r = rcp(b);
    const float q = a * r;
    return fmaf(fmaf(-b, q, a), r, q);
}

// Hybrid Smagorinsky tau (mesh.cpp:877-910) as tx = tau - 1/2:
// tau = (sqrt(tau0^2 + X) + tau0) / 2 with X = 18 Cs^2 Q / rho, written
// tx = tx0 + X / (2 (sqrt(tau0^2 + X) + tau0)) so that tx0 and the eddy
// part keep their digits.
IPPI_HD float les_tx(float pxx, float pxy, float pyy, float rho, float tx0, float cs) {
#pragma clang fp contract(fast)
    const float tau0 = 0.5f + tx0;
    float ir = rcp(rho);
    float c = 1.5f * ir * rcp(tau0);
    float S2 = c * c * (pxx * pxx + 2.0f * pxy * pxy + pyy * pyy);
    float Q = fsqrt(2.0f * S2);
    float X = 18.0f * cs * cs * Q * ir;
    return tx0 + X * rcp(2.0f * (fsqrt(tau0 * tau0 + X) + tau0));
}

// Set the mass and momentum of a collision increment d to their exact
// values: zero mass, momentum (tx, ty) (zero without forcing; with the
// immersed-boundary velocity shift u = (j + F/2)/rho it is omega F/2 for
// BGK and F/2 for RR). In FP32 the rounding of the equilibrium is nearly the
// same at every node of a near-uniform flow, so its residual would add up
// coherently over the domain and over time. The projection
// d -= w (m + 3 e.(j - t)) removes it (sum w = 1, sum w e e = I/3).
IPPI_HD void conserve(float* d, float tx, float ty) {
#pragma clang fp contract(fast)
    float c = (d[5] + d[6]) + (d[7] + d[8]);
    float m = d[0] + ((d[1] + d[3]) + (d[2] + d[4])) + c;
    float jx = (d[1] - d[3]) + ((d[5] - d[7]) + (d[8] - d[6]));
    float jy = (d[2] - d[4]) + ((d[5] - d[7]) + (d[6] - d[8]));
    jx = 3.0f * (jx - tx);
    jy = 3.0f * (jy - ty);
#pragma unroll
    for (int k = 0; k < 9; k++) d[k] -= WF[k] * (m + (float)EX[k] * jx + (float)EY[k] * jy);
}

// Regularized collision (NodeAndLattice.cpp:228-286): fcol = feq +
// (1 - omega) f1 with 1 - omega = dl - 1, taken as dl f1 - f1.
IPPI_HD void rr(float* g, const float* q, float pxx, float pxy, float pyy, float dl, float fx, float fy) {
#pragma clang fp contract(fast)
    const float third = 1.0f / 3.0f;
    float d[9];
#pragma unroll
    for (int k = 0; k < 9; k++) {
        float q0 = (float)(EX[k] * EX[k]) - third;
        float q3 = (float)(EY[k] * EY[k]) - third;
        float s = q0 * pxx + q3 * pyy;
        if (EX[k] * EY[k] != 0) s += 2.0f * (float)(EX[k] * EY[k]) * pxy;
        const float f1 = (4.5f * WF[k]) * s;
        d[k] = (q[k] - g[k]) + (dl * f1 - f1);
    }
    conserve(d, 0.5f * fx, 0.5f * fy);
#pragma unroll
    for (int k = 0; k < 9; k++) g[k] += d[k];
}

// Guo source (NodeAndLattice.cpp:162-195) with 1 - 1/(2 TauAtBeginning)
// = dl_begin / 2.
IPPI_HD void guo(float ux, float uy, float fx, float fy, float dl_begin, float* g) {
#pragma clang fp contract(fast)
    float omega = 0.5f * dl_begin;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        float ex = (float)EX[k], ey = (float)EY[k];
        float eu = ex * ux + ey * uy;
        float ef = ex * fx + ey * fy;
        float uf = ux * fx + uy * fy;
        g[k] += WF[k] * omega * (3.0f * (ef - uf) + 9.0f * eu * ef);
    }
}

}  // namespace sh
}  // namespace lbm
