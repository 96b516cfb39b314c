// libippi_lbm.so -- MI355X (gfx950) D2Q9 time step of IPPI-LBM behind the
// C ABI in include/ippi_lbm.h.
//
// One time step of the reference (Mesh::OneLBMStep, IPPI-LBM/mesh.cpp:357-513)
// is collide -> stream -> boundary conditions. Here the populations stored
// in HBM are the post-collision ones and one fused kernel does
//     pull-stream (+ boundary rule) -> moments -> [IB velocity correction]
//     -> relaxation time (const / field / LES / MLP) -> collision
//     [+ Guo forcing] -> store,
// i.e. the streaming of step t and the collision of step t+1, reading and
// writing 9 populations per node once (A/B buffers, SoA f[k][y][x]).
// The immersed boundary (mesh.cpp:514-832, 1244-1281, wired before the
// collision) runs as two small kernels before each fused step: marker
// interpolation + direct forcing, then a gather-style spread into a dense
// per-body patch that the fused kernel reads.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ippi_lbm.h"
#include "lbm_node.hpp"

using namespace lbm;

#define IPPI_MAXB 32

namespace {

struct Patch {
    int x0, y0, w, h, off;
};

struct StepArgs {
    int nx, ny;
    double tau0, cs, uinx, uiny;
    float tau0f, csf, uinxf, uinyf;
    int les;
    int tau_field;          // 1: per-node base tau from tau_ptr
    const void* tau_ptr;    // double* (FP64) / float* (FP32)
    const double* patch_f;  // [node][2]
    int* flag;              // divergence flag
    int nbody;
    Patch patch[IPPI_MAXB];
};

// ----------------------------------------------------------------- streaming
// Pull of the post-collision populations into node (x, y), with the
// boundary rules of the configured kind (not used for NEE columns).
// 32-bit index math: ippi_create limits 9*nx*ny to < 2^32.
template <typename T, int BC>
__device__ __forceinline__ void pull(const T* __restrict__ s, int x, int y, int nx, int ny, unsigned N, T* f) {
    const unsigned n = (unsigned)y * nx + x;
    if (BC == IPPI_BC_PERIODIC_BB) {
        const unsigned xm = x == 0 ? nx - 1 : x - 1;
        const unsigned xp = x == nx - 1 ? 0 : x + 1;
        const unsigned r0 = (unsigned)y * nx, rm = r0 - nx, rp = r0 + nx;
        const bool bot = y == 0, top = y == ny - 1;
        // halfway bounce-back below row 0 / above row ny-1
        const unsigned i2 = bot ? 4 * N + n : 2 * N + rm + x;
        const unsigned i5 = bot ? 7 * N + n : 5 * N + rm + xm;
        const unsigned i6 = bot ? 8 * N + n : 6 * N + rm + xp;
        const unsigned i4 = top ? 2 * N + n : 4 * N + rp + x;
        const unsigned i7 = top ? 5 * N + n : 7 * N + rp + xp;
        const unsigned i8 = top ? 6 * N + n : 8 * N + rp + xm;
        f[0] = s[n];
        f[1] = s[1 * N + r0 + xm];
        f[3] = s[3 * N + r0 + xp];
        f[2] = s[i2];
        f[4] = s[i4];
        f[5] = s[i5];
        f[6] = s[i6];
        f[7] = s[i7];
        f[8] = s[i8];
    } else {
        const unsigned r0 = (unsigned)y * nx;
        f[0] = s[n];
        f[1] = s[1 * N + r0 + x - 1];
        f[3] = s[3 * N + r0 + x + 1];
        if (y == 0) {  // specular free-slip row (mesh.cpp:325-337)
            const unsigned rp = r0 + nx;
            f[2] = s[4 * N + rp + x];
            f[4] = s[4 * N + rp + x];
            f[5] = s[8 * N + rp + x - 1];
            f[6] = s[7 * N + rp + x + 1];
            f[7] = s[7 * N + rp + x + 1];
            f[8] = s[8 * N + rp + x - 1];
        } else if (y == ny - 1) {  // mirror image for the top row
            const unsigned rm = r0 - nx;
            f[2] = s[2 * N + rm + x];
            f[4] = s[2 * N + rm + x];
            f[5] = s[5 * N + rm + x - 1];
            f[6] = s[6 * N + rm + x + 1];
            f[7] = s[6 * N + rm + x + 1];
            f[8] = s[5 * N + rm + x - 1];
        } else {
            const unsigned rm = r0 - nx, rp = r0 + nx;
            f[2] = s[2 * N + rm + x];
            f[4] = s[4 * N + rp + x];
            f[5] = s[5 * N + rm + x - 1];
            f[6] = s[6 * N + rm + x + 1];
            f[7] = s[7 * N + rp + x + 1];
            f[8] = s[8 * N + rp + x - 1];
        }
    }
}

// Post-streaming state of node (x, y): populations (full f for FP64, f - w
// for FP32) and moments. For IPPI_BC_INFLOW the columns 0 and nx-1 are
// non-equilibrium extrapolation nodes (mesh.cpp:239-286): rho = 1, inlet
// u = u_in, outlet u = u(nx-2), f = feq(1, u) + f_n - feq_n with n the
// neighbouring column.
struct Mom {
    double rho, ux, uy;
};

__device__ __forceinline__ void node_state(const double* __restrict__ s, int x, int y, int nx, int ny, unsigned N,
                                           int bc, double uinx, double uiny, double* f, double& rho, double& ux,
                                           double& uy) {
    if (bc == IPPI_BC_PERIODIC_BB) {
        pull<double, IPPI_BC_PERIODIC_BB>(s, x, y, nx, ny, N, f);
        ref::macro(f, rho, ux, uy);
        return;
    }
    const bool nee = (x == 0) || (x == nx - 1);
    const int xs = x == 0 ? 1 : (x == nx - 1 ? nx - 2 : x);
    pull<double, IPPI_BC_INFLOW>(s, xs, y, nx, ny, N, f);
    ref::macro(f, rho, ux, uy);
    if (nee) {
#pragma clang fp contract(off)
        double bx = x == 0 ? uinx : ux, by = x == 0 ? uiny : uy;
        double qb[9], qn[9];
        ref::feq(1.0, bx, by, qb);
        ref::feq(rho, ux, uy, qn);
#pragma unroll
        for (int k = 0; k < 9; k++) f[k] = qb[k] + f[k] - qn[k];
        rho = 1.0;
        ux = bx;
        uy = by;
    }
}

__device__ __forceinline__ void node_state(const float* __restrict__ s, int x, int y, int nx, int ny, unsigned N,
                                           int bc, float uinx, float uiny, float* g, float& drho, float& rho,
                                           float& ux, float& uy) {
#pragma clang fp contract(fast)
    if (bc == IPPI_BC_PERIODIC_BB) {
        pull<float, IPPI_BC_PERIODIC_BB>(s, x, y, nx, ny, N, g);
        sh::macro(g, drho, rho, ux, uy);
        return;
    }
    const bool nee = (x == 0) || (x == nx - 1);
    const int xs = x == 0 ? 1 : (x == nx - 1 ? nx - 2 : x);
    pull<float, IPPI_BC_INFLOW>(s, xs, y, nx, ny, N, g);
    sh::macro(g, drho, rho, ux, uy);
    if (nee) {
        float bx = x == 0 ? uinx : ux, by = x == 0 ? uiny : uy;
        float qb[9], qn[9];
        sh::geq(0.f, 1.f, bx, by, qb);
        sh::geq(drho, rho, ux, uy, qn);
#pragma unroll
        for (int k = 0; k < 9; k++) g[k] = qb[k] + (g[k] - qn[k]);
        drho = 0.f;
        rho = 1.f;
        ux = bx;
        uy = by;
    }
}

// IB force of node (x, y) from the per-body patches (the block-level test
// keeps the per-node test off all blocks that touch no patch).
__device__ __forceinline__ bool patch_force(const StepArgs& a, int x, int y, int bx0, int bx1, int by0, int by1,
                                            double& fx, double& fy) {
    for (int b = 0; b < a.nbody; ++b) {
        const Patch p = a.patch[b];
        if (bx1 < p.x0 || bx0 >= p.x0 + p.w || by1 < p.y0 || by0 >= p.y0 + p.h) continue;
        if (x >= p.x0 && x < p.x0 + p.w && y >= p.y0 && y < p.y0 + p.h) {
            const size_t i = (size_t)p.off + (size_t)(y - p.y0) * p.w + (x - p.x0);
            fx = a.patch_f[2 * i];
            fy = a.patch_f[2 * i + 1];
            return true;
        }
    }
    return false;
}

constexpr int BX = 256;  // nodes per block along x (one row segment)

// ------------------------------------------------------------- FP64 step
// PULL = false: collide the fnew state stored at the node itself with the
// moments from the macro arrays (first collision after initialisation).
template <int COLL, int BC, bool PULL>
__global__ void __launch_bounds__(BX) k_step_f64(const double* __restrict__ src, double* __restrict__ dst,
                                                 const double* __restrict__ mac, StepArgs a) {
#pragma clang fp contract(off)
    const int x = blockIdx.x * BX + threadIdx.x;
    const int y = blockIdx.y;
    const int nx = a.nx, ny = a.ny;
    if (x >= nx) return;
    const unsigned N = (unsigned)nx * ny;
    const unsigned n = (unsigned)y * nx + x;
    double f[9], rho, ux, uy;
    if (PULL) {
        node_state(src, x, y, nx, ny, N, BC, a.uinx, a.uiny, f, rho, ux, uy);
    } else {
#pragma unroll
        for (int k = 0; k < 9; k++) f[k] = src[k * N + n];
        rho = mac[n];
        ux = mac[N + n];
        uy = mac[2 * N + n];
    }
    if (!(rho >= 0.0 && rho <= 10.0)) *a.flag = 1;  // ErrorCatch, mesh.cpp:2101-2111
    double fx = 0.0, fy = 0.0;
    bool forced = false;
    if (a.nbody > 0) {
        const int bx0 = blockIdx.x * BX;
        forced = patch_force(a, x, y, bx0, bx0 + BX - 1, y, y, fx, fy);
        if (forced) {  // UpdateVelocity, mesh.cpp:1254-1255
            ux = ux + 0.5 * 1 * fx / rho;
            uy = uy + 0.5 * 1 * fy / rho;
        }
    }
    const double tb = a.tau_field ? static_cast<const double*>(a.tau_ptr)[n] : a.tau0;
    double tau = tb;
    if (a.les) tau = ref::les_tau(f, rho, ux, uy, tb, a.cs);
    double o[9];
    if (COLL == IPPI_COLL_MRT)
        ref::mrt(f, rho, ux, uy, tau, o);
    else if (COLL == IPPI_COLL_RR)
        ref::rr(f, rho, ux, uy, tau, o);
    else
        ref::bgk(f, rho, ux, uy, tau, o);
    if (forced) {
        double g[9];
        ref::guo(ux, uy, fx, fy, tb, g);
#pragma unroll
        for (int k = 0; k < 9; k++) o[k] = o[k] + g[k];
    }
#pragma unroll
    for (int k = 0; k < 9; k++) __builtin_nontemporal_store(o[k], &dst[k * N + n]);
}

// ------------------------------------------------------------- FP32 step
template <int COLL, int BC, bool PULL>
__global__ void __launch_bounds__(BX) k_step_f32(const float* __restrict__ src, float* __restrict__ dst,
                                                 const double* __restrict__ mac, StepArgs a) {
#pragma clang fp contract(fast)
    const int x = blockIdx.x * BX + threadIdx.x;
    const int y = blockIdx.y;
    const int nx = a.nx, ny = a.ny;
    if (x >= nx) return;
    const unsigned N = (unsigned)nx * ny;
    const unsigned n = (unsigned)y * nx + x;
    float g[9], drho, rho, ux, uy;
    if (PULL) {
        node_state(src, x, y, nx, ny, N, BC, a.uinxf, a.uinyf, g, drho, rho, ux, uy);
    } else {
#pragma unroll
        for (int k = 0; k < 9; k++) g[k] = src[k * N + n];
        rho = (float)mac[n];
        drho = (float)(mac[n] - 1.0);
        ux = (float)mac[N + n];
        uy = (float)mac[2 * N + n];
    }
    if (!(rho >= 0.f && rho <= 10.f)) *a.flag = 1;
    double fxd = 0.0, fyd = 0.0;
    bool forced = false;
    float fx = 0.f, fy = 0.f;
    if (a.nbody > 0) {
        const int bx0 = blockIdx.x * BX;
        forced = patch_force(a, x, y, bx0, bx0 + BX - 1, y, y, fxd, fyd);
        if (forced) {
            fx = (float)fxd;
            fy = (float)fyd;
            float ir = 0.5f * sh::rcp(rho);
            ux += fx * ir;
            uy += fy * ir;
        }
    }
    const float tb = a.tau_field ? static_cast<const float*>(a.tau_ptr)[n] : a.tau0f;
    float tau = tb;
    float q[9];
    sh::geq(drho, rho, ux, uy, q);
    if (COLL == IPPI_COLL_MRT) {
        if (a.les) {
            float pxx, pxy, pyy;
            sh::neq_stress(g, q, pxx, pxy, pyy);
            tau = sh::les_tau(pxx, pxy, pyy, rho, tb, a.csf);
        }
        sh::mrt(g, drho, rho, ux, uy, tau);
    } else if (COLL == IPPI_COLL_RR) {
        float pxx, pxy, pyy;
        sh::neq_stress(g, q, pxx, pxy, pyy);
        if (a.les) tau = sh::les_tau(pxx, pxy, pyy, rho, tb, a.csf);
        sh::rr(g, q, pxx, pxy, pyy, tau, fx, fy);
    } else {
        if (a.les) {
            float pxx, pxy, pyy;
            sh::neq_stress(g, q, pxx, pxy, pyy);
            tau = sh::les_tau(pxx, pxy, pyy, rho, tb, a.csf);
        }
        sh::bgk(g, q, tau, fx, fy);
    }
    if (forced) sh::guo(ux, uy, fx, fy, tb, g);
#pragma unroll
    for (int k = 0; k < 9; k++) __builtin_nontemporal_store(g[k], &dst[k * N + n]);
}

// ----------------------------------------------------------- init kernels
// rho, u from the macro arrays (mode 0) or from rho0/u0 + seeded noise
// (mode 1); f = feq (InitialLatticeFeq form), stored as the fnew state.
template <typename T>
__global__ void k_init(T* __restrict__ dst, double* __restrict__ mac, size_t N, int mode,
                       unsigned long long seed, double rho0, double ux0, double uy0, double amp) {
#pragma clang fp contract(off)
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    double rho, ux, uy;
    if (mode == 1) {
        rho = rho0;
        ux = ux0 + amp * noise(seed, (long long)n, 0);
        uy = uy0 + amp * noise(seed, (long long)n, 1);
        mac[n] = rho;
        mac[N + n] = ux;
        mac[2 * N + n] = uy;
    } else {
        rho = mac[n];
        ux = mac[N + n];
        uy = mac[2 * N + n];
    }
    double f[9];
    ref::feq_init(rho, ux, uy, f);
#pragma unroll
    for (int k = 0; k < 9; k++) {
        if constexpr (sizeof(T) == 8)
            dst[k * N + n] = f[k];
        else
            dst[k * N + n] = (float)(f[k] - WD[k]);
    }
}

// Post-streaming populations and moments of the current state (reads only);
// f_out may be null. FP32 moments are accumulated in FP64 from the stored
// deviations.
template <typename T, int BC>
__global__ void k_macro(const T* __restrict__ src, double* __restrict__ mac, double* __restrict__ f_out, int nx,
                        int ny, double uinx, double uiny) {
#pragma clang fp contract(off)
    const int x = blockIdx.x * BX + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= nx) return;
    const unsigned N = (unsigned)nx * ny;
    const unsigned n = (unsigned)y * nx + x;
    double f[9], rho, ux, uy;
    if constexpr (sizeof(T) == 8) {
        node_state(src, x, y, nx, ny, N, BC, uinx, uiny, f, rho, ux, uy);
    } else {
        float g[9], dr, r, vx, vy;
        node_state(src, x, y, nx, ny, N, BC, (float)uinx, (float)uiny, g, dr, r, vx, vy);
        double s = 0.0, jx, jy;
        double gd[9];
#pragma unroll
        for (int k = 0; k < 9; k++) gd[k] = (double)g[k];
        for (int k = 0; k < 9; k++) s += gd[k];
        jx = ((gd[1] - gd[3]) + (gd[5] - gd[7])) + (gd[8] - gd[6]);
        jy = ((gd[2] - gd[4]) + (gd[5] - gd[7])) + (gd[6] - gd[8]);
        const bool nee = BC == IPPI_BC_INFLOW && (x == 0 || x == nx - 1);
        rho = nee ? 1.0 : 1.0 + s;
        ux = nee ? (double)vx : jx / rho;
        uy = nee ? (double)vy : jy / rho;
#pragma unroll
        for (int k = 0; k < 9; k++) f[k] = gd[k] + WD[k];
    }
    mac[n] = rho;
    mac[N + n] = ux;
    mac[2 * N + n] = uy;
    if (f_out) {
#pragma unroll
        for (int k = 0; k < 9; k++) f_out[k * N + n] = f[k];
    }
}

// ------------------------------------------------------------ IB kernels
// InterpolateVel (4-point delta, mesh.cpp:555-605) + CalParticleForce
// (DirectForce, mesh.cpp:514-532). One thread per marker; the 16 weights
// phi(x - X) phi(y - Y) are precomputed on the host in the reference's
// X-outer / Y-inner order. Velocities at the support nodes come from the
// current state (pull) or from the macro arrays (first step).
struct MarkerDev {
    const int* base;    // [m][2]: XInt-1, YInt-1
    const double* w;    // [m][16]
    const double* ud;   // [m][2] desired velocity
    double* ui;         // [m][2] interpolated velocity
    double* F;          // [m][2] marker force
    int nm;
};

// One thread per (marker, support node): 16 markers per 256-thread block.
// Contributions u(X,Y) phi phi go to LDS and one thread per marker adds
// them in the reference's X-outer / Y-inner order.
constexpr int IB_MPB = 16;

template <typename T, int BC, bool PULL>
__global__ void __launch_bounds__(256) k_ib_markers(const T* __restrict__ src, const double* __restrict__ mac,
                                                    MarkerDev md, int nx, int ny) {
#pragma clang fp contract(off)
    __shared__ double cx[IB_MPB][16], cy[IB_MPB][16];
    const int lm = threadIdx.x >> 4, sidx = threadIdx.x & 15;
    const int m = blockIdx.x * IB_MPB + lm;
    const unsigned N = (unsigned)nx * ny;
    if (m < md.nm) {
        const int X = md.base[2 * m] + (sidx >> 2), Y = md.base[2 * m + 1] + (sidx & 3);
        const unsigned n = (unsigned)Y * nx + X;
        double ux, uy;
        if (PULL) {
            if constexpr (sizeof(T) == 8) {
                double f[9], r;
                pull<double, BC>(src, X, Y, nx, ny, N, f);
                ref::macro(f, r, ux, uy);
            } else {
                float g[9], dr, r, vx, vy;
                pull<float, BC>(src, X, Y, nx, ny, N, g);
                sh::macro(g, dr, r, vx, vy);
                ux = vx;
                uy = vy;
            }
        } else {
            if constexpr (sizeof(T) == 8) {
                ux = mac[N + n];
                uy = mac[2 * N + n];
            } else {
                ux = (double)(float)mac[N + n];
                uy = (double)(float)mac[2 * N + n];
            }
        }
        const double D = md.w[16 * m + sidx];
        cx[lm][sidx] = (ux * D);
        cy[lm][sidx] = (uy * D);
    }
    __syncthreads();
    if (m < md.nm && sidx == 0) {
        double uix = 0, uiy = 0;
#pragma unroll
        for (int s = 0; s < 16; s++) {
            uix += cx[lm][s];
            uiy += cy[lm][s];
        }
        md.ui[2 * m] = uix;
        md.ui[2 * m + 1] = uiy;
        md.F[2 * m] = 2 * (md.ud[2 * m] - uix);
        md.F[2 * m + 1] = 2 * (md.ud[2 * m + 1] - uiy);
    }
}

// SpreadForce (mesh.cpp:659-741) as a gather: each patch node sums its
// (marker, weight) list in increasing marker order, the order of the
// reference's marker loop. Block 0 also writes the per-body force record
// (minus the sum of the marker forces, in marker order), staging the
// marker forces through LDS in chunks.
constexpr int IB_CHUNK = 2048;

__global__ void __launch_bounds__(256) k_ib_spread(const int* __restrict__ ptr, const int* __restrict__ mk,
                                                   const double* __restrict__ w, const double* __restrict__ F,
                                                   double* __restrict__ patch_f, int npn,
                                                   const int* __restrict__ body_first, int nbody, int nm,
                                                   double* __restrict__ hist) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < npn) {
        double fx = 0.0, fy = 0.0;
        for (int e = ptr[i]; e < ptr[i + 1]; e++) {
            const int m = mk[e];
            fx += (F[2 * m] * w[e]);
            fy += (F[2 * m + 1] * w[e]);
        }
        patch_f[2 * i] = fx;
        patch_f[2 * i + 1] = fy;
    }
    if (hist && blockIdx.x == 0) {
        __shared__ double sf[2 * IB_CHUNK];
        const int b = threadIdx.x;
        double fx = 0.0, fy = 0.0;
        const int b0 = b < nbody ? body_first[b] : 0, b1 = b < nbody ? body_first[b + 1] : 0;
        for (int c0 = 0; c0 < nm; c0 += IB_CHUNK) {
            const int c1 = min(nm, c0 + IB_CHUNK);
            __syncthreads();
            for (int j = c0 * 2 + (int)threadIdx.x; j < c1 * 2; j += blockDim.x) sf[j - 2 * c0] = F[j];
            __syncthreads();
            if (b < nbody) {
                for (int m = max(b0, c0); m < min(b1, c1); m++) {
                    fx += sf[2 * (m - c0)];
                    fy += sf[2 * (m - c0) + 1];
                }
            }
        }
        if (b < nbody) {
            hist[2 * b] = -fx;
            hist[2 * b + 1] = -fy;
        }
    }
}

// Deterministic FP64 sum (fixed block partition, fixed tree).
__global__ void k_sum_partial(const double* __restrict__ v, size_t n, double* __restrict__ part) {
    __shared__ double sm[256];
    const size_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const size_t b0 = blockIdx.x * chunk, b1 = std::min(n, b0 + chunk);
    double s = 0.0;
    for (size_t i = b0 + threadIdx.x; i < b1; i += 256) s += v[i];
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

}  // namespace

// ====================================================================== host

struct ippi_handle {
    ippi_config c{};
    size_t N = 0;
    size_t esz = 8;
    hipStream_t s = nullptr;
    void* buf[2] = {nullptr, nullptr};
    int cur = 0;
    int phase = 0;  // 0: buf[cur] holds fnew + macro arrays; 1: buf[cur] holds fcol
    int64_t steps = 0;
    double* mac = nullptr;  // rho, ux, uy (3N)
    void* tau = nullptr;
    int* flag = nullptr;
    double* part = nullptr;
    // immersed boundary
    struct Body {
        int first, count;
        Patch p;
    };
    std::vector<Body> bodies;
    std::vector<double> mk_xy, mk_ud;
    std::vector<int> mk_base;
    std::vector<double> mk_w;
    bool ib_dirty = false;
    int *d_base = nullptr, *d_ptr = nullptr, *d_mk = nullptr, *d_bfirst = nullptr;
    double *d_w = nullptr, *d_ud = nullptr, *d_ui = nullptr, *d_F = nullptr, *d_cw = nullptr, *d_pf = nullptr;
    int npn = 0;
    double* hist = nullptr;
    int hist_cap = 0;
    int64_t nrec = 0;
    // optional per-launch timing (HIP events on this handle's stream)
    int timing = 0;
    std::vector<hipEvent_t> ev_step, ev_ib;  // start/end pairs
    std::string err;
};

namespace {

int fail(ippi_handle* h, int code, const std::string& msg) {
    if (h) h->err = msg;
    return code;
}

#define HIPC(h, expr)                                                                         \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(h, e_ == hipErrorOutOfMemory ? IPPI_ERR_NOMEM : IPPI_ERR_HIP,         \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                   \
    } while (0)

template <typename P>
void dfree(P*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

int check_flag(ippi_handle* h) {
    int v = 0;
    HIPC(h, hipMemcpyAsync(&v, h->flag, sizeof(int), hipMemcpyDeviceToHost, h->s));
    HIPC(h, hipStreamSynchronize(h->s));
    if (v) return fail(h, IPPI_ERR_DIVERGED, "density left [0, 10] or became NaN (ErrorCatch)");
    return IPPI_OK;
}

template <typename T>
using StepKernel = void (*)(const T*, T*, const double*, StepArgs);

template <typename T, bool PULL>
StepKernel<T> pick_step(int coll, int bc) {
#define IPPI_PICK(C, B)                                                                   \
    if (coll == C && bc == B) {                                                           \
        if constexpr (sizeof(T) == 8)                                                     \
            return (StepKernel<T>)k_step_f64<C, B, PULL>;                                 \
        else                                                                              \
            return (StepKernel<T>)k_step_f32<C, B, PULL>;                                 \
    }
    IPPI_PICK(IPPI_COLL_BGK, IPPI_BC_PERIODIC_BB)
    IPPI_PICK(IPPI_COLL_MRT, IPPI_BC_PERIODIC_BB)
    IPPI_PICK(IPPI_COLL_RR, IPPI_BC_PERIODIC_BB)
    IPPI_PICK(IPPI_COLL_BGK, IPPI_BC_INFLOW)
    IPPI_PICK(IPPI_COLL_MRT, IPPI_BC_INFLOW)
    IPPI_PICK(IPPI_COLL_RR, IPPI_BC_INFLOW)
#undef IPPI_PICK
    return nullptr;
}

StepArgs make_args(ippi_handle* h) {
    StepArgs a{};
    a.nx = h->c.nx;
    a.ny = h->c.ny;
    a.tau0 = h->c.tau0;
    a.cs = h->c.cs;
    a.uinx = h->c.u_in_x;
    a.uiny = h->c.u_in_y;
    a.tau0f = (float)h->c.tau0;
    a.csf = (float)h->c.cs;
    a.uinxf = (float)h->c.u_in_x;
    a.uinyf = (float)h->c.u_in_y;
    a.les = h->c.les;
    a.tau_field = h->c.tau_mode == IPPI_TAU_FIELD;
    a.tau_ptr = h->tau;
    a.patch_f = h->d_pf;
    a.flag = h->flag;
    a.nbody = (int)h->bodies.size();
    for (size_t b = 0; b < h->bodies.size(); b++) a.patch[b] = h->bodies[b].p;
    return a;
}

// Upload marker weights and the spread lists after bodies were added.
int upload_ib(ippi_handle* h) {
    if (!h->ib_dirty) return IPPI_OK;
    const int nm = (int)(h->mk_ud.size() / 2);
    const int nb = (int)h->bodies.size();
    // patch layout and gather lists
    int off = 0;
    for (auto& b : h->bodies) {
        b.p.off = off;
        off += b.p.w * b.p.h;
    }
    h->npn = off;
    std::vector<std::vector<std::pair<int, double>>> rows(off);
    for (int bi = 0; bi < nb; bi++) {
        const auto& b = h->bodies[bi];
        for (int m = b.first; m < b.first + b.count; m++) {
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++) {
                    const int X = h->mk_base[2 * m] + i, Y = h->mk_base[2 * m + 1] + j;
                    const int r = b.p.off + (Y - b.p.y0) * b.p.w + (X - b.p.x0);
                    rows[r].push_back({m, h->mk_w[16 * m + 4 * i + j]});
                }
        }
    }
    std::vector<int> ptr(off + 1, 0), mk;
    std::vector<double> cw;
    for (int r = 0; r < off; r++) {
        for (auto& e : rows[r]) {
            mk.push_back(e.first);
            cw.push_back(e.second);
        }
        ptr[r + 1] = (int)mk.size();
    }
    std::vector<int> bfirst(nb + 1);
    for (int bi = 0; bi < nb; bi++) bfirst[bi] = h->bodies[bi].first;
    bfirst[nb] = nm;
    dfree(h->d_base); dfree(h->d_ptr); dfree(h->d_mk); dfree(h->d_bfirst);
    dfree(h->d_w); dfree(h->d_ud); dfree(h->d_ui); dfree(h->d_F); dfree(h->d_cw); dfree(h->d_pf);
    dfree(h->hist);
    HIPC(h, hipMalloc(&h->d_base, sizeof(int) * 2 * nm));
    HIPC(h, hipMalloc(&h->d_w, sizeof(double) * 16 * nm));
    HIPC(h, hipMalloc(&h->d_ud, sizeof(double) * 2 * nm));
    HIPC(h, hipMalloc(&h->d_ui, sizeof(double) * 2 * nm));
    HIPC(h, hipMalloc(&h->d_F, sizeof(double) * 2 * nm));
    HIPC(h, hipMalloc(&h->d_ptr, sizeof(int) * (off + 1)));
    HIPC(h, hipMalloc(&h->d_mk, sizeof(int) * std::max<size_t>(1, mk.size())));
    HIPC(h, hipMalloc(&h->d_cw, sizeof(double) * std::max<size_t>(1, cw.size())));
    HIPC(h, hipMalloc(&h->d_pf, sizeof(double) * 2 * std::max(1, off)));
    HIPC(h, hipMalloc(&h->d_bfirst, sizeof(int) * (nb + 1)));
    h->hist_cap = h->c.force_history > 0 ? h->c.force_history : 65536;
    HIPC(h, hipMalloc(&h->hist, sizeof(double) * 2 * (size_t)nb * h->hist_cap));
    HIPC(h, hipMemcpy(h->d_base, h->mk_base.data(), sizeof(int) * 2 * nm, hipMemcpyHostToDevice));
    HIPC(h, hipMemcpy(h->d_w, h->mk_w.data(), sizeof(double) * 16 * nm, hipMemcpyHostToDevice));
    HIPC(h, hipMemcpy(h->d_ud, h->mk_ud.data(), sizeof(double) * 2 * nm, hipMemcpyHostToDevice));
    HIPC(h, hipMemset(h->d_ui, 0, sizeof(double) * 2 * nm));
    HIPC(h, hipMemset(h->d_F, 0, sizeof(double) * 2 * nm));
    HIPC(h, hipMemcpy(h->d_ptr, ptr.data(), sizeof(int) * (off + 1), hipMemcpyHostToDevice));
    if (!mk.empty()) {
        HIPC(h, hipMemcpy(h->d_mk, mk.data(), sizeof(int) * mk.size(), hipMemcpyHostToDevice));
        HIPC(h, hipMemcpy(h->d_cw, cw.data(), sizeof(double) * cw.size(), hipMemcpyHostToDevice));
    }
    HIPC(h, hipMemcpy(h->d_bfirst, bfirst.data(), sizeof(int) * (nb + 1), hipMemcpyHostToDevice));
    h->nrec = 0;
    h->ib_dirty = false;
    return IPPI_OK;
}

template <typename T>
int launch_ib(ippi_handle* h, bool pull) {
    const int nm = (int)(h->mk_ud.size() / 2);
    MarkerDev md{h->d_base, h->d_w, h->d_ud, h->d_ui, h->d_F, nm};
    const T* src = (const T*)h->buf[h->cur];
    const int nx = h->c.nx, ny = h->c.ny;
    dim3 gm((nm + IB_MPB - 1) / IB_MPB);
    if (h->c.bc == IPPI_BC_PERIODIC_BB) {
        if (pull)
            hipLaunchKernelGGL((k_ib_markers<T, IPPI_BC_PERIODIC_BB, true>), gm, dim3(256), 0, h->s, src, h->mac, md, nx, ny);
        else
            hipLaunchKernelGGL((k_ib_markers<T, IPPI_BC_PERIODIC_BB, false>), gm, dim3(256), 0, h->s, src, h->mac, md, nx, ny);
    } else {
        if (pull)
            hipLaunchKernelGGL((k_ib_markers<T, IPPI_BC_INFLOW, true>), gm, dim3(256), 0, h->s, src, h->mac, md, nx, ny);
        else
            hipLaunchKernelGGL((k_ib_markers<T, IPPI_BC_INFLOW, false>), gm, dim3(256), 0, h->s, src, h->mac, md, nx, ny);
    }
    HIPC(h, hipGetLastError());
    const int nb = (int)h->bodies.size();
    double* rec = h->hist + 2 * (size_t)nb * (size_t)(h->nrec % h->hist_cap);
    dim3 gs((std::max(h->npn, 1) + 255) / 256);
    hipLaunchKernelGGL(k_ib_spread, gs, dim3(256), 0, h->s, h->d_ptr, h->d_mk, h->d_cw, h->d_F, h->d_pf, h->npn,
                       h->d_bfirst, nb, nm, rec);
    HIPC(h, hipGetLastError());
    h->nrec++;
    return IPPI_OK;
}

hipEvent_t new_event(ippi_handle* h) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    (void)hipEventRecord(e, h->s);
    return e;
}

void clear_events(std::vector<hipEvent_t>& v) {
    for (auto e : v)
        if (e) (void)hipEventDestroy(e);
    v.clear();
}

template <typename T>
int run_steps(ippi_handle* h, int nsteps) {
    const int nx = h->c.nx, ny = h->c.ny;
    dim3 grid((nx + BX - 1) / BX, ny);
    const bool ib = !h->bodies.empty();
    if (ib) {
        int r = upload_ib(h);
        if (r) return r;
    }
    StepArgs a = make_args(h);
    for (int t = 0; t < nsteps; t++) {
        const bool pull = h->phase == 1;
        if (ib) {
            if (h->timing) h->ev_ib.push_back(new_event(h));
            int r = launch_ib<T>(h, pull);
            if (r) return r;
            if (h->timing) h->ev_ib.push_back(new_event(h));
        }
        StepKernel<T> k = pull ? pick_step<T, true>(h->c.collision, h->c.bc) : pick_step<T, false>(h->c.collision, h->c.bc);
        if (!k) return fail(h, IPPI_ERR_ARG, "unsupported collision / boundary combination");
        if (h->timing && pull) h->ev_step.push_back(new_event(h));
        hipLaunchKernelGGL(k, grid, dim3(BX), 0, h->s, (const T*)h->buf[h->cur], (T*)h->buf[h->cur ^ 1], h->mac, a);
        HIPC(h, hipGetLastError());
        if (h->timing && pull) h->ev_step.push_back(new_event(h));
        h->cur ^= 1;
        h->phase = 1;
        h->steps++;
    }
    return IPPI_OK;
}

template <typename T>
int run_macro(ippi_handle* h, double* f_out) {
    if (h->phase == 0) {
        if (f_out) {  // fnew state stored directly
            std::vector<T> tmp(9 * h->N);
            HIPC(h, hipMemcpyAsync(tmp.data(), h->buf[h->cur], sizeof(T) * 9 * h->N, hipMemcpyDeviceToHost, h->s));
            HIPC(h, hipStreamSynchronize(h->s));
            for (size_t k = 0; k < 9; k++)
                for (size_t n = 0; n < h->N; n++)
                    f_out[k * h->N + n] = sizeof(T) == 8 ? (double)tmp[k * h->N + n] : (double)tmp[k * h->N + n] + WD[k];
        }
        return IPPI_OK;
    }
    double* d_f = nullptr;
    if (f_out) HIPC(h, hipMalloc(&d_f, sizeof(double) * 9 * h->N));
    dim3 grid((h->c.nx + BX - 1) / BX, h->c.ny);
    if (h->c.bc == IPPI_BC_PERIODIC_BB)
        hipLaunchKernelGGL((k_macro<T, IPPI_BC_PERIODIC_BB>), grid, dim3(BX), 0, h->s, (const T*)h->buf[h->cur], h->mac,
                           d_f, h->c.nx, h->c.ny, h->c.u_in_x, h->c.u_in_y);
    else
        hipLaunchKernelGGL((k_macro<T, IPPI_BC_INFLOW>), grid, dim3(BX), 0, h->s, (const T*)h->buf[h->cur], h->mac,
                           d_f, h->c.nx, h->c.ny, h->c.u_in_x, h->c.u_in_y);
    HIPC(h, hipGetLastError());
    if (f_out) {
        HIPC(h, hipMemcpyAsync(f_out, d_f, sizeof(double) * 9 * h->N, hipMemcpyDeviceToHost, h->s));
        HIPC(h, hipStreamSynchronize(h->s));
        dfree(d_f);
    }
    return IPPI_OK;
}

int macro(ippi_handle* h, double* f_out) {
    return h->esz == 8 ? run_macro<double>(h, f_out) : run_macro<float>(h, f_out);
}

}  // namespace

// ================================================================== C ABI

extern "C" {

int ippi_version(void) { return 1; }

const char* ippi_last_error(const ippi_handle* h) { return h ? h->err.c_str() : "null handle"; }

int ippi_create(const ippi_config* cfg, ippi_handle** out) {
    if (!cfg || !out) return IPPI_ERR_ARG;
    *out = nullptr;
    if (cfg->nx < 4 || cfg->ny < 4 || cfg->collision < 0 || cfg->collision > 2 || cfg->precision < 0 ||
        cfg->precision > 1 || cfg->bc < 0 || cfg->bc > 1 || cfg->tau_mode < 0 || cfg->tau_mode > 2 ||
        !(cfg->tau0 > 0.5))
        return IPPI_ERR_ARG;
    if (9 * (uint64_t)cfg->nx * (uint64_t)cfg->ny >= ((uint64_t)1 << 32)) return IPPI_ERR_ARG;
    ippi_handle* h = new ippi_handle();
    h->c = *cfg;
    h->N = (size_t)cfg->nx * cfg->ny;
    h->esz = cfg->precision == IPPI_FP64 ? 8 : 4;
    auto bail = [&](int code) {
        ippi_destroy(h);
        return code;
    };
    if (hipSetDevice(cfg->device) != hipSuccess) return bail(IPPI_ERR_HIP);
    if (hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking) != hipSuccess) return bail(IPPI_ERR_HIP);
    for (int i = 0; i < 2; i++)
        if (hipMalloc(&h->buf[i], h->esz * 9 * h->N) != hipSuccess) return bail(IPPI_ERR_NOMEM);
    if (hipMalloc(&h->mac, sizeof(double) * 3 * h->N) != hipSuccess) return bail(IPPI_ERR_NOMEM);
    if (hipMalloc(&h->flag, sizeof(int)) != hipSuccess) return bail(IPPI_ERR_NOMEM);
    if (hipMalloc(&h->part, sizeof(double) * 1024) != hipSuccess) return bail(IPPI_ERR_NOMEM);
    if (hipMemset(h->flag, 0, sizeof(int)) != hipSuccess) return bail(IPPI_ERR_HIP);
    *out = h;
    return IPPI_OK;
}

int ippi_destroy(ippi_handle* h) {
    if (!h) return IPPI_OK;
    if (h->s) (void)hipStreamSynchronize(h->s);
    clear_events(h->ev_step);
    clear_events(h->ev_ib);
    dfree(h->buf[0]); dfree(h->buf[1]); dfree(h->mac); dfree(h->tau); dfree(h->flag); dfree(h->part);
    dfree(h->d_base); dfree(h->d_ptr); dfree(h->d_mk); dfree(h->d_bfirst);
    dfree(h->d_w); dfree(h->d_ud); dfree(h->d_ui); dfree(h->d_F); dfree(h->d_cw); dfree(h->d_pf);
    dfree(h->hist);
    if (h->s) (void)hipStreamDestroy(h->s);
    delete h;
    return IPPI_OK;
}

static int init_common(ippi_handle* h, int mode, uint64_t seed, double rho0, double ux0, double uy0, double amp) {
    dim3 g((unsigned)((h->N + 255) / 256));
    if (h->esz == 8)
        hipLaunchKernelGGL(k_init<double>, g, dim3(256), 0, h->s, (double*)h->buf[0], h->mac, h->N, mode,
                           (unsigned long long)seed, rho0, ux0, uy0, amp);
    else
        hipLaunchKernelGGL(k_init<float>, g, dim3(256), 0, h->s, (float*)h->buf[0], h->mac, h->N, mode,
                           (unsigned long long)seed, rho0, ux0, uy0, amp);
    HIPC(h, hipGetLastError());
    HIPC(h, hipMemsetAsync(h->flag, 0, sizeof(int), h->s));
    h->cur = 0;
    h->phase = 0;
    h->steps = 0;
    h->nrec = 0;
    return IPPI_OK;
}

int ippi_init_fields(ippi_handle* h, uint64_t seed, double rho0, double ux0, double uy0, double noise_amp) {
    if (!h) return IPPI_ERR_ARG;
    return init_common(h, 1, seed, rho0, ux0, uy0, noise_amp);
}

int ippi_set_fields(ippi_handle* h, const double* rho, const double* ux, const double* uy) {
    if (!h || !rho || !ux || !uy) return IPPI_ERR_ARG;
    HIPC(h, hipMemcpyAsync(h->mac, rho, sizeof(double) * h->N, hipMemcpyHostToDevice, h->s));
    HIPC(h, hipMemcpyAsync(h->mac + h->N, ux, sizeof(double) * h->N, hipMemcpyHostToDevice, h->s));
    HIPC(h, hipMemcpyAsync(h->mac + 2 * h->N, uy, sizeof(double) * h->N, hipMemcpyHostToDevice, h->s));
    int r = init_common(h, 0, 0, 0, 0, 0, 0);
    if (r) return r;
    HIPC(h, hipStreamSynchronize(h->s));
    return IPPI_OK;
}

int ippi_add_body(ippi_handle* h, double cx, double cy, double diameter, int M, double udx, double udy, int* body) {
    if (!h || M <= 0 || !(diameter > 0)) return IPPI_ERR_ARG;
    if ((int)h->bodies.size() >= IPPI_MAXB) return fail(h, IPPI_ERR_ARG, "too many bodies");
    const double R = 0.5 * diameter;
    ippi_handle::Body b{};
    b.first = (int)(h->mk_ud.size() / 2);
    b.count = M;
    int x0 = 1 << 30, x1 = -(1 << 30), y0 = 1 << 30, y1 = -(1 << 30);
    std::vector<double> xy(2 * M);
    std::vector<int> base(2 * M);
    std::vector<double> w(16 * M);
    for (int i = 0; i < M; i++) {
        // InitialIBParticle, mesh.cpp:180-181
        const double x = cx + R * sin(2. * M_PI * (double)i / M);
        const double y = cy + R * cos(2. * M_PI * (double)i / M);
        if (!(x >= 0 && y >= 0)) return fail(h, IPPI_ERR_ARG, "marker outside the domain");
        const int XI = (int)(x), YI = (int)(y);
        xy[2 * i] = x;
        xy[2 * i + 1] = y;
        base[2 * i] = XI - 1;
        base[2 * i + 1] = YI - 1;
        for (int a = 0; a < 4; a++)
            for (int c = 0; c < 4; c++) {
#pragma clang fp contract(off)
                const double dx = x - (XI - 1 + a), dy = y - (YI - 1 + c);
                w[16 * i + 4 * a + c] = delta4(dx) * delta4(dy);
            }
        x0 = std::min(x0, XI - 1);
        x1 = std::max(x1, XI + 2);
        y0 = std::min(y0, YI - 1);
        y1 = std::max(y1, YI + 2);
    }
    if (x0 < 1 || y0 < 1 || x1 > h->c.nx - 2 || y1 > h->c.ny - 2)
        return fail(h, IPPI_ERR_ARG, "immersed-boundary support must lie in the interior");
    b.p = Patch{x0, y0, x1 - x0 + 1, y1 - y0 + 1, 0};
    for (const auto& o : h->bodies) {
        const Patch& q = o.p;
        if (!(b.p.x0 + b.p.w <= q.x0 || q.x0 + q.w <= b.p.x0 || b.p.y0 + b.p.h <= q.y0 || q.y0 + q.h <= b.p.y0))
            return fail(h, IPPI_ERR_ARG, "body supports overlap");
    }
    h->bodies.push_back(b);
    h->mk_xy.insert(h->mk_xy.end(), xy.begin(), xy.end());
    h->mk_base.insert(h->mk_base.end(), base.begin(), base.end());
    h->mk_w.insert(h->mk_w.end(), w.begin(), w.end());
    for (int i = 0; i < M; i++) {
        h->mk_ud.push_back(udx);
        h->mk_ud.push_back(udy);
    }
    h->ib_dirty = true;
    if (body) *body = (int)h->bodies.size() - 1;
    return IPPI_OK;
}

int ippi_num_bodies(const ippi_handle* h, int* n) {
    if (!h || !n) return IPPI_ERR_ARG;
    *n = (int)h->bodies.size();
    return IPPI_OK;
}

int ippi_set_tau_field(ippi_handle* h, const double* tau) {
    if (!h || !tau) return IPPI_ERR_ARG;
    if (h->c.tau_mode != IPPI_TAU_FIELD) return fail(h, IPPI_ERR_STATE, "tau_mode is not IPPI_TAU_FIELD");
    for (size_t n = 0; n < h->N; n++)
        if (!(tau[n] > 0.5)) return fail(h, IPPI_ERR_ARG, "tau field must exceed 0.5");
    if (!h->tau) HIPC(h, hipMalloc(&h->tau, h->esz * h->N));
    if (h->esz == 8) {
        HIPC(h, hipMemcpy(h->tau, tau, sizeof(double) * h->N, hipMemcpyHostToDevice));
    } else {
        std::vector<float> t(h->N);
        for (size_t n = 0; n < h->N; n++) t[n] = (float)tau[n];
        HIPC(h, hipMemcpy(h->tau, t.data(), sizeof(float) * h->N, hipMemcpyHostToDevice));
    }
    return IPPI_OK;
}

int ippi_set_mlp(ippi_handle* h, const double* blob, int n) {
    if (!h || !blob) return IPPI_ERR_ARG;
    (void)n;
    return fail(h, IPPI_ERR_STATE, "relaxation-time MLP not available in this build");
}

int ippi_step(ippi_handle* h, int nsteps) {
    if (!h || nsteps < 0) return IPPI_ERR_ARG;
    if (h->c.tau_mode == IPPI_TAU_FIELD && !h->tau) return fail(h, IPPI_ERR_STATE, "tau field not set");
    if (h->c.tau_mode == IPPI_TAU_NN) return fail(h, IPPI_ERR_STATE, "relaxation-time MLP not set");
    return h->esz == 8 ? run_steps<double>(h, nsteps) : run_steps<float>(h, nsteps);
}

int ippi_set_timing(ippi_handle* h, int enable) {
    if (!h) return IPPI_ERR_ARG;
    HIPC(h, hipStreamSynchronize(h->s));
    clear_events(h->ev_step);
    clear_events(h->ev_ib);
    h->timing = enable ? 1 : 0;
    return IPPI_OK;
}

int ippi_get_timing(ippi_handle* h, double* step_ms, int64_t* step_launches, double* ib_ms, int64_t* ib_launches) {
    if (!h) return IPPI_ERR_ARG;
    HIPC(h, hipStreamSynchronize(h->s));
    auto sum = [&](std::vector<hipEvent_t>& v, double* ms, int64_t* cnt) -> int {
        double t = 0.0;
        int64_t c = 0;
        for (size_t i = 0; i + 1 < v.size(); i += 2) {
            if (!v[i] || !v[i + 1]) return fail(h, IPPI_ERR_HIP, "event creation failed");
            float e = 0.f;
            HIPC(h, hipEventElapsedTime(&e, v[i], v[i + 1]));
            t += e;
            c++;
        }
        if (ms) *ms = t;
        if (cnt) *cnt = c;
        clear_events(v);
        return IPPI_OK;
    };
    int r = sum(h->ev_step, step_ms, step_launches);
    if (r) return r;
    return sum(h->ev_ib, ib_ms, ib_launches);
}

int ippi_synchronize(ippi_handle* h) {
    if (!h) return IPPI_ERR_ARG;
    HIPC(h, hipStreamSynchronize(h->s));
    return check_flag(h);
}

int ippi_steps_done(const ippi_handle* h, int64_t* steps) {
    if (!h || !steps) return IPPI_ERR_ARG;
    *steps = h->steps;
    return IPPI_OK;
}

int ippi_get_fields(ippi_handle* h, double* rho, double* ux, double* uy) {
    if (!h) return IPPI_ERR_ARG;
    int r = macro(h, nullptr);
    if (r) return r;
    if (rho) HIPC(h, hipMemcpyAsync(rho, h->mac, sizeof(double) * h->N, hipMemcpyDeviceToHost, h->s));
    if (ux) HIPC(h, hipMemcpyAsync(ux, h->mac + h->N, sizeof(double) * h->N, hipMemcpyDeviceToHost, h->s));
    if (uy) HIPC(h, hipMemcpyAsync(uy, h->mac + 2 * h->N, sizeof(double) * h->N, hipMemcpyDeviceToHost, h->s));
    HIPC(h, hipStreamSynchronize(h->s));
    return check_flag(h);
}

int ippi_get_populations(ippi_handle* h, double* f) {
    if (!h || !f) return IPPI_ERR_ARG;
    int r = macro(h, f);
    if (r) return r;
    return check_flag(h);
}

int ippi_get_body_forces(ippi_handle* h, double* out, int max_records, int* nrec) {
    if (!h || !nrec) return IPPI_ERR_ARG;
    const int nb = (int)h->bodies.size();
    const int64_t have = std::min<int64_t>(h->nrec, h->hist_cap);
    const int n = (int)std::min<int64_t>(have, max_records);
    *nrec = n;
    if (n == 0 || nb == 0 || !out) return IPPI_OK;
    std::vector<double> all(2 * (size_t)nb * h->hist_cap);
    HIPC(h, hipStreamSynchronize(h->s));
    HIPC(h, hipMemcpy(all.data(), h->hist, sizeof(double) * all.size(), hipMemcpyDeviceToHost));
    // the last n records, oldest first
    for (int i = 0; i < n; i++) {
        const int64_t rec = h->nrec - n + i;
        const size_t slot = (size_t)(rec % h->hist_cap);
        std::memcpy(out + 2 * (size_t)nb * i, all.data() + 2 * (size_t)nb * slot, sizeof(double) * 2 * nb);
    }
    return check_flag(h);
}

int ippi_get_markers(ippi_handle* h, double* out, int max_markers, int* n) {
    if (!h || !n) return IPPI_ERR_ARG;
    const int nm = (int)(h->mk_ud.size() / 2);
    *n = std::min(nm, max_markers);
    if (!out || *n == 0 || !h->d_ui) return IPPI_OK;
    std::vector<double> ui(2 * nm), F(2 * nm);
    HIPC(h, hipStreamSynchronize(h->s));
    HIPC(h, hipMemcpy(ui.data(), h->d_ui, sizeof(double) * 2 * nm, hipMemcpyDeviceToHost));
    HIPC(h, hipMemcpy(F.data(), h->d_F, sizeof(double) * 2 * nm, hipMemcpyDeviceToHost));
    for (int m = 0; m < *n; m++) {
        out[6 * m + 0] = h->mk_xy[2 * m];
        out[6 * m + 1] = h->mk_xy[2 * m + 1];
        out[6 * m + 2] = ui[2 * m];
        out[6 * m + 3] = ui[2 * m + 1];
        out[6 * m + 4] = F[2 * m];
        out[6 * m + 5] = F[2 * m + 1];
    }
    return IPPI_OK;
}

int ippi_mass(ippi_handle* h, double* mass) {
    if (!h || !mass) return IPPI_ERR_ARG;
    int r = macro(h, nullptr);
    if (r) return r;
    const int nblk = 1024;
    hipLaunchKernelGGL(k_sum_partial, dim3(nblk), dim3(256), 0, h->s, (const double*)h->mac, h->N, h->part);
    HIPC(h, hipGetLastError());
    std::vector<double> p(nblk);
    HIPC(h, hipMemcpyAsync(p.data(), h->part, sizeof(double) * nblk, hipMemcpyDeviceToHost, h->s));
    HIPC(h, hipStreamSynchronize(h->s));
    double s = 0.0;
    for (double v : p) s += v;
    *mass = s;
    return check_flag(h);
}

// Checkpoint layout: header (magic, nx, ny, precision, cur, phase, steps,
// nrec), populations of buf[cur], macro arrays.
struct SaveHdr {
    uint64_t magic;
    int32_t nx, ny, precision, phase;
    int64_t steps, nrec;
};

int ippi_save(ippi_handle* h, void* buf, size_t* size) {
    if (!h || !size) return IPPI_ERR_ARG;
    const size_t need = sizeof(SaveHdr) + h->esz * 9 * h->N + sizeof(double) * 3 * h->N;
    if (!buf) {
        *size = need;
        return IPPI_OK;
    }
    if (*size < need) return fail(h, IPPI_ERR_ARG, "checkpoint buffer too small");
    SaveHdr hd{0x4950504931ull, h->c.nx, h->c.ny, h->c.precision, h->phase, h->steps, h->nrec};
    char* p = (char*)buf;
    std::memcpy(p, &hd, sizeof(hd));
    HIPC(h, hipStreamSynchronize(h->s));
    HIPC(h, hipMemcpy(p + sizeof(hd), h->buf[h->cur], h->esz * 9 * h->N, hipMemcpyDeviceToHost));
    HIPC(h, hipMemcpy(p + sizeof(hd) + h->esz * 9 * h->N, h->mac, sizeof(double) * 3 * h->N, hipMemcpyDeviceToHost));
    *size = need;
    return IPPI_OK;
}

int ippi_load(ippi_handle* h, const void* buf, size_t size) {
    if (!h || !buf) return IPPI_ERR_ARG;
    const size_t need = sizeof(SaveHdr) + h->esz * 9 * h->N + sizeof(double) * 3 * h->N;
    if (size < need) return fail(h, IPPI_ERR_ARG, "checkpoint too small");
    SaveHdr hd;
    std::memcpy(&hd, buf, sizeof(hd));
    if (hd.magic != 0x4950504931ull || hd.nx != h->c.nx || hd.ny != h->c.ny || hd.precision != h->c.precision)
        return fail(h, IPPI_ERR_ARG, "checkpoint does not match this configuration");
    const char* p = (const char*)buf;
    HIPC(h, hipStreamSynchronize(h->s));
    h->cur = 0;
    HIPC(h, hipMemcpy(h->buf[0], p + sizeof(hd), h->esz * 9 * h->N, hipMemcpyHostToDevice));
    HIPC(h, hipMemcpy(h->mac, p + sizeof(hd) + h->esz * 9 * h->N, sizeof(double) * 3 * h->N, hipMemcpyHostToDevice));
    HIPC(h, hipMemset(h->flag, 0, sizeof(int)));
    h->phase = hd.phase;
    h->steps = hd.steps;
    return IPPI_OK;
}

}  // extern "C"
