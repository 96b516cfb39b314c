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
// collision) runs inside the fused step kernel where the bodies allow it
// (k_step_f32_ib / k_step_f64_ib: one block per body computes the marker
// velocities and forces while the lattice rows stream, and the rows of the
// bodies' patches wait for them), otherwise as one small kernel before each
// fused step (k_ib). Either way the step kernel gathers each patch node's
// force from the marker forces (the spread, mesh.cpp:659-741).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/ippi_lbm.h"
#include "adjoint.hpp"
#include "lbm_node.hpp"
#include "mlp.hpp"

using namespace lbm;

#define IPPI_MAXB 32
// waves per SIMD the network variants of the step kernels are compiled for
#ifndef IPPI_NN_WAVES
#define IPPI_NN_WAVES 4
#endif
#ifndef IPPI_INK_MAXACT_NN
#define IPPI_INK_MAXACT_NN 3072
#endif

namespace {

struct Patch {
    int x0, y0, w, h, off;
    int roff;  // offset of the body's region (patch grown by one node) in the region numbering
    int m0;    // first marker of the body
};

// Reference state of the FP32 storage (s = f - c, c = feq(1, (U, 0)); see
// lbm_node.hpp): U in both precisions, c in both precisions, and the
// correction bb = c_7 - c_5 that halfway bounce-back adds where it returns
// a diagonal population reversed (c_5 = c_8 and c_6 = c_7, so a reversed
// population changes its offset by +-bb; the axis ones keep theirs).
struct RefState {
    double U;
    float Uf;
    float bb;
    double c[9];
    float cf[9];
};

struct StepArgs {
    int nx, ny;
    RefState ref;           // FP32 storage offsets
    float duinxf;           // FP32 inlet velocity minus U
    float qin[9];           // FP32 inlet equilibrium deviation feq(1, u_in) - c, computed once on the device
    double tau0, cs, uinx, uiny;
    float tx0f, csf, uinxf, uinyf;  // FP32 path: tx0f = tau0 - 1/2
    int les;
    int tau_field;          // 1: per-node base tau from tau_ptr
    const void* tau_ptr;    // double* tau (FP64) / float* tau - 1/2 (FP32)
    const double* patch_f;  // [node][2]
    int* flag;              // divergence flag
    int nbody;
    int nt;                 // non-temporal stores
    int round32;            // FP64 path: each stored population rounded as the FP32 path stores it
                            // (f - w to float and back): a diagnostic of the storage format alone
    double nn_max;          // upper bound of accepted network values
    float nn_maxtxf;        // nn_max - 1/2
    MlpDev mlp;             // relaxation-time network (tau_mode NN)
    int plate_x;            // IPPI_BC_FLATPLATE: first no-slip column of the bottom wall
    const double* corner;   // IPPI_BC_FLATPLATE: populations of (0, 0) and (nx-1, 0) ([2][9], full f)
    const double* mac;      // macro arrays (rho, ux, uy)
    // populations the next IB kernel pulls, copied into slot[k][active node]
    // by the node they come from (scat = nullptr: none)
    const int* scat;        // [9][region node] active node pulling population k from it, or -1
    void* slot;             // [9][active] (the population type)
    int nreg, nact;
    // patch forces gathered from the marker forces F (ib_gather = 1) instead
    // of read from patch_f: pslot maps a patch node to its active node (-1:
    // no marker reaches it), whose gather list (ecnt, emk, ecw, row width
    // ell) names the markers in increasing order
    int ib_gather, ell;
    const int* pslot;
    const int* ecnt;
    const int* emk;
    const double* ecw;
    const double* F;
    // IB inside the step kernel (k_step_f32_ib): row 0 of the grid computes
    // the marker forces, the rows of the band [band0, band1) wait for them
    const struct IbBlk* ib_lead;  // [body] the body's active-node range and markers
    const int* ib_act;            // active node -> patch node
    const int* ib_pidx;
    const double* ib_w;
    const double* ib_ud;
    double* ib_ui;
    double* ib_hist;        // [body][2] of this step's record
    unsigned* ib_flag;      // [body] sequence number of the last step whose forces are stored
    unsigned ib_seq;        // this step's sequence number
    unsigned ib_wait;       // the value the band waits for (ib_seq; ib_seq + 1 only to test the time-out)
    int band0, band1;
    // boundary-row / column networks (the reference's NNmodel_Type01/03/05/07):
    // accepted tau of this step (0: rejected) at row 1 [nx], row ny-2 [nx],
    // column 0 [ny], column nx-1 [ny]; bnn_on: bit t set for type 2t+1
    const double* bnn;
    int bnn_on;
    int rot;                // lattice row of grid row 1 (the rows follow cyclically)
    int xcd;                // XCD-grouped block order (xcd_block)
    int ink_bodies;         // k_step_f32_ib / k_step_f64_ib: body blocks (0: forces from k_ib)
    int brows;              // grid rows of the body blocks before the lattice rows (the lattice's own width)
    const unsigned* rowmask;  // [ny] bit b set where body b's patch covers the row (nullptr: test every body)
    Patch patch[IPPI_MAXB];
};

// Accepted boundary-network tau of node (x, y) (0: none): the row networks
// (T07 at row 1, T03 at row ny-2) decide where they meet a column network,
// as their calls come after T01 and T05 in the reference's step.
__device__ __forceinline__ double bnn_tau_at(const StepArgs& a, int x, int y) {
    const int nx = a.nx, ny = a.ny;
    if ((a.bnn_on & 8) && y == 1) return a.bnn[x];
    if ((a.bnn_on & 2) && y == ny - 2) return a.bnn[nx + x];
    if (y >= 1 && y <= ny - 2) {
        if ((a.bnn_on & 1) && x == 0) return a.bnn[2 * nx + y];
        if ((a.bnn_on & 4) && x == nx - 1) return a.bnn[2 * nx + ny + y];
    }
    return 0.0;
}

// ----------------------------------------------------------------- streaming
// Loads use a wave-uniform base pointer (one population array, in SGPRs) and
// a 32-bit per-lane element offset, so no 64-bit address arithmetic is done
// per lane. ippi_create limits 9*nx*ny to < 2^32.
template <typename T>
__device__ __forceinline__ T ld(const T* __restrict__ base, unsigned idx) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + (size_t)(idx * (unsigned)sizeof(T)));
}
// Stores: up to twice the 256 MB MALL (Infinity Cache) for both population
// buffers the next step still finds much of what this one writes there, so
// the stores allocate; beyond that they stream past the caches
// (non-temporal). Measured: non-temporal costs 43% at 151 MB and 4% at
// 302 MB, and gains 2% at 604 MB and above.
// The rule is a template parameter: with a runtime flag the compiler merges
// the two identical stores and drops the non-temporal hint.
template <bool NT, typename T>
__device__ __forceinline__ void st(T* __restrict__ base, unsigned idx, T v) {
    T* p = reinterpret_cast<T*>(reinterpret_cast<char*>(base) + (size_t)(idx * (unsigned)sizeof(T)));
    if (NT)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

// The nine populations of node n.
template <bool NT, typename T>
__device__ __forceinline__ void st9(T* __restrict__ dst, unsigned N, unsigned n, const T* v) {
#pragma unroll
    for (int k = 0; k < 9; k++) st<NT>(dst + k * N, n, v[k]);
}

// Pull of the post-collision populations into node (x, y), with the
// boundary rules of the configured kind (not used for NEE columns). Each
// population's source (array, row, column offset) is chosen with
// row-uniform selects, so the nine loads issue back to back with no branch
// between them (y is uniform across a wave in every caller).
//   IPPI_BC_PERIODIC_BB: periodic in x; halfway bounce-back below row 0 and
//     above row ny-1 (the population leaving the node comes back reversed).
//   IPPI_BC_INFLOW: rows 0 and ny-1 are specular free-slip rows
//     (mesh.cpp:325-337 and its mirror image).
// FP32 storage (s = f - c): a population returned reversed by bounce-back
// changes offset (bb = c_7 - c_5, see RefState); the mirror of the
// free-slip rows keeps it (c is symmetric in y).
template <typename T, int BC>
__device__ __forceinline__ void pull(const T* __restrict__ s, int x, int y, int nx, int ny, unsigned N, T* f,
                                     float bb = 0.0f) {
    const bool bot = y == 0, top = y == ny - 1;
    unsigned xc[3];  // lane columns for dx = -1, 0, +1
    if (BC == IPPI_BC_PERIODIC_BB) {
        xc[0] = x == 0 ? nx - 1 : x - 1;
        xc[1] = x;
        xc[2] = x == nx - 1 ? 0 : x + 1;
    } else {
        xc[0] = x - 1;
        xc[1] = x;
        xc[2] = x + 1;
    }
#pragma unroll
    for (int k = 0; k < 9; k++) {
        int pk = k, dy = -EY[k], dx = -EX[k];
        if (BC == IPPI_BC_PERIODIC_BB) {
            if (EY[k] == 1 && bot) {  // source below row 0: bounce-back
                pk = OPP[k];
                dy = 0;
                dx = 0;
            }
            if (EY[k] == -1 && top) {
                pk = OPP[k];
                dy = 0;
                dx = 0;
            }
        } else {
            if (EY[k] == 1 && bot) {  // specular: k comes from its mirror k' one row up
                pk = k == 2 ? 4 : (k == 5 ? 8 : 7);
                dy = 1;
            }
            if (EY[k] == -1 && bot) {
                dy = 1;
            }
            if (EY[k] == -1 && top) {
                pk = k == 4 ? 2 : (k == 7 ? 6 : 5);
                dy = -1;
            }
            if (EY[k] == 1 && top) {
                dy = -1;
            }
        }
        const unsigned row = (unsigned)(y + dy) * nx;
        f[k] = ld(s + (unsigned)pk * N, row + xc[dx + 1]);
    }
    if constexpr (sizeof(T) == 4 && BC == IPPI_BC_PERIODIC_BB) {
        if (bot) {
            f[5] += bb;
            f[6] -= bb;
        }
        if (top) {
            f[7] -= bb;
            f[8] += bb;
        }
    }
}

// Post-streaming state of node (x, y): populations (full f for FP64, f - w
// for FP32) and moments. For IPPI_BC_INFLOW the columns 0 and nx-1 are
// non-equilibrium extrapolation nodes (mesh.cpp:239-286): rho = 1, inlet
// u = u_in, outlet u = u(nx-2), f = feq(1, u) + f_n - feq_n with n the
// neighbouring column.
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

// FP32 node state: stored deviations g = f - c and their moments (see
// lbm_node.hpp). `Mo` keeps the partial sums shared by the moments, the
// stress and the MRT transform, the deviations drho, (dux, duy) of the
// density and velocity from the reference state (1, (U, 0)), and the full
// velocity ux = U + dux, uy = duy.
struct Mo {
    float drho, rho, ir, dux, duy, ux, uy;
    float s13, s24, d13, d24, s57, s68, sx, sy;
};

__device__ __forceinline__ void moments32(const float* g, float U, Mo& o) {
#pragma clang fp contract(fast)
    o.s13 = g[1] + g[3];
    o.s24 = g[2] + g[4];
    o.d13 = g[1] - g[3];
    o.d24 = g[2] - g[4];
    o.s57 = g[5] + g[7];
    o.s68 = g[6] + g[8];
    const float d57 = g[5] - g[7];
    o.sx = d57 + (g[8] - g[6]);  // g5 - g6 - g7 + g8
    o.sy = d57 + (g[6] - g[8]);  // g5 + g6 - g7 - g8
    o.drho = g[0] + ((o.s13 + o.s24) + (o.s57 + o.s68));
    const float djx = o.d13 + o.sx, djy = o.d24 + o.sy;
    o.rho = 1.0f + o.drho;
    if (o.rho < 1e-6f) {
        o.ir = 0.f;
        o.dux = -U;
        o.duy = 0.f;
    } else {
        o.ir = sh::rcp(o.rho);
        o.dux = (djx - U * o.drho) * o.ir;
        o.duy = djy * o.ir;
    }
    o.ux = U + o.dux;
    o.uy = o.duy;
}

// Non-equilibrium extrapolation in deviations: g = q(b) + (g - q(n)), the
// node taking rho_b = 1 + drb and u_b = (U + dbx, dby).
// qbc: the boundary node's equilibrium deviation when the caller has it
// (the inlet's, StepArgs::qin), else computed here; rho1: drb is 0 (the
// same values without the c drb terms, which the compiler would otherwise
// evaluate for every node ahead of the boundary branch)
__device__ __forceinline__ void nee32(float* g, Mo& o, float drb, float dbx, float dby, const RefState& r,
                                      const float* qbc = nullptr, bool rho1 = false) {
#pragma clang fp contract(fast)
    float qb[9], qn[9];
    if (qbc) {
#pragma unroll
        for (int k = 0; k < 9; k++) qb[k] = qbc[k];
    } else if (rho1) {
        sh::geq_dev0(dbx, dby, r.Uf, qb);
    } else {
        sh::geq_dev(drb, 1.0f + drb, dbx, dby, r.Uf, r.cf, qb);
    }
    sh::geq_dev(o.drho, o.rho, o.dux, o.duy, r.Uf, r.cf, qn);
#pragma unroll
    for (int k = 0; k < 9; k++) g[k] = qb[k] + (g[k] - qn[k]);
    moments32(g, r.Uf, o);  // partial sums of the new populations
    o.drho = drb;
    o.rho = 1.0f + drb;
    o.ir = sh::rcp(o.rho);
    o.dux = dbx;
    o.duy = dby;
    o.ux = r.Uf + dbx;
    o.uy = dby;
}

// Non-equilibrium second moments sum e e f - (rho u u + rho/3 I) from the
// deviations: the stored sums less (drho U^2 + rho X + drho/3, ...) with
// X = dux (2 U + dux), i.e. rho u_x^2 - U^2 (c's own moments cancel).
__device__ __forceinline__ void neq32(const Mo& o, float U, float& nxx, float& nxy, float& nyy) {
#pragma clang fp contract(fast)
    const float d3 = o.drho * (1.0f / 3.0f);
    const float X = o.dux * (2.0f * U + o.dux), Jy = o.rho * o.duy;
    nxx = (o.s13 + (o.s57 + o.s68)) - ((o.drho * (U * U) + o.rho * X) + d3);
    nyy = (o.s24 + (o.s57 + o.s68)) - (Jy * o.duy + d3);
    nxy = (o.s57 - o.s68) - Jy * (U + o.dux);
}

template <int BC>
__device__ __forceinline__ void node_state32(const float* __restrict__ s, int x, int y, int nx, int ny, unsigned N,
                                             float duinx, float uiny, const RefState& r, float* g, Mo& o) {
#pragma clang fp contract(fast)
    if (BC == IPPI_BC_PERIODIC_BB) {
        pull<float, IPPI_BC_PERIODIC_BB>(s, x, y, nx, ny, N, g, r.bb);
        moments32(g, r.Uf, o);
        return;
    }
    const bool nee = (x == 0) || (x == nx - 1);
    const int xs = x == 0 ? 1 : (x == nx - 1 ? nx - 2 : x);
    pull<float, IPPI_BC_INFLOW>(s, xs, y, nx, ny, N, g);
    moments32(g, r.Uf, o);
    if (nee) {
        if (x == 0)
            nee32(g, o, 0.f, duinx, uiny, r);
        else
            nee32(g, o, 0.f, o.dux, o.duy, r);
    }
}

// ------------------------------------------------------------- flat plate
// IPPI_BC_FLATPLATE, the reference's own boundary set (Mesh::BoundaryCondition,
// mesh.cpp:228-356, styles of Mesh::JudgeStyle, mesh.cpp:1391-1454), x being
// the reference's global column. State of node (x, y) after streaming and
// the boundary rules, with the macros its collision uses:
//   1 <= x <= nx-2, 1 <= y <= ny-2: pull, macros from the populations;
//   x = 0 (inlet), 1 <= y <= ny-2: rho = 1, u = u_in, non-equilibrium
//     extrapolation (NEE) from (1, y);
//   x = nx-1 (outlet), 1 <= y <= ny-2: rho = 1, u = u(nx-2, y), NEE from (nx-2, y);
//   y = ny-1 (top), all x: rho = 1, u = u(x, ny-2), NEE from (x, ny-2), where
//     (0, ny-2) and (nx-1, ny-2) are the inlet and outlet nodes just built;
//   y = 0, 1 <= x < plate_x: specular free slip ('m'), macros from the populations;
//   y = 0, plate_x <= x <= nx-2: rho = rho(x, 1), u = 0, NEE from (x, 1) ('d');
//   (0, 0) and (nx-1, 0): no rule reaches them, so their populations and
//     macros never change after initialisation. Returns false for them, with
//     their stored post-collision populations in f (copied through).
// FP64: the reference's arithmetic (NEE as feq_b + f_n - feq_n).
__device__ __forceinline__ bool fp_state64(const double* __restrict__ s, int x, int y, int nx, int ny, unsigned N,
                                           double uinx, double uiny, int plate_x, double* f, double& rho, double& ux,
                                           double& uy) {
#pragma clang fp contract(off)
    auto nee = [&](double rb, double bx, double by) {
        double qb[9], qn[9];
        ref::feq(rb, bx, by, qb);
        ref::feq(rho, ux, uy, qn);
#pragma unroll
        for (int k = 0; k < 9; k++) f[k] = qb[k] + f[k] - qn[k];
        rho = rb;
        ux = bx;
        uy = by;
    };
    if (y == 0) {
        if (x == 0 || x == nx - 1) {
#pragma unroll
            for (int k = 0; k < 9; k++) f[k] = ld(s + k * N, (unsigned)x);
            rho = 1.0;
            ux = 0.0;
            uy = 0.0;
            return false;
        }
        pull<double, IPPI_BC_INFLOW>(s, x, x < plate_x ? 0 : 1, nx, ny, N, f);
        ref::macro(f, rho, ux, uy);
        if (x >= plate_x) nee(rho, 0.0, 0.0);
        return true;
    }
    const int xs = x == 0 ? 1 : (x == nx - 1 ? nx - 2 : x);
    pull<double, IPPI_BC_INFLOW>(s, xs, y == ny - 1 ? ny - 2 : y, nx, ny, N, f);
    ref::macro(f, rho, ux, uy);
    if (x == 0)
        nee(1.0, uinx, uiny);
    else if (x == nx - 1)
        nee(1.0, ux, uy);
    if (y == ny - 1) nee(1.0, ux, uy);
    return true;
}

// FP32 (deviation) counterpart of fp_state64.
__device__ __forceinline__ bool fp_state32(const float* __restrict__ s, int x, int y, int nx, int ny, unsigned N,
                                           float duinx, float uiny, int plate_x, const RefState& r, float* g, Mo& o) {
#pragma clang fp contract(fast)
    if (y == 0) {
        if (x == 0 || x == nx - 1) {
#pragma unroll
            for (int k = 0; k < 9; k++) g[k] = ld(s + k * N, (unsigned)x);
            moments32(g, r.Uf, o);
            return false;
        }
        pull<float, IPPI_BC_INFLOW>(s, x, x < plate_x ? 0 : 1, nx, ny, N, g);
        moments32(g, r.Uf, o);
        if (x >= plate_x) nee32(g, o, o.drho, -r.Uf, 0.f, r);  // 'd': rho from (x, 1), u = 0
        return true;
    }
    const int xs = x == 0 ? 1 : (x == nx - 1 ? nx - 2 : x);
    pull<float, IPPI_BC_INFLOW>(s, xs, y == ny - 1 ? ny - 2 : y, nx, ny, N, g);
    moments32(g, r.Uf, o);
    if (x == 0)
        nee32(g, o, 0.f, duinx, uiny, r);
    else if (x == nx - 1)
        nee32(g, o, 0.f, o.dux, o.duy, r);
    if (y == ny - 1) nee32(g, o, 0.f, o.dux, o.duy, r);
    return true;
}

// IB force of node (x, y) from the per-body patches (the block-level test
// keeps the per-node test off all blocks that touch no patch).
// COH: the marker forces were stored by another block of the same kernel
// (k_step_f32_ib) and are read with agent-coherent loads.
template <bool COH = false>
__device__ __forceinline__ bool patch_force(const StepArgs& a, int x, int y, int bx0, int bx1, int by0, int by1,
                                            double& fx, double& fy) {
    // the bodies whose patches cover the block's row (one scalar load), in increasing order
    unsigned m = a.nbody >= 32 ? ~0u : (1u << a.nbody) - 1u;
    if (a.rowmask && by0 == by1) m &= __builtin_amdgcn_readfirstlane(a.rowmask[by0]);
    while (m) {
        const int b = __builtin_ctz(m);
        m &= m - 1u;
        const Patch p = a.patch[b];
        if (bx1 < p.x0 || bx0 >= p.x0 + p.w || by1 < p.y0 || by0 >= p.y0 + p.h) continue;
        if (x >= p.x0 && x < p.x0 + p.w && y >= p.y0 && y < p.y0 + p.h) {
            const size_t i = (size_t)p.off + (size_t)(y - p.y0) * p.w + (x - p.x0);
            if (!a.ib_gather) {
                fx = a.patch_f[2 * i];
                fy = a.patch_f[2 * i + 1];
                return true;
            }
            // SpreadForce as a gather: sum of F_m phi phi over the markers
            // reaching the node, in increasing marker order (zero if none),
            // the order and arithmetic of the IB kernel's own spread
            fx = 0.0;
            fy = 0.0;
            const int ia = a.pslot[i];
            if (ia >= 0) {
                const size_t E = (size_t)a.ell;
                const int cnt = a.ecnt[ia];
                for (int e0 = 0; e0 < cnt; e0 += 2) {
                    int mk[2];
                    double w[2], Fx[2], Fy[2];
#pragma unroll
                    for (int j = 0; j < 2; j++) {  // padding entries are valid (marker 0, weight 0)
                        mk[j] = p.m0 + a.emk[E * ia + min(e0 + j, (int)E - 1)];
                        w[j] = a.ecw[E * ia + min(e0 + j, (int)E - 1)];
                    }
#pragma unroll
                    for (int j = 0; j < 2; j++) {
                        if (COH) {
                            double* Fp = const_cast<double*>(a.F) + 2 * mk[j];
                            Fx[j] = __hip_atomic_load(Fp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            Fy[j] = __hip_atomic_load(Fp + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        } else {
                            Fx[j] = a.F[2 * mk[j]];
                            Fy[j] = a.F[2 * mk[j] + 1];
                        }
                    }
#pragma unroll
                    for (int j = 0; j < 2; j++) {
                        if (e0 + j < cnt) {
                            fx += (Fx[j] * w[j]);
                            fy += (Fy[j] * w[j]);
                        }
                    }
                }
            }
            return true;
        }
    }
    return false;
}

constexpr int BX = 256;  // nodes per block along x (one row segment)

// XCD-grouped block order. Consecutive workgroups go to the eight XCDs in
// turn, so workgroup L runs on XCD L % 8. Giving the i-th workgroup of an
// XCD column i % gx of row 8 (i / gx) + L % 8 puts all blocks of a row on
// one XCD: the cache lines that neighbouring blocks' x-shifted loads share
// are then fetched into one L2, not two. With G rows of blocks, the rows
// past the last whole group of 8 follow in plain order. Used while the
// stores are temporal (both population buffers
// within 512 MB). Measured per step, grouped against plain order: the
// 2048 x 1024 cylinder 25.9 against 26.4 us (FP32) and 63.3 against 65.5
// us (FP64), 1024 x 512 FP32 9.8 against 10.3 us; with non-temporal stores
// (4096 x 2048) grouping costs 6% (FP32 110.3 against 103.7 us) and 3%
// with the network, so those keep the plain order.
__device__ __forceinline__ void xcd_block(int L, int gx, int G, int& bx, int& by) {
    const int full = gx * (G & ~7);  // blocks in whole groups of 8 rows
    if (L >= full) {
        bx = (L - full) % gx;
        by = (G & ~7) + (L - full) / gx;
        return;
    }
    const int k = L >> 3;
    bx = k % gx;
    by = 8 * (k / gx) + (L & 7);
}


// Slots of node (x, y) for the next IB kernel: sl[k] is the slot of the
// active node that pulls population k from here (-1: none). Loaded at the
// start of the node's work, so the loads overlap the population loads. A
// node lies in at most one body's region for each k that has a slot.
__device__ __forceinline__ void ib_slot_ids(const StepArgs& a, int x, int y, int bx0, int* sl) {
#pragma unroll
    for (int k = 0; k < 9; k++) sl[k] = -1;
    if (!a.scat) return;
    const int bx1 = bx0 + BX - 1;
    for (int b = 0; b < a.nbody; ++b) {
        const Patch p = a.patch[b];
        if (y < p.y0 - 1 || y > p.y0 + p.h || bx1 < p.x0 - 1 || bx0 > p.x0 + p.w) continue;
        if (x >= p.x0 - 1 && x <= p.x0 + p.w) {
            const int ri = p.roff + (y - p.y0 + 1) * (p.w + 2) + (x - p.x0 + 1);
#pragma unroll
            for (int k = 0; k < 9; k++) {
                const int v = a.scat[k * a.nreg + ri];
                sl[k] = v >= 0 ? v : sl[k];
            }
        }
    }
}
// Population k of the stored state v to slot sl[k].
template <typename T>
__device__ __forceinline__ void ib_slots(const StepArgs& a, const int* sl, const T* v) {
    T* slot = static_cast<T*>(a.slot);
#pragma unroll
    for (int k = 0; k < 9; k++)
        if (sl[k] >= 0) slot[(size_t)k * a.nact + sl[k]] = v[k];
}

// Rows per thread of the fused pull-step kernels: each thread walks NPT rows
// of one column and issues the loads of row j+1 before the arithmetic of
// row j, so every wave keeps two rows of populations in flight. Measured on
// the headline lattice: FP32 is fastest with one row (two or four cost
// 3-15%), FP64 without the network with four (5% faster than one). The FP64
// value is the most rows a kernel is built for; rows64() picks per lattice.
template <typename T, bool PULL, bool NN>
constexpr int step_npt() {
    return PULL && sizeof(T) == 8 && !NN ? 4 : 1;
}
// Four rows per thread leave small lattices with too few blocks for the
// GPU: 1024 x 512 (512 blocks of four rows) takes 27.3 us per FP64 step with
// the IB inside against 19.5 us with one row, 256 x 128 18.5 against 10.4;
// 2048 x 1024 (2048 blocks) 66.9 against 69.0. Four rows from 2^20 nodes on.
constexpr size_t F64_ROWS4_MIN_NODES = (size_t)1 << 20;
inline int rows64_for(int nx, int ny) { return (size_t)nx * ny >= F64_ROWS4_MIN_NODES ? 4 : 1; }
// The same for boundary set BC. The flat-plate rules load their own
// populations, so no row's loads go out ahead of the previous row's work
// there and one row per thread is fastest in both precisions (2401 x 501
// FP64: 30.6 us against 31.9 with two rows and 33.2 with four; FP32: 17.5
// against 26.6 and 24.4).
template <typename T, bool PULL, bool NN, int BC>
constexpr int step_npt_bc() {
    return BC == IPPI_BC_FLATPLATE ? 1 : step_npt<T, PULL, NN>();
}

// ------------------------------------------------------------- FP64 step
// Raw pulled populations of node (x, y) (for IPPI_BC_INFLOW the NEE columns
// pull the neighbouring column), or the stored fnew state when !PULL.
template <typename T, int BC, bool PULL>
__device__ __forceinline__ void load_raw(const T* __restrict__ src, int x, int y, int nx, int ny, unsigned N, T* f,
                                         float bb = 0.0f) {
    if (PULL) {
        const int xs = BC == IPPI_BC_INFLOW ? (x == 0 ? 1 : (x == nx - 1 ? nx - 2 : x)) : x;
        pull<T, BC>(src, xs, y, nx, ny, N, f, bb);
    } else {
        const unsigned n = (unsigned)y * nx + x;
#pragma unroll
        for (int k = 0; k < 9; k++) f[k] = ld(src + k * N, n);
    }
}

// Collision of one node. The equilibrium is computed once and shared by the
// LES and the BGK / RR collisions (the reference recomputes the identical
// values); the MRT keeps the reference's moment-space evaluation.
// NN tau of the FP64 path: features as the oracle's or_mlp_features
// (rho, ux, uy, |S| with S = -3/(2 rho) Pi_neq, Pi_xx - Pi_yy, Pi_xy).
__device__ __forceinline__ double nn_tau_f64(const double* f, const double* q, double rho, double ux, double uy,
                                             const double* __restrict__ blob) {
#pragma clang fp contract(off)
    double d[9];
#pragma unroll
    for (int k = 0; k < 9; k++) d[k] = f[k] - q[k];
    const double pxx = ((((d[1] + d[3]) + d[5]) + d[6]) + d[7]) + d[8];
    const double pxy = ((d[5] - d[6]) + d[7]) - d[8];
    const double pyy = ((((d[2] + d[4]) + d[5]) + d[6]) + d[7]) + d[8];
    const double c = -3.0 / (2 * rho);
    const double sxx = pxx * c, sxy = pxy * c, syy = pyy * c;
    double in[6];
    in[0] = rho;
    in[1] = ux;
    in[2] = uy;
    in[3] = sqrt(2.0 * (sxx * sxx + 2.0 * sxy * sxy + syy * syy));
    in[4] = pxx - pyy;
    in[5] = pxy;
    return mlp_tau_f64(in, blob);
}

template <int COLL, int BC, bool PULL, bool NN, bool NT, bool COH = false>
__device__ __forceinline__ void node_f64(double* f, int x, int y, unsigned N, const double* __restrict__ src,
                                         const double* __restrict__ mac, double* __restrict__ dst, const StepArgs& a,
                                         bool active, int bx0) {
#pragma clang fp contract(off)
    const int nx = a.nx;
    const unsigned n = (unsigned)y * nx + x;
    double rho, ux, uy;
    bool fresh = true;  // false: flat-plate corner, stored populations copied through
    if (PULL && BC == IPPI_BC_FLATPLATE) {
        fresh = fp_state64(src, x, y, nx, a.ny, N, a.uinx, a.uiny, a.plate_x, f, rho, ux, uy);
    } else if (PULL) {
        ref::macro(f, rho, ux, uy);
        if (BC == IPPI_BC_INFLOW && (x == 0 || x == nx - 1)) {  // NEE, mesh.cpp:239-286
            double bx = x == 0 ? a.uinx : ux, by = x == 0 ? a.uiny : uy;
            double qb[9], qn[9];
            ref::feq(1.0, bx, by, qb);
            ref::feq(rho, ux, uy, qn);
#pragma unroll
            for (int k = 0; k < 9; k++) f[k] = qb[k] + f[k] - qn[k];
            rho = 1.0;
            ux = bx;
            uy = by;
        }
    } else {
        rho = mac[n];
        ux = mac[N + n];
        uy = mac[2 * N + n];
    }
    if (!(rho >= 0.0 && rho <= 10.0)) *a.flag = 1;  // ErrorCatch, mesh.cpp:2101-2111
    double fx = 0.0, fy = 0.0;
    bool forced = false;
    if (a.nbody > 0) {
        forced = patch_force<COH>(a, x, y, bx0, bx0 + BX - 1, y, y, fx, fy);
        if (forced) {  // UpdateVelocity, mesh.cpp:1254-1255
            ux = ux + 0.5 * 1 * fx / rho;
            uy = uy + 0.5 * 1 * fy / rho;
        }
    }
    const double tb = a.tau_field ? ld(static_cast<const double*>(a.tau_ptr), n) : a.tau0;
    double tau = tb;
    double o[9];
    double q[9];
    ref::feq(rho, ux, uy, q);
    bool nn_ok = false;
    if (NN) {  // guard of mesh.cpp:2821-2827: NN value if >= 0.501, else LES / base
        const double t = nn_tau_f64(f, q, rho, ux, uy, a.mlp.blob);
        nn_ok = t >= 0.501 && t <= a.nn_max;
        if (nn_ok) tau = t;
    }
    if (a.bnn_on) {  // a boundary network's accepted value (NNflag, mesh.cpp:393)
        const double tbn = bnn_tau_at(a, x, y);
        if (tbn > 0.0) {
            tau = tbn;
            nn_ok = true;
        }
    }
    if (COLL == IPPI_COLL_MRT) {
        if (a.les && !nn_ok) tau = ref::les_tau_q(f, q, rho, tb, a.cs);
        ref::mrt(f, rho, ux, uy, tau, o);
    } else {
        if (a.les && !nn_ok) tau = ref::les_tau_q(f, q, rho, tb, a.cs);
        if (COLL == IPPI_COLL_RR)
            ref::rr_q(f, q, tau, o);
        else
            ref::bgk_q(f, q, tau, o);
    }
    if (forced) {
        double g[9];
        ref::guo(ux, uy, fx, fy, tb, g);
#pragma unroll
        for (int k = 0; k < 9; k++) o[k] = o[k] + g[k];
    }
    if (active) {
        double v[9];
#pragma unroll
        for (int k = 0; k < 9; k++) v[k] = fresh ? o[k] : f[k];
        if (a.round32) {  // block-uniform
#pragma unroll
            for (int k = 0; k < 9; k++) v[k] = (double)(float)(v[k] - a.ref.c[k]) + a.ref.c[k];
        }
        st9<NT>(dst, N, n, v);
    }
}

// PULL = false: collide the fnew state stored at the node itself with the
// moments from the macro arrays (first collision after initialisation).
template <int COLL, int BC, bool PULL, int NPT, bool NN, bool NT>
__global__ void __launch_bounds__(BX) k_step_f64(const double* __restrict__ src, double* __restrict__ dst,
                                                 const double* __restrict__ mac, StepArgs a) {
    const int nx = a.nx, ny = a.ny;
    int bxi = blockIdx.x, byi = blockIdx.y;
    if (a.xcd) xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x, gridDim.y, bxi, byi);
    const bool active = (int)(bxi * BX + threadIdx.x) < nx;
    const int x = min((int)(bxi * BX + threadIdx.x), nx - 1);
    const int y0 = byi * NPT;
    const unsigned N = (unsigned)nx * ny;
    constexpr bool RAW = !(PULL && BC == IPPI_BC_FLATPLATE);  // the flat-plate rules load for themselves
    double cur[9];
    if (RAW) load_raw<double, BC, PULL>(src, x, y0, nx, ny, N, cur);
#pragma unroll
    for (int j = 0; j < NPT; j++) {
        const int y = y0 + j;
        if (y >= ny) break;
        const bool more = j + 1 < NPT && y + 1 < ny;
        double nxt[9];
        if (RAW && j + 1 < NPT) load_raw<double, BC, PULL>(src, x, min(y + 1, ny - 1), nx, ny, N, nxt);
        node_f64<COLL, BC, PULL, NN, NT>(cur, x, y, N, src, mac, dst, a, active, bxi * BX);
        if (more) {
#pragma unroll
            for (int k = 0; k < 9; k++) cur[k] = nxt[k];
        }
    }
}

// ------------------------------------------------------------- FP32 step
// Shifted populations g = f - w. BGK and MRT are evaluated in moment space
// (BGK = MRT with every non-conserved rate 1/tau): the conserved moments are
// never touched, so FP32 rounding cannot drift mass or momentum. RR is
// evaluated on populations with the increment projected onto exact mass and
// momentum. The LES stress comes from the second moments directly.
// NN: 0 no network, 1 network with compensated (hi/lo) bf16 products,
// 2 network with plain bf16 products (IPPI_TAU_NN_BF16)
template <int COLL, int BC, bool PULL, int NN, bool NT, bool COH = false>
__device__ __forceinline__ void node_f32(float* g, int x, int y, unsigned N, const float* __restrict__ src,
                                         const double* __restrict__ mac, float* __restrict__ dst, const StepArgs& a,
                                         bool active, __bf16* lds, int bx0, const u32x4* lf = nullptr) {
#pragma clang fp contract(fast)
    const int nx = a.nx;
    const unsigned n = (unsigned)y * nx + x;
    int sl[9];  // not in the network variants: VALU-bound, they need every register
    if (!NN) ib_slot_ids(a, x, y, bx0, sl);
    Mo o;
    bool fresh = true;  // false: flat-plate corner, stored populations copied through
    float keep[BC == IPPI_BC_FLATPLATE ? 9 : 1];
    const RefState& r = a.ref;
    const float U = r.Uf;
    if (PULL && BC == IPPI_BC_FLATPLATE) {
        fresh = fp_state32(src, x, y, nx, a.ny, N, a.duinxf, a.uinyf, a.plate_x, r, g, o);
#pragma unroll
        for (int k = 0; k < (BC == IPPI_BC_FLATPLATE ? 9 : 1); k++) keep[k] = g[k];
    } else {
        moments32(g, U, o);
    }
    if (PULL && BC != IPPI_BC_FLATPLATE) {
        // NEE, mesh.cpp:239-286. The outer test is block-uniform (only the
        // first and last blocks of a row hold x = 0 or x = nx - 1), so the
        // other blocks skip the inlet equilibrium altogether.
        if (BC == IPPI_BC_INFLOW && (bx0 == 0 || bx0 + BX >= nx)) {
            if (x == 0)
                nee32(g, o, 0.f, a.duinxf, a.uinyf, r, a.qin);
            else if (x == nx - 1)
                nee32(g, o, 0.f, o.dux, o.duy, r, nullptr, true);
        }
    } else if (!PULL) {
        o.rho = (float)mac[n];
        o.drho = (float)(mac[n] - 1.0);
        o.ir = sh::rcp(o.rho);
        o.dux = (float)(mac[N + n] - r.U);
        o.duy = (float)mac[2 * N + n];
    }
    if (!(o.rho >= 0.f && o.rho <= 10.f)) *a.flag = 1;
    float fx = 0.f, fy = 0.f;
    bool forced = false;
    if (a.nbody > 0) {
        double fxd = 0.0, fyd = 0.0;
        forced = patch_force<COH>(a, x, y, bx0, bx0 + BX - 1, y, y, fxd, fyd);
        if (forced) {
            fx = (float)fxd;
            fy = (float)fyd;
            const float h = 0.5f * o.ir;
            o.dux += fx * h;
            o.duy += fy * h;
        }
    }
    // relaxation as tx = tau - 1/2 (see lbm_node.hpp)
    float txb = a.tx0f;
    if (a.tau_field) txb = ld(static_cast<const float*>(a.tau_ptr), n);  // uniform branch, no load otherwise
    float tx = txb;
    const float rho = o.rho, drho = o.drho, dux = o.dux, duy = o.duy;
    const float ux = U + dux, uy = duy;
    // deviations from the reference state's: momentum J = rho u - (U, 0),
    // rho |u|^2 - U^2 = drho U^2 + rho Q, rho u_x^2 - U^2 = drho U^2 + rho X
    const float Jx = drho * U + rho * dux, Jy = rho * duy;
    const float X = dux * (2.0f * U + dux), Q = X + duy * duy;
    const float dU2 = drho * (U * U);
    float nxx, nxy, nyy;  // non-equilibrium second moments
    neq32(o, U, nxx, nxy, nyy);
    bool nn_ok = false;
    if (NN) {  // guard of mesh.cpp:2821-2827: NN value if >= 0.501, else LES / base
        // features as deviations (rho - 1, u - (U, 0)); the host shifts mu to match
        float feat[6];
        feat[0] = drho;
        feat[1] = dux;
        feat[2] = duy;
        feat[3] = 1.5f * o.ir * sh::fsqrt(2.0f * (nxx * nxx + 2.0f * nxy * nxy + nyy * nyy));
        feat[4] = nxx - nyy;
        feat[5] = nxy;
        const float t = mlp_tau_nn<NN>(feat, a.mlp, lds, lf);  // tau - 1/2
        nn_ok = t >= 0.001f && t <= a.nn_maxtxf;
        if (nn_ok) tx = t;
    }
    if (a.bnn_on) {  // a boundary network's accepted value (NNflag, mesh.cpp:393)
        const double tbn = bnn_tau_at(a, x, y);
        if (tbn > 0.0) {
            tx = (float)(tbn - 0.5);
            nn_ok = true;
        }
    }
    if (a.les && !nn_ok) tx = sh::les_tx(nxx, nxy, nyy, rho, txb, a.csf);
    const float dl = sh::dl_of(tx);  // 2 - omega
    if (COLL == IPPI_COLL_RR) {
        float q[9];
        sh::geq_dev(drho, rho, dux, duy, U, r.cf, q);
        sh::rr(g, q, nxx, nxy, nyy, dl, fx, fy);
    } else {
        // moments of g (Lallemand-Luo basis, NodeAndLattice.cpp:18-27)
        const float s4 = o.s13 + o.s24, s8 = o.s57 + o.s68;
        const float m1 = -4.0f * g[0] - s4 + 2.0f * s8;
        const float m2 = 4.0f * g[0] - 2.0f * s4 + s8;
        const float m4 = -2.0f * o.d13 + o.sx;
        const float m6 = -2.0f * o.d24 + o.sy;
        const float m7 = o.s13 - o.s24;
        const float m8 = o.s57 - o.s68;
        // equilibrium moments less those of c: energy drho U^2 + rho Q
        // terms, p_xx rho u_x^2 - rho u_y^2 - U^2, p_xy rho u_x u_y
        const float e1 = -2.0f * drho + 3.0f * (dU2 + rho * Q);
        const float e2 = drho - 3.0f * (dU2 + rho * Q);
        // c_l = s_l (m_l - meq_l) / |M_l|^2; a rate s_l = omega = 2 - dl is
        // applied as 2 x - dl x (one rounding, no rounded omega)
        auto om_x = [dl](float x) { return 2.0f * x - dl * x; };
        const float x1 = (m1 - e1) * (1.0f / 36.0f), x2 = (m2 - e2) * (1.0f / 36.0f);
        const float x4 = (m4 + Jx) * (1.0f / 12.0f), x6 = (m6 + Jy) * (1.0f / 12.0f);
        float c1, c2, c4, c6;
        if (COLL == IPPI_COLL_MRT) {  // fixed rates, NodeAndLattice.cpp:306
            c1 = 1.2f * x1;
            c2 = 1.1f * x2;
            c4 = 1.2f * x4;
            c6 = 1.2f * x6;
        } else {
            c1 = om_x(x1);
            c2 = om_x(x2);
            c4 = om_x(x4);
            c6 = om_x(x6);
        }
        const float c7 = om_x((m7 - ((dU2 + rho * X) - Jy * duy)) * 0.25f);
        const float c8 = om_x((m8 - Jy * ux) * 0.25f);
        // BGK relaxes the momentum towards rho u (shifted by F/2 under IB
        // forcing): exact target -omega F/2; MRT does not relax it.
        float c3 = 0.f, c5 = 0.f;
        if (COLL == IPPI_COLL_BGK) {
            c3 = om_x(fx * (-1.0f / 12.0f));
            c5 = om_x(fy * (-1.0f / 12.0f));
        }
        const float A = -c1 - 2.0f * c2, B = 2.0f * c1 + c2;
        const float p46 = c4 + c6, q64 = c6 - c4;
        g[0] -= 4.0f * (c2 - c1);
        g[1] -= A - 2.0f * c4 + c7 + c3;
        g[3] -= A + 2.0f * c4 + c7 - c3;
        g[2] -= A - 2.0f * c6 - c7 + c5;
        g[4] -= A + 2.0f * c6 - c7 - c5;
        g[5] -= B + p46 + c8 + c3 + c5;
        g[7] -= B - p46 + c8 - c3 - c5;
        g[6] -= B + q64 - c8 - c3 + c5;
        g[8] -= B - q64 - c8 + c3 - c5;
    }
    if (forced) sh::guo(ux, uy, fx, fy, sh::dl_of(txb), g);
    if (BC == IPPI_BC_FLATPLATE && !fresh) {
#pragma unroll
        for (int k = 0; k < 9; k++) g[k] = keep[BC == IPPI_BC_FLATPLATE ? k : 0];
    }
    if (active) {
        st9<NT>(dst, N, n, g);
        if (!NN && a.scat) ib_slots<float>(a, sl, g);
    }
}

// At least 4 waves per SIMD for the network variants: they are VALU-bound
// and need the occupancy to overlap transcendental latency and the
// population loads (caps them at 128 VGPRs). The others are latency-bound
// streams and keep the full 8 waves (64 VGPRs).
template <int COLL, int BC, bool PULL, int NPT, int NN, bool NT>
__global__ void __launch_bounds__(BX) __attribute__((amdgpu_waves_per_eu(NN ? IPPI_NN_WAVES : 8)))
k_step_f32(const float* __restrict__ src, float* __restrict__ dst,
                                                 const double* __restrict__ mac, StepArgs a) {
    const int nx = a.nx, ny = a.ny;
    // all lanes stay resident (the MLP's MFMAs are wave-wide); lanes past
    // the last column recompute column nx-1 and do not store
    int bxi = blockIdx.x, byi = blockIdx.y;
    if (a.xcd) xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x, gridDim.y, bxi, byi);
    const bool active = (int)(bxi * BX + threadIdx.x) < nx;
    const int x = min((int)(bxi * BX + threadIdx.x), nx - 1);
    const int y0 = byi * NPT;
    __shared__ __bf16 lds_mlp[NN ? BX * 16 : 1];
    __shared__ u32x4 lds_frag[NN == 3 ? MLP_FRAGE_CHUNKS * 64 : 1];
    __bf16* lds = lds_mlp + (NN ? (threadIdx.x & ~63) * 16 : 0);
    const unsigned N = (unsigned)nx * ny;
    constexpr bool RAW = !(PULL && BC == IPPI_BC_FLATPLATE);  // the flat-plate rules load for themselves
    if (NN == 3) mlp_frage_to_lds<BX>(a.mlp, lds_frag);  // in flight with the population loads
    float cur[9];
    if (RAW) load_raw<float, BC, PULL>(src, x, y0, nx, ny, N, cur, a.ref.bb);
    if (NN == 3) __syncthreads();
#pragma unroll
    for (int j = 0; j < NPT; j++) {
        const int y = y0 + j;
        if (y >= ny) break;
        const bool more = j + 1 < NPT && y + 1 < ny;
        float nxt[9];
        if (RAW && j + 1 < NPT) load_raw<float, BC, PULL>(src, x, min(y + 1, ny - 1), nx, ny, N, nxt, a.ref.bb);
        node_f32<COLL, BC, PULL, NN, NT>(cur, x, y, N, src, mac, dst, a, active, lds, bxi * BX, lds_frag);
        if (more) {
#pragma unroll
            for (int k = 0; k < 9; k++) cur[k] = nxt[k];
        }
    }
}


// Inputs of the last pull step's collision at the nine nodes population k
// of node (x, y) streams in from, s_k = (x, y) - e_k, in FP64: fnew_k(s_k)
// (pulled from the state that step read, prev), feq_k(s_k) from the macros
// of s_k's pulled populations, and s_k's relaxation time (base, or the
// hybrid Smagorinsky value; the IB velocity correction is not applied). One
// thread per (node, k); the nodes lie two or more rows and columns inside
// the domain, so every pull is the plain stencil.
template <typename T>
__global__ void k_stencil(const T* __restrict__ prev, const int* __restrict__ xy, int n, StepArgs a,
                          double* __restrict__ fnew, double* __restrict__ feq, double* __restrict__ tau) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 9 * n) return;
    const int node = i / 9, k = i - 9 * node;
    const int nx = a.nx, sx = xy[2 * node] - EX[k], sy = xy[2 * node + 1] - EY[k];
    const unsigned N = (unsigned)nx * a.ny, sn = (unsigned)sy * nx + sx;
    double f[9];
#pragma unroll
    for (int q = 0; q < 9; q++) {
        const T v = ld(prev + q * N, (unsigned)(sy - EY[q]) * nx + (sx - EX[q]));
        f[q] = sizeof(T) == 8 ? (double)v : (double)v + a.ref.c[q];
    }
    double rho, ux, uy, fq[9];
    ref::macro(f, rho, ux, uy);
    ref::feq(rho, ux, uy, fq);
    double tb = a.tau0;
    if (a.tau_field)
        tb = sizeof(T) == 8 ? (double)ld(static_cast<const double*>(a.tau_ptr), sn)
                            : 0.5 + (double)ld(static_cast<const float*>(a.tau_ptr), sn);
    fnew[i] = f[k];
    feq[i] = fq[k];
    tau[i] = a.les ? ref::les_tau_q(f, fq, rho, tb, a.cs) : tb;
}

// Bandwidth probe: the fused step's memory pattern with no arithmetic.
// MODE 1 loads the nine populations exactly as the pull of an inflow step
// does (shifted rows and columns, SGPR base + 32-bit offsets) and stores
// them with the step's store rule; MODE 0 copies them in place. The time
// per launch is the ceiling the fused kernel's memory traffic can reach.
// MODE 2 is MODE 1 with the XCD-grouped block order.
template <typename T, int MODE, bool NT>
__global__ void __launch_bounds__(BX) k_probe(const T* __restrict__ src, T* __restrict__ dst, int nx, int ny) {
    int bx = blockIdx.x, y = blockIdx.y;
    if (MODE == 2) xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x, gridDim.y, bx, y);
    const bool active = (int)(bx * BX + threadIdx.x) < nx;
    const int x = min((int)(bx * BX + threadIdx.x), nx - 1);
    const unsigned N = (unsigned)nx * ny, n = (unsigned)y * nx + x;
    T f[9];
    if (MODE >= 1) {
        load_raw<T, IPPI_BC_INFLOW, true>(src, x, y, nx, ny, N, f);
    } else {
#pragma unroll
        for (int k = 0; k < 9; k++) f[k] = ld(src + k * N, n);
    }
    if (active) {
#pragma unroll
        for (int k = 0; k < 9; k++) st<NT>(dst + k * N, n, f[k]);
    }
}

// Bandwidth ceiling: a flat copy of one array with 16 bytes per lane (grid
// stride), the simplest HBM stream the chip runs (pattern probe mode 3).
typedef float f32x4v __attribute__((ext_vector_type(4)));
template <bool NT>
__global__ void __launch_bounds__(256) k_copy16(const f32x4v* __restrict__ s, f32x4v* __restrict__ d, size_t n4) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n4; i += (size_t)gridDim.x * 256) {
        const f32x4v v = s[i];
        if (NT)
            __builtin_nontemporal_store(v, d + i);
        else
            d[i] = v;
    }
}

// Relaxation time the next collision uses at each node, from the current
// post-streaming state (IB velocity correction not applied): base value,
// LES, or the network with its guard.
template <int BC, int NN>
__global__ void __launch_bounds__(BX) k_tau32(const float* __restrict__ src, double* __restrict__ out, StepArgs a) {
#pragma clang fp contract(fast)
    const int nx = a.nx, ny = a.ny;
    const bool active = (int)(blockIdx.x * BX + threadIdx.x) < nx;
    const int x = min((int)(blockIdx.x * BX + threadIdx.x), nx - 1);
    const int y = blockIdx.y;
    const unsigned N = (unsigned)nx * ny, n = (unsigned)y * nx + x;
    __shared__ __bf16 lds_mlp[NN ? BX * 16 : 1];
    __shared__ u32x4 lds_frag[NN == 3 ? MLP_FRAGE_CHUNKS * 64 : 1];
    if (NN == 3) {
        mlp_frage_to_lds<BX>(a.mlp, lds_frag);
        __syncthreads();
    }
    float g[9];
    Mo o;
    if constexpr (BC == IPPI_BC_FLATPLATE) {
        if (y == 0 && (x == 0 || x == nx - 1)) {  // corner: initial populations, macros from mac
#pragma unroll
            for (int k = 0; k < 9; k++) g[k] = (float)(a.corner[9 * (x != 0) + k] - a.ref.c[k]);
            moments32(g, a.ref.Uf, o);
            o.rho = (float)a.mac[n];
            o.drho = (float)(a.mac[n] - 1.0);
            o.ir = sh::rcp(o.rho);
            o.dux = (float)(a.mac[N + n] - a.ref.U);
            o.duy = (float)a.mac[2 * N + n];
        } else {
            fp_state32(src, x, y, nx, ny, N, a.duinxf, a.uinyf, a.plate_x, a.ref, g, o);
        }
    } else {
        node_state32<BC>(src, x, y, nx, ny, N, a.duinxf, a.uinyf, a.ref, g, o);
    }
    const float txb = a.tau_field ? ld(static_cast<const float*>(a.tau_ptr), n) : a.tx0f;
    float nxx, nxy, nyy;
    neq32(o, a.ref.Uf, nxx, nxy, nyy);
    float tx = txb;
    bool nn_ok = false;
    if (NN) {  // features as deviations, as in node_f32
        float feat[6] = {o.drho, o.dux, o.duy,
                         1.5f * o.ir * sh::fsqrt(2.0f * (nxx * nxx + 2.0f * nxy * nxy + nyy * nyy)), nxx - nyy, nxy};
        const float t = mlp_tau_nn<NN>(feat, a.mlp, lds_mlp + (threadIdx.x & ~63) * 16, lds_frag);
        nn_ok = t >= 0.001f && t <= a.nn_maxtxf;
        if (nn_ok) tx = t;
    }
    if (a.bnn_on) {
        const double tbn = bnn_tau_at(a, x, y);
        if (tbn > 0.0) {
            tx = (float)(tbn - 0.5);
            nn_ok = true;
        }
    }
    if (a.les && !nn_ok) tx = sh::les_tx(nxx, nxy, nyy, o.rho, txb, a.csf);
    if (active) out[n] = 0.5 + (double)tx;
}

template <int BC, bool NN>
__global__ void __launch_bounds__(BX) k_tau64(const double* __restrict__ src, double* __restrict__ out, StepArgs a) {
#pragma clang fp contract(off)
    const int nx = a.nx, ny = a.ny;
    const int x = blockIdx.x * BX + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= nx) return;
    const unsigned N = (unsigned)nx * ny, n = (unsigned)y * nx + x;
    double f[9], rho, ux, uy;
    if constexpr (BC == IPPI_BC_FLATPLATE) {
        if (y == 0 && (x == 0 || x == nx - 1)) {  // corner: initial populations, macros from mac
#pragma unroll
            for (int k = 0; k < 9; k++) f[k] = a.corner[9 * (x != 0) + k];
            rho = a.mac[n];
            ux = a.mac[N + n];
            uy = a.mac[2 * N + n];
        } else {
            fp_state64(src, x, y, nx, ny, N, a.uinx, a.uiny, a.plate_x, f, rho, ux, uy);
        }
    } else {
        node_state(src, x, y, nx, ny, N, BC, a.uinx, a.uiny, f, rho, ux, uy);
    }
    const double tb = a.tau_field ? ld(static_cast<const double*>(a.tau_ptr), n) : a.tau0;
    double q[9];
    ref::feq(rho, ux, uy, q);
    double tau = tb;
    bool nn_ok = false;
    if (NN) {
        const double t = nn_tau_f64(f, q, rho, ux, uy, a.mlp.blob);
        nn_ok = t >= 0.501 && t <= a.nn_max;
        if (nn_ok) tau = t;
    }
    if (a.bnn_on) {
        const double tbn = bnn_tau_at(a, x, y);
        if (tbn > 0.0) {
            tau = tbn;
            nn_ok = true;
        }
    }
    if (a.les && !nn_ok) tau = ref::les_tau_q(f, q, rho, tb, a.cs);
    out[n] = tau;
}


// --------------------------------------------------- training-pair kernels
// Restriction of this (fine) solver's post-streaming state to the grid of
// half the resolution: coarse node (I, J) takes the mass and momentum of
// fine nodes (2I..2I+1, 2J..2J+1) and the average of their non-equilibrium
// parts f - feq(rho_f, u_f), scaled by s = 2 tau_c / tau_f (the stress in
// lattice units of a grid twice as coarse at the same lattice velocity):
// f_c = feq(rho_c, u_c) + s <f - feq>. The unresolved stress
// <feq> - feq(<.>) is left out, as the coarse solver's own state has none.
template <typename T, int BC>
__global__ void __launch_bounds__(BX) k_restrict(const T* __restrict__ src, double* __restrict__ fc,
                                                 double* __restrict__ mc, int nx, int ny, double uinx,
                                                 double uiny, double s, int want_f, RefState ref) {
#pragma clang fp contract(off)
    const int nxc = nx / 2, nyc = ny / 2;
    const int I = blockIdx.x * BX + threadIdx.x, J = blockIdx.y;
    if (I >= nxc) return;
    const unsigned N = (unsigned)nx * ny, Nc = (unsigned)nxc * nyc, nc = (unsigned)J * nxc + I;
    double fa[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, na[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int dj = 0; dj < 2; dj++)
        for (int di = 0; di < 2; di++) {
            const int x = 2 * I + di, y = 2 * J + dj;
            double f[9], r, vx, vy;
            if constexpr (sizeof(T) == 8) {
                node_state(src, x, y, nx, ny, N, BC, uinx, uiny, f, r, vx, vy);
            } else {
                float g[9];
                Mo o;
                node_state32<BC>(src, x, y, nx, ny, N, (float)(uinx - ref.U), (float)uiny, ref, g, o);
#pragma unroll
                for (int k = 0; k < 9; k++) f[k] = (double)g[k] + ref.c[k];
                ref::macro(f, r, vx, vy);
            }
            double q[9];
            ref::feq(r, vx, vy, q);
#pragma unroll
            for (int k = 0; k < 9; k++) {
                fa[k] += 0.25 * f[k];
                na[k] += 0.25 * (f[k] - q[k]);
            }
        }
    double rho, ux, uy;
    ref::macro(fa, rho, ux, uy);
    mc[nc] = rho;
    mc[Nc + nc] = ux;
    mc[2 * Nc + nc] = uy;
    if (want_f) {
        double q[9];
        ref::feq(rho, ux, uy, q);
#pragma unroll
        for (int k = 0; k < 9; k++) fc[(size_t)k * Nc + nc] = q[k] + s * na[k];
    }
}

// The network's inputs at coarse node p, from p's restricted state: rho,
// ux, uy, |S| (the strain estimate of the LES, from the non-equilibrium
// stress), Pi_xx - Pi_yy and Pi_xy of f - feq.
__device__ __forceinline__ void pair_features(const double* __restrict__ fc, const double* __restrict__ mc,
                                              unsigned Nc, unsigned p, float* __restrict__ feat) {
#pragma clang fp contract(off)
    double d[9], q[9];
    const double rho = mc[p], ux = mc[Nc + p], uy = mc[2 * Nc + p];
    ref::feq(rho, ux, uy, q);
#pragma unroll
    for (int k = 0; k < 9; k++) d[k] = fc[(size_t)k * Nc + p] - q[k];
    const double pxx = ((((d[1] + d[3]) + d[5]) + d[6]) + d[7]) + d[8];
    const double pxy = ((d[5] - d[6]) + d[7]) - d[8];
    const double pyy = ((((d[2] + d[4]) + d[5]) + d[6]) + d[7]) + d[8];
    const double c = -3.0 / (2 * rho);
    const double sxx = pxx * c, sxy = pxy * c, syy = pyy * c;
    feat[6 * p + 0] = (float)rho;
    feat[6 * p + 1] = (float)ux;
    feat[6 * p + 2] = (float)uy;
    feat[6 * p + 3] = (float)sqrt(2.0 * (sxx * sxx + 2.0 * sxy * sxy + syy * syy));
    feat[6 * p + 4] = (float)(pxx - pyy);
    feat[6 * p + 5] = (float)pxy;
}

// Labels: for coarse node p, the relaxation rate omega whose collision at
// the nine upstream nodes q_k = p - e_k, once streamed into p, best matches
// the target mass and momentum (least squares, closed form). This is the
// search of FindAndWriteTau (mesh.cpp:2625-2715: recompute the upstream
// collisions with tau + dtau, stream, compare with the high-fidelity
// moments) solved exactly instead of over 20001 candidates. BGK: post =
// f - omega (f - feq); RR: post = feq + f1 - omega f1 (f1 the regularised
// non-equilibrium of NodeAndLattice.cpp:258-283).
__global__ void __launch_bounds__(BX) k_labels(const double* __restrict__ fc, const double* __restrict__ mc,
                                               const double* __restrict__ tgt, int nxc, int nyc, int rr,
                                               const int4* __restrict__ excl, int nexcl, float* __restrict__ feat,
                                               float* __restrict__ label, float* __restrict__ weight) {
#pragma clang fp contract(off)
    const int I = blockIdx.x * BX + threadIdx.x, J = blockIdx.y;
    if (I >= nxc) return;
    const unsigned Nc = (unsigned)nxc * nyc, p = (unsigned)J * nxc + I;
    pair_features(fc, mc, Nc, p, feat);
    bool ok = I >= 2 && I < nxc - 2 && J >= 2 && J < nyc - 2;
    for (int b = 0; b < nexcl; b++) {
        const int4 e = excl[b];
        if (I >= e.x && I <= e.z && J >= e.y && J <= e.w) ok = false;
    }
    double a0 = 0, a1 = 0, a2 = 0, r0 = 0, r1 = 0, r2 = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const unsigned qn = (unsigned)(J - EY[k]) * nxc + (unsigned)(I - EX[k]);
        const unsigned qq = ok ? qn : p;
        double g[9], e[9];
#pragma unroll
        for (int m = 0; m < 9; m++) g[m] = fc[(size_t)m * Nc + qq];
        const double rq = mc[qq], xq = mc[Nc + qq], yq = mc[2 * Nc + qq];
        ref::feq(rq, xq, yq, e);
        double slope, base;
        if (rr) {
            double d[9];
#pragma unroll
            for (int m = 0; m < 9; m++) d[m] = g[m] - e[m];
            const double nxx = ((((d[1] + d[3]) + d[5]) + d[6]) + d[7]) + d[8];
            const double nxy = ((d[5] - d[6]) + d[7]) - d[8];
            const double nyy = ((((d[2] + d[4]) + d[5]) + d[6]) + d[7]) + d[8];
            const double q0 = EX[k] * EX[k] - 1. / 3, q3 = EY[k] * EY[k] - 1. / 3, q1 = EX[k] * EY[k];
            slope = 4.5 * WD[k] * (q0 * nxx + 2.0 * q1 * nxy + q3 * nyy);
            base = e[k] + slope;
        } else {
            slope = g[k] - e[k];
            base = g[k];
        }
        a0 += slope;
        a1 += EX[k] * slope;
        a2 += EY[k] * slope;
        r0 += base;
        r1 += EX[k] * base;
        r2 += EY[k] * base;
    }
    const double tr = tgt[p], tx = tgt[Nc + p] * tr, ty = tgt[2 * Nc + p] * tr;
    r0 -= tr;
    r1 -= tx;
    r2 -= ty;
    const double aa = a0 * a0 + a1 * a1 + a2 * a2;
    const double om = aa > 0 ? (a0 * r0 + a1 * r1 + a2 * r2) / aa : 0.0;
    // the fit is linear in omega = 1/tau: labels stay in omega (no window on
    // tau, which would cut the noise distribution one-sidedly near tau = 1/2)
    ok = ok && aa > 1e-30 && om > 0.0 && om < 4.0;
    label[p] = ok ? (float)om : 0.0f;
    weight[p] = ok ? (float)aa : 0.0f;
}

// Effective-relaxation-time labels (stress matching). For coarse node p and
// its 2x2 fine block B, with each fine node's post-streaming state, its
// non-equilibrium stress Pi = sum e e (f - feq) and the relaxation time tau_l
// of its next collision (LES included):
//   fine viscous stress  sigma = -(1 - 1/(2 tau_l)) Pi,
//   fine strain          S     = -Pi / (2 rho c_s^2 tau_l),
//   subgrid stress       R     = <rho u u>_B - rho_c u_c u_c,
//   target stress        T     = <sigma>_B - R,
//   coarse strain        C     = 2 <S>_B (coarse time unit = 2 fine),
// the coarse viscosity nu minimising |dev T - 2 rho_c nu dev C|^2 is
// nu = dev T : dev C / (2 rho_c dev C : dev C), and tau = 1/2 + 3 nu: the
// relaxation time whose viscous stress on the coarse grid carries the
// momentum flux of the resolved fine solution (the a-priori dynamic
// procedure of LES). Label omega = 1 / clamp(tau, 0.5005, 2); weight
// dev C : dev C (0 within two coarse cells of a body's support).
template <typename T, int BC>
__global__ void __launch_bounds__(BX) k_stress_labels(const T* __restrict__ src, const double* __restrict__ tauf,
                                                      const double* __restrict__ fc, const double* __restrict__ mc,
                                                      int nx, int ny, double uinx, double uiny, RefState ref,
                                                      const int4* __restrict__ excl, int nexcl,
                                                      float* __restrict__ feat, float* __restrict__ label,
                                                      float* __restrict__ weight) {
#pragma clang fp contract(off)
    const int nxc = nx / 2, nyc = ny / 2;
    const int I = blockIdx.x * BX + threadIdx.x, J = blockIdx.y;
    if (I >= nxc) return;
    const unsigned N = (unsigned)nx * ny, Nc = (unsigned)nxc * nyc, p = (unsigned)J * nxc + I;
    pair_features(fc, mc, Nc, p, feat);
    // block sums: rho, rho u, rho u u, sigma, S
    double sr = 0, smx = 0, smy = 0, rxx = 0, rxy = 0, ryy = 0, gxx = 0, gxy = 0, gyy = 0, cxx = 0, cxy = 0, cyy = 0;
    for (int dj = 0; dj < 2; dj++)
        for (int di = 0; di < 2; di++) {
            const int x = 2 * I + di, y = 2 * J + dj;
            double f[9], r, vx, vy;
            if constexpr (sizeof(T) == 8) {
                node_state(src, x, y, nx, ny, N, BC, uinx, uiny, f, r, vx, vy);
            } else {
                float g[9];
                Mo o;
                node_state32<BC>(src, x, y, nx, ny, N, (float)(uinx - ref.U), (float)uiny, ref, g, o);
#pragma unroll
                for (int k = 0; k < 9; k++) f[k] = (double)g[k] + ref.c[k];
                ref::macro(f, r, vx, vy);
            }
            double q[9], d[9];
            ref::feq(r, vx, vy, q);
#pragma unroll
            for (int k = 0; k < 9; k++) d[k] = f[k] - q[k];
            const double pxx = ((((d[1] + d[3]) + d[5]) + d[6]) + d[7]) + d[8];
            const double pxy = ((d[5] - d[6]) + d[7]) - d[8];
            const double pyy = ((((d[2] + d[4]) + d[5]) + d[6]) + d[7]) + d[8];
            const double tl = tauf[(unsigned)y * nx + x];
            const double gs = -(1.0 - 0.5 / tl), cs = -3.0 / (2.0 * r * tl);
            sr += r;
            smx += r * vx;
            smy += r * vy;
            rxx += r * vx * vx;
            rxy += r * vx * vy;
            ryy += r * vy * vy;
            gxx += gs * pxx;
            gxy += gs * pxy;
            gyy += gs * pyy;
            cxx += cs * pxx;
            cxy += cs * pxy;
            cyy += cs * pyy;
        }
    const double rc = 0.25 * sr, ux = smx / sr, uy = smy / sr;
    // T = <sigma> - (<rho u u> - rho_c u_c u_c); C = 2 <S>
    const double txx = 0.25 * gxx - (0.25 * rxx - rc * ux * ux);
    const double txy = 0.25 * gxy - (0.25 * rxy - rc * ux * uy);
    const double tyy = 0.25 * gyy - (0.25 * ryy - rc * uy * uy);
    const double Cxx = 0.5 * cxx, Cxy = 0.5 * cxy, Cyy = 0.5 * cyy;
    const double tdx = 0.5 * (txx - tyy), cdx = 0.5 * (Cxx - Cyy);  // deviatoric xx (= -yy)
    const double num = 2.0 * tdx * cdx + 2.0 * txy * Cxy;
    const double den = 2.0 * cdx * cdx + 2.0 * Cxy * Cxy;
    bool ok = den > 0.0;
    for (int b = 0; b < nexcl; b++) {
        const int4 e = excl[b];
        if (I >= e.x && I <= e.z && J >= e.y && J <= e.w) ok = false;
    }
    const double tau = ok ? 0.5 + 3.0 * (num / (2.0 * rc * den)) : 1.0;
    const double tcl = fmin(fmax(tau, 0.5005), 2.0);
    label[p] = ok ? (float)(1.0 / tcl) : 0.0f;
    weight[p] = ok ? (float)den : 0.0f;
}

// ----------------------------------------------------------- init kernels
// rho, u from the macro arrays (mode 0) or from rho0/u0 + seeded noise
// (mode 1); f = feq (InitialLatticeFeq form), stored as the fnew state.
template <typename T>
__global__ void k_init(T* __restrict__ dst, double* __restrict__ mac, size_t N, int mode,
                       unsigned long long seed, double rho0, double ux0, double uy0, double amp, RefState ref) {
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
            dst[k * N + n] = (float)(f[k] - ref.c[k]);
    }
}

// Post-streaming populations and moments of the current state (reads only);
// f_out may be null. FP32 moments are accumulated in FP64 from the stored
// deviations.
// Running means from the macro arrays (state before the first step).
__global__ void k_mean_phase0(const double* __restrict__ mac, double* __restrict__ mean, size_t N, double kk) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= N) return;
#pragma unroll
    for (int c = 0; c < 3; c++) mean[c * N + i] = (mean[c * N + i] * kk + mac[c * N + i]) / (kk + 1.0);
}

// Saves the populations of the flat-plate corners (0, 0) and (nx-1, 0)
// from the stored initial state (full f, FP64): no rule updates them, so
// they are the populations these nodes keep.
template <typename T>
__global__ void k_corner(const T* __restrict__ src, double* __restrict__ corner, int nx, unsigned N, RefState ref) {
    const int t = threadIdx.x;
    if (t >= 18) return;
    const int k = t % 9;
    double v = (double)src[(size_t)k * N + (t < 9 ? 0u : (unsigned)(nx - 1))];
    if constexpr (sizeof(T) == 4) v += ref.c[k];
    corner[t] = v;
}

// Populations (full f, FP64) and macros rho, u of node (x, y) of the
// current post-streaming state, as the read-back reports them: FP32
// deviations are accumulated in FP64. The flat-plate corners keep their
// initial populations and the macros in mac.
template <typename T, int BC>
__device__ __forceinline__ void macro_at(const T* __restrict__ src, const double* __restrict__ mac,
                                         const double* __restrict__ corner, int x, int y, int nx, int ny,
                                         double uinx, double uiny, int plate_x, const RefState& ref, double* f,
                                         double& rho, double& ux, double& uy) {
#pragma clang fp contract(off)
    const unsigned N = (unsigned)nx * ny;
    const unsigned n = (unsigned)y * nx + x;
    if constexpr (BC == IPPI_BC_FLATPLATE) {
        if (y == 0 && (x == 0 || x == nx - 1)) {  // corner: state of the initialisation, macros kept in mac
#pragma unroll
            for (int k = 0; k < 9; k++) f[k] = corner[9 * (x != 0) + k];
            rho = mac[n];
            ux = mac[N + n];
            uy = mac[2 * N + n];
            return;
        }
        if constexpr (sizeof(T) == 8) {
            fp_state64(src, x, y, nx, ny, N, uinx, uiny, plate_x, f, rho, ux, uy);
        } else {
            float g[9];
            Mo o;
            fp_state32(src, x, y, nx, ny, N, (float)(uinx - ref.U), (float)uiny, plate_x, ref, g, o);
            const bool assigned = x == 0 || x == nx - 1 || y == ny - 1 || (y == 0 && x >= plate_x);
            double gd[9], sm = 0.0;
#pragma unroll
            for (int k = 0; k < 9; k++) gd[k] = (double)g[k];
            for (int k = 0; k < 9; k++) sm += gd[k];
            const double jx = ((gd[1] - gd[3]) + (gd[5] - gd[7])) + (gd[8] - gd[6]);
            const double jy = ((gd[2] - gd[4]) + (gd[5] - gd[7])) + (gd[6] - gd[8]);
            rho = assigned ? 1.0 + (double)o.drho : 1.0 + sm;
            ux = assigned ? ref.U + (double)o.dux : (ref.U + jx) / rho;
            uy = assigned ? (double)o.duy : jy / rho;
#pragma unroll
            for (int k = 0; k < 9; k++) f[k] = gd[k] + ref.c[k];
        }
    } else if constexpr (sizeof(T) == 8) {
        node_state(src, x, y, nx, ny, N, BC, uinx, uiny, f, rho, ux, uy);
    } else {
        float g[9];
        Mo o;
        node_state32<BC>(src, x, y, nx, ny, N, (float)(uinx - ref.U), (float)uiny, ref, g, o);
        double s = 0.0, jx, jy;
        double gd[9];
#pragma unroll
        for (int k = 0; k < 9; k++) gd[k] = (double)g[k];
        for (int k = 0; k < 9; k++) s += gd[k];
        jx = ((gd[1] - gd[3]) + (gd[5] - gd[7])) + (gd[8] - gd[6]);
        jy = ((gd[2] - gd[4]) + (gd[5] - gd[7])) + (gd[6] - gd[8]);
        const bool nee = BC == IPPI_BC_INFLOW && (x == 0 || x == nx - 1);
        rho = nee ? 1.0 : 1.0 + s;
        ux = nee ? ref.U + (double)o.dux : (ref.U + jx) / rho;
        uy = nee ? (double)o.duy : jy / rho;
#pragma unroll
        for (int k = 0; k < 9; k++) f[k] = gd[k] + ref.c[k];
    }
}

template <typename T, int BC>
__global__ void k_macro(const T* __restrict__ src, double* __restrict__ mac, double* __restrict__ f_out, int nx,
                        int ny, double uinx, double uiny, int plate_x, const double* __restrict__ corner,
                        double* __restrict__ mean, double kk, RefState ref) {
#pragma clang fp contract(off)
    const int x = blockIdx.x * BX + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= nx) return;
    const unsigned N = (unsigned)nx * ny;
    const unsigned n = (unsigned)y * nx + x;
    double f[9], rho, ux, uy;
    macro_at<T, BC>(src, mac, corner, x, y, nx, ny, uinx, uiny, plate_x, ref, f, rho, ux, uy);
    mac[n] = rho;
    mac[N + n] = ux;
    mac[2 * N + n] = uy;
    if (mean) {  // running means, ComputeAvgVelocity (mesh.cpp:1954-1974)
        mean[n] = (mean[n] * kk + rho) / (kk + 1.0);
        mean[N + n] = (mean[N + n] * kk + ux) / (kk + 1.0);
        mean[2 * N + n] = (mean[2 * N + n] * kk + uy) / (kk + 1.0);
    }
    if (f_out) {
#pragma unroll
        for (int k = 0; k < 9; k++) f_out[(size_t)k * N + n] = f[k];
    }
}

// ------------------------------------------- boundary-row / column networks
// The reference's per-node networks at the domain's edge rows and columns
// (Mesh::NNmodel_Type01/03/05/07, mesh.cpp:2767-2898, 3128-3260): inputs
// (u.x, u.y, rho) of five neighbours as float, Linear(15,128) -> 6 x
// Linear(128,128) -> Linear(128,1) in FP32, ReLU between layers; the node
// takes the output as its tau where it is >= 0.501 (mesh.cpp:2821-2827).
// One block of BNN_T threads serves BNN_NODES nodes: thread t computes
// neuron t % 128 for the nodes of half t / 128. Each layer's weights are
// staged through LDS in chunks of BNN_KC inputs with 16-byte loads, so a
// block reads them once and many loads are in flight together; every
// output is a chain of fused multiply-adds over the inputs in increasing
// order, so the result is that of a sequential FP32 loop bit for bit.
// Weights are stored input-major (Wt[k][o]).
constexpr int BNN_NIN = 15, BNN_NH = 128, BNN_NHID = 6;
constexpr int BNN_NPARAM = BNN_NIN * BNN_NH + BNN_NH + BNN_NHID * (BNN_NH * BNN_NH + BNN_NH) + BNN_NH + 1;  // 101249
constexpr int BNN_T = 2 * BNN_NH, BNN_NODES = 16, BNN_HALF = BNN_NODES / 2, BNN_KC = 64;

struct BnnNode {
    int x, y, slot, t;  // node, index into the tau arrays (-1: padding), type index (0: T01, 1: T03, 2: T05, 3: T07)
};
// The node list holds the set types one after the other, each padded to a
// multiple of 16 entries (copies of its last node with slot -1), so a tile
// of 16 nodes has one type.
constexpr int BNN_PAD = 16;
struct BnnDev {
    const BnnNode* nodes;
    int nnodes;
    const float* w[4];   // per type, device layout (input-major), nullptr when not set
    double* tau;         // [2 nx + 2 ny]
    int tend[4];         // end of each type's entries in the list (types not set: empty ranges)
    const float* wm[4];  // per type, the MFMA kernel's layout (bnm_layout), nullptr when not set
};
// The MFMA kernel's copy of a network, ordered so that each 16-byte load of
// a wave reads 1 KB in one piece. Lane l = 16 kq + r of wave w takes the
// weights of output 16 w + r and inputs 4 j + kq. Layer 0: [w 8][l 64][j 4]
// = W0[4j + kq][16 w + r] (zero for input 15), then its 128 biases; hidden
// layer h (1..6) at BNM_OFF_H1 + (h - 1) BNM_LAYER_H: [w 8][q 8][l 64][e 4]
// = Wh[4 (4q + e) + kq][16 w + r], then its 128 biases; the output row and
// bias (129) at BNM_OFF_OUT.
constexpr int BNM_OFF_B0 = 8 * 4 * 16 * 4, BNM_LAYER_H = 8 * 4 * 16 * 32 + BNN_NH;
constexpr int BNM_OFF_H1 = BNM_OFF_B0 + BNN_NH, BNM_OFF_OUT = BNM_OFF_H1 + BNN_NHID * BNM_LAYER_H;
constexpr int BNM_NPARAM = BNM_OFF_OUT + BNN_NH + 1;
static_assert(BNM_OFF_H1 % 4 == 0 && BNM_LAYER_H % 4 == 0, "16-byte loads of the MFMA layout");

// neighbour offsets in the reference's input order (mesh.cpp:3149-3165,
// 2788-2804, 3216-3232, 2854-2870)
__constant__ int BNN_DX[4][5] = {{0, 1, 1, 1, 0}, {1, 1, 0, -1, -1}, {0, -1, -1, -1, 0}, {-1, -1, 0, 1, 1}};
__constant__ int BNN_DY[4][5] = {{1, 1, 0, -1, -1}, {0, -1, -1, -1, 0}, {1, 1, 0, -1, -1}, {0, 1, 1, 1, 0}};

// LDS copy of rows [k0, k0 + nk) of an input-major weight matrix (nk x 128)
__device__ __forceinline__ void bnn_stage(const float* __restrict__ Wt, int k0, int nk, float* lw) {
    const float4* src = reinterpret_cast<const float4*>(Wt + (size_t)k0 * BNN_NH);
    float4* dst = reinterpret_cast<float4*>(lw);
    const int n4 = nk * BNN_NH / 4;
    for (int i = threadIdx.x; i < n4; i += BNN_T) dst[i] = src[i];
}

template <typename T, int BC>
__global__ void __launch_bounds__(BNN_T) k_bnn(const T* __restrict__ src, const double* __restrict__ mac,
                                              const double* __restrict__ corner, int pull, StepArgs a, BnnDev d) {
    __shared__ __attribute__((aligned(16))) float lw[BNN_KC * BNN_NH];  // staged weights
    __shared__ __attribute__((aligned(16))) float act[2][BNN_NH][BNN_NODES];  // [k][node]
    __shared__ __attribute__((aligned(16))) float xin[BNN_NIN][BNN_NODES];
    __shared__ int typ[BNN_NODES];
    const int t = threadIdx.x, o = t % BNN_NH, h0 = (t / BNN_NH) * BNN_HALF;
    const int i0 = blockIdx.x * BNN_NODES;
    const int nx = a.nx, ny = a.ny;
    if (t < BNN_NODES) typ[t] = d.nodes[min(i0 + t, d.nnodes - 1)].t;
    // inputs: (u.x, u.y, rho) of the 5 neighbours of each node, as float
    if (t < 5 * BNN_NODES) {
        const int j = t / BNN_NODES, q = t % BNN_NODES;
        const BnnNode b = d.nodes[min(i0 + q, d.nnodes - 1)];
        int xn = b.x + BNN_DX[b.t][j], yn = b.y + BNN_DY[b.t][j];
        xn = xn < 0 ? xn + nx : (xn >= nx ? xn - nx : xn);  // the ring's ghost columns
        double f[9], rho, ux, uy;
        if (pull) {
            macro_at<T, BC>(src, mac, corner, xn, yn, nx, ny, a.uinx, a.uiny, a.plate_x, a.ref, f, rho, ux, uy);
        } else {  // first step: the macros of the initial state
            const unsigned N = (unsigned)nx * ny, n = (unsigned)yn * nx + xn;
            rho = mac[n];
            ux = mac[N + n];
            uy = mac[2 * N + n];
        }
        xin[3 * j][q] = (float)ux;
        xin[3 * j + 1][q] = (float)uy;
        xin[3 * j + 2][q] = (float)rho;
    }
    __syncthreads();
    // a block's nodes may be of two types: each type's weights are applied
    // to all of them and kept for that type's nodes
    for (int tp = 0; tp < 4; tp++) {
        bool any = false;
#pragma unroll
        for (int q = 0; q < BNN_NODES; q++) any = any || typ[q] == tp;
        if (!any || !d.w[tp]) continue;  // block-uniform
        const float* W = d.w[tp];
        float v[BNN_HALF];
        // layer 0: 15 inputs
        __syncthreads();  // the previous pass's readers of lw and act are done
        bnn_stage(W, 0, BNN_NIN, lw);
        __syncthreads();
        {
            const float bo = W[BNN_NIN * BNN_NH + o];
#pragma unroll
            for (int q = 0; q < BNN_HALF; q++) v[q] = bo;
#pragma unroll
            for (int k = 0; k < BNN_NIN; k++) {
                const float w = lw[k * BNN_NH + o];
#pragma unroll
                for (int q = 0; q < BNN_HALF; q++) v[q] = __builtin_fmaf(w, xin[k][h0 + q], v[q]);
            }
        }
#pragma unroll
        for (int q = 0; q < BNN_HALF; q++) act[0][o][h0 + q] = v[q] > 0.0f ? v[q] : 0.0f;
        size_t off = BNN_NIN * BNN_NH + BNN_NH;
        int cur = 0;
        for (int l = 0; l < BNN_NHID; l++) {
            const float* Wl = W + off;
            const float bo = Wl[BNN_NH * BNN_NH + o];
#pragma unroll
            for (int q = 0; q < BNN_HALF; q++) v[q] = bo;
            for (int k0 = 0; k0 < BNN_NH; k0 += BNN_KC) {
                __syncthreads();  // lw free, and act[cur] written
                bnn_stage(Wl, k0, BNN_KC, lw);
                __syncthreads();
#pragma unroll 8
                for (int k = 0; k < BNN_KC; k++) {
                    const float w = lw[k * BNN_NH + o];
#pragma unroll
                    for (int q = 0; q < BNN_HALF; q++) v[q] = __builtin_fmaf(w, act[cur][k0 + k][h0 + q], v[q]);
                }
            }
#pragma unroll
            for (int q = 0; q < BNN_HALF; q++) act[cur ^ 1][o][h0 + q] = v[q] > 0.0f ? v[q] : 0.0f;
            cur ^= 1;
            off += BNN_NH * BNN_NH + BNN_NH;
        }
        __syncthreads();
        if (t < BNN_NODES && typ[t] == tp && i0 + t < d.nnodes && d.nodes[i0 + t].slot >= 0) {
            const float* Wl = W + off;
            float r = Wl[BNN_NH];
            for (int k = 0; k < BNN_NH; k++) r = __builtin_fmaf(Wl[k], act[cur][k][t], r);
            const double tau = (double)r;
            d.tau[d.nodes[i0 + t].slot] = tau >= 0.501 ? tau : 0.0;
        }
    }
}

// The same networks on FP32 MFMA tiles (v_mfma_f32_16x16x4_f32), for FP32
// and FP64 handles: a block of 8 waves serves one tile of 16 nodes of one
// type; wave w computes outputs 16w..16w+15 of every hidden layer, with
// FP32 products and FP32 accumulation, the bias as the accumulator's start.
// Lane l of an MFMA supplies A[l%16][l/16] (the weight of output 16w + l%16
// and input 4j + l/16, read from L2), B[l/16][l%16] (the input activation
// of node l%16, from LDS) and receives D[4(l/16)+i][l%16]. Activations
// ping-pong between two LDS arrays [128][16]; the output layer is a
// sequential FMA chain per node. An FP32 MFMA adds its four products to the
// accumulator one after the other with a rounding each, so a chain of
// MFMAs over k = 0, 4, 8, ... from the bias is k_bnn's FMA chain: the
// outputs are bit for bit k_bnn's, and the oracle's (asserted in the
// tests). The block's type follows from its index, so its first weights
// are requested before the neighbour gather; each wave holds its 16
// outputs' weights and biases of the next layer in registers, loaded at the
// start of the current layer (the scheduling fence keeps the compiler from
// sinking those loads to the layer's end). One tile per block: the ~5,800
// edge nodes of a 2401 x 501 lattice make 363 blocks for the 256 CUs
// (64-node blocks left 165 idle); at 124 VGPRs two blocks share a CU, so
// one block's gather and barriers overlap the other's MFMAs.
constexpr int BNM_T = 512, BNM_NODES = BNN_PAD, BNM_LD = BNM_NODES + 1;  // activation rows padded: no bank conflicts
typedef float f32x4b __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void bnm_ld4(const float* __restrict__ p, float* v) {  // one 16-byte load
    const f32x4b q = *reinterpret_cast<const f32x4b*>(p);
#pragma unroll
    for (int e = 0; e < 4; e++) v[e] = q[e];
}
// One hidden layer: reads X[P ^ 1], writes X[P]; its A operands and biases
// in av, bv, which then take the next layer's from Wn (if not null).
template <int P>
__device__ __forceinline__ void bnm_hidden(float (*X)[BNN_NH][BNM_LD], float (&av)[BNN_NH / 4], float (&bv)[4],
                                           const float* __restrict__ Wn, int w, int l, int bofs) {
    const int kq = l >> 4;
    f32x4b c;
#pragma unroll
    for (int i = 0; i < 4; i++) c[i] = bv[i];
    float avc[BNN_NH / 4];
#pragma unroll
    for (int j = 0; j < BNN_NH / 4; j++) avc[j] = av[j];
    if (Wn) {  // block-uniform
#pragma unroll
        for (int q = 0; q < BNN_NH / 16; q++) bnm_ld4(Wn + 4 * ((8 * w + q) * 64 + l), av + 4 * q);
        bnm_ld4(Wn + BNM_LAYER_H - BNN_NH + bofs, bv);
    }
    __builtin_amdgcn_sched_barrier(0);  // the loads stay here, a layer ahead of their MFMAs
    // the B operands of 16 k-steps read together, then their MFMAs
#pragma unroll
    for (int j0 = 0; j0 < BNN_NH / 4; j0 += 16) {
        float b[16];
#pragma unroll
        for (int jj = 0; jj < 16; jj++) b[jj] = X[P ^ 1][4 * (j0 + jj) + kq][l & 15];
#pragma unroll
        for (int jj = 0; jj < 16; jj++) c = __builtin_amdgcn_mfma_f32_16x16x4f32(avc[j0 + jj], b[jj], c, 0, 0, 0);
        __builtin_amdgcn_sched_group_barrier(0x100, 16, 0);  // LDS reads
        __builtin_amdgcn_sched_group_barrier(0x008, 16, 0);  // then the MFMAs
    }
    // ReLU: outputs 16 w + 4 kq + i of node l % 16
#pragma unroll
    for (int i = 0; i < 4; i++) X[P][bofs + i][l & 15] = fmaxf(c[i], 0.0f);
    __syncthreads();  // this layer's outputs written and its inputs read by every wave
}

template <typename T, int BC>
__global__ void __launch_bounds__(BNM_T) k_bnn_mfma(const T* __restrict__ src, const double* __restrict__ mac,
                                                   const double* __restrict__ corner, int pull, StepArgs a,
                                                   BnnDev d) {
    __shared__ __attribute__((aligned(16))) float xin[16][BNM_LD];       // layer-0 inputs (row 15: zero)
    __shared__ __attribute__((aligned(16))) float X[2][BNN_NH][BNM_LD];  // activations [k][node]
    const int t = threadIdx.x, l = t & 63, w = t >> 6;
    const int i0 = blockIdx.x * BNM_NODES;
    const int nx = a.nx, ny = a.ny;
    int tp = 0;  // the tile's type (block-uniform)
    while (tp < 3 && i0 >= d.tend[tp]) tp++;
    const float* W = d.wm[tp];
    const int kq = l >> 4;             // this lane's A column / B row within a k-step
    const int bofs = 16 * w + 4 * kq;  // this lane's four outputs (biases)
    // node i0 + t % 16 of the list, computed as the host lays it out (each
    // type's row or column in order, padded with copies of its last node
    // whose slot is -1), so the gather starts without a load
    BnnNode b;
    {
        const int q0 = i0 + (t & (BNM_NODES - 1)) - (tp == 0 ? 0 : d.tend[tp - 1]);
        const bool col = tp == 0 || tp == 2;
        const int cnt = col ? ny - 2 : nx, q = min(q0, cnt - 1);
        b.x = col ? (tp == 0 ? 0 : nx - 1) : q;
        b.y = col ? 1 + q : (tp == 3 ? 1 : ny - 2);
        b.slot = q0 >= cnt ? -1 : (col ? 2 * nx + (tp == 2 ? ny : 0) + b.y : (tp == 1 ? nx : 0) + b.x);
        b.t = tp;
    }
    // layer 0 (K = 15, padded to 16 with a zero weight and a zero input) and the first hidden
    // layer's weights and biases, requested before the gather
    float a0[4], b0[4], av[BNN_NH / 4], bv[4];
    bnm_ld4(W + 4 * (64 * w + l), a0);
    bnm_ld4(W + BNM_OFF_B0 + bofs, b0);
#pragma unroll
    for (int q = 0; q < BNN_NH / 16; q++) bnm_ld4(W + BNM_OFF_H1 + 4 * ((8 * w + q) * 64 + l), av + 4 * q);
    bnm_ld4(W + BNM_OFF_H1 + BNM_LAYER_H - BNN_NH + bofs, bv);
    // inputs: (u.x, u.y, rho) of the 5 neighbours of each node, as float
    if (t < 5 * BNM_NODES) {
        const int j = t / BNM_NODES, q = t % BNM_NODES;
        int xn = b.x + BNN_DX[b.t][j], yn = b.y + BNN_DY[b.t][j];
        xn = xn < 0 ? xn + nx : (xn >= nx ? xn - nx : xn);  // the ring's ghost columns
        double f[9], rho, ux, uy;
        if (pull) {
            macro_at<T, BC>(src, mac, corner, xn, yn, nx, ny, a.uinx, a.uiny, a.plate_x, a.ref, f, rho, ux, uy);
        } else {  // first step: the macros of the initial state
            const unsigned N = (unsigned)nx * ny, n = (unsigned)yn * nx + xn;
            rho = mac[n];
            ux = mac[N + n];
            uy = mac[2 * N + n];
        }
        xin[3 * j][q] = (float)ux;
        xin[3 * j + 1][q] = (float)uy;
        xin[3 * j + 2][q] = (float)rho;
    } else if (t < 6 * BNM_NODES) {
        xin[BNN_NIN][t - 5 * BNM_NODES] = 0.0f;
    }
    __syncthreads();
    {  // layer 0 -> X[0]
        f32x4b c;
#pragma unroll
        for (int i = 0; i < 4; i++) c[i] = b0[i];
#pragma unroll
        for (int j = 0; j < 4; j++) c = __builtin_amdgcn_mfma_f32_16x16x4f32(a0[j], xin[4 * j + kq][l & 15], c, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; i++) X[0][bofs + i][l & 15] = fmaxf(c[i], 0.0f);
        __syncthreads();
    }
    // hidden layers in pairs (odd: X[0] -> X[1], even: X[1] -> X[0])
#pragma unroll
    for (int hl = 1; hl < BNN_NHID; hl += 2) {
        const float* Wh = W + BNM_OFF_H1 + hl * BNM_LAYER_H;  // hidden layer hl + 1
        bnm_hidden<1>(X, av, bv, hl + 1 <= BNN_NHID ? Wh : nullptr, w, l, bofs);
        bnm_hidden<0>(X, av, bv, hl + 2 <= BNN_NHID ? Wh + BNM_LAYER_H : nullptr, w, l, bofs);
    }
    const int cur = 0;  // the last hidden layer's outputs
    if (t < BNM_NODES) {
        const int slot = b.slot;  // b is node i0 + t for t < 16 (-1: padding)
        const float* Wl = W + BNM_OFF_OUT;
        float r = Wl[BNN_NH];
        for (int k = 0; k < BNN_NH; k++) r = __builtin_fmaf(Wl[k], X[cur][k][t], r);
        const double tau = (double)r;
        if (slot >= 0) d.tau[slot] = tau >= 0.501 ? tau : 0.0;
    }
}

// ------------------------------------------------------------- IB kernel
// One kernel per step for the immersed boundary, wired before the
// collision:
//   InterpolateVel (4-point delta, mesh.cpp:555-605)
//   CalParticleForce (DirectForce, mesh.cpp:514-532)
//   SpreadForce (4-point delta, mesh.cpp:659-741), as a gather
// A body's "active" nodes are the union of its markers' 4x4 supports; they
// are exactly the nodes the spread reaches. Each block serves one body and
// a slice of its active nodes:
//   A. the post-streaming velocity of every active node of the body goes to
//      LDS (four nodes in flight per lane);
//   B. one lane per marker sums u(X,Y) phi phi over its 16 support nodes in
//      the reference's X-outer / Y-inner order (node indices and weights were
//      prefetched before A), F = 2 (U_d - U_interp) to LDS;
//   C. each active node of the slice sums its (marker, weight) list in
//      increasing marker order, the order of the reference's marker loop.
// The first block of a body writes the marker state and the body-force
// record (minus the sum of the marker forces, fixed-order tree in LDS).
constexpr int IB_T = 512;        // threads per IB block = active nodes per block
constexpr int IB_MAXM = 1024;    // markers per body
constexpr int IB_MAXACT = 6144;  // active nodes per body (96 KB of LDS velocities: gfx950 gives a block up to 160 KB)

// Everything a block needs, fetched with one scalar load.
struct alignas(16) IbBlk {
    int body, a0, a1, m0;      // body, slice of active nodes [a0, a1), first marker
    int M, act0, nact, lead;   // markers, first active node of the body, active count, first block of body
    int x0, y0, w, pad;        // patch origin, width and offset in the patch numbering
};

struct IbDev {
    const double* w;     // [m][16] phi(x-X) phi(y-Y), X-outer / Y-inner
    const int* pidx;     // [m][16] support node -> index in the body's active list
    const double* ud;    // [m][2] desired velocity
    double* ui;          // [m][2] interpolated velocity
    double* F;           // [m][2] marker force
    const IbBlk* blk;    // per block
    const int* act;      // active node -> patch node (global patch numbering)
    const int* ecnt;     // entries per active node
    const int* emk;      // [active][E] marker index relative to the body's first marker
    const double* ecw;   // [active][E] weight (0 in padding)
    double* patch_f;     // [patch node][2]
    double* hist;        // [body][2] of this step's record
    RefState ref;        // FP32 storage offsets (the interpolated deviation gets U back)
};

// Velocity of node (X, Y) for the interpolation: FP64 the full velocity;
// FP32 its deviation (u - U, u_y) from the reference state, to which the
// interpolation adds U at the end (the 16 weights of a marker sum to 1).
template <typename T, bool PULL>
__device__ __forceinline__ void support_u(const T* __restrict__ src, const double* __restrict__ mac, int X, int Y,
                                          int nx, unsigned N, double& ux, double& uy, const RefState& rs) {
#pragma clang fp contract(off)
    // supports lie in the interior, so both boundary kinds pull the plain stencil
    const unsigned n = (unsigned)Y * nx + X;
    if (PULL) {
        T f[9];
        const unsigned m = n - nx, p = n + nx;
        f[0] = ld(src, n);
        f[1] = ld(src + 1 * N, n - 1);
        f[3] = ld(src + 3 * N, n + 1);
        f[2] = ld(src + 2 * N, m);
        f[4] = ld(src + 4 * N, p);
        f[5] = ld(src + 5 * N, m - 1);
        f[6] = ld(src + 6 * N, m + 1);
        f[7] = ld(src + 7 * N, p + 1);
        f[8] = ld(src + 8 * N, p - 1);
        if constexpr (sizeof(T) == 8) {
            double r;
            ref::macro(f, r, ux, uy);
        } else {
            Mo o;
            moments32(f, rs.Uf, o);
            ux = o.dux;
            uy = o.duy;
        }
    } else {
        if constexpr (sizeof(T) == 8) {
            ux = mac[N + n];
            uy = mac[2 * N + n];
        } else {  // as the FP32 step's first collision reads them
            ux = (double)(float)(mac[N + n] - rs.U);
            uy = (double)(float)mac[2 * N + n];
        }
    }
}

// SLOT: the pulled populations of the active nodes were copied into
// slot[k][active node] by the step kernel that stored them (one coalesced
// load each instead of the node list and the stencil gathers).
// Velocity from the nine pulled populations, as support_u computes it.
template <typename T>
__device__ __forceinline__ void slot_u(const T* f, double& ux, double& uy, const RefState& rs) {
#pragma clang fp contract(off)
    if constexpr (sizeof(T) == 8) {
        double r;
        ref::macro(f, r, ux, uy);
    } else {
        Mo o;
        moments32(f, rs.Uf, o);
        ux = o.dux;
        uy = o.duy;
    }
}

// gather: the step kernel gathers the patch forces from F itself, so only
// the first block of each body runs (marker forces and record), without
// the spread.
template <typename T, bool PULL, int E, bool SLOT>
__global__ void __launch_bounds__(IB_T) k_ib(const T* __restrict__ src, const double* __restrict__ mac, IbDev d,
                                             int nx, int ny, const T* __restrict__ slot, int nact_all, int gather) {
#pragma clang fp contract(off)
    __shared__ double U[2 * IB_MAXACT];
    __shared__ double FB[2 * IB_MAXM];
    const IbBlk bk = d.blk[blockIdx.x];
    if (gather && !bk.lead) return;  // block-uniform
    const int t = threadIdx.x;
    const unsigned N = (unsigned)nx * ny;
    // prefetch everything that does not depend on the populations
    const int mt = min(t, bk.M - 1);  // this lane's first marker (clamped)
    int pi[16];
    double pw[16];
#pragma unroll
    for (int s = 0; s < 16; s++) {
        pi[s] = d.pidx[16 * (bk.m0 + mt) + s];
        pw[s] = d.w[16 * (bk.m0 + mt) + s];
    }
    const double udx = d.ud[2 * (bk.m0 + mt)], udy = d.ud[2 * (bk.m0 + mt) + 1];
    const int ia = min(bk.a0 + t, bk.a1 - 1);  // this lane's active node of the slice (clamped)
    const int pn = d.act[ia];
    const int cnt = d.ecnt[ia];
    int em[E];
    double ew[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
        em[e] = d.emk[(size_t)E * ia + e];
        ew[e] = d.ecw[(size_t)E * ia + e];
    }
    // A. velocity at every active node of the body, four nodes in flight per lane
    if (SLOT) {
        for (int c0 = 0; c0 < bk.nact; c0 += 4 * IB_T) {
            double ux[4], uy[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int i = bk.act0 + min(c0 + j * IB_T + t, bk.nact - 1);
                T f[9];
#pragma unroll
                for (int k = 0; k < 9; k++) f[k] = slot[(size_t)k * nact_all + i];
                slot_u(f, ux[j], uy[j], d.ref);
            }
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int i = c0 + j * IB_T + t;
                if (i < bk.nact) {
                    U[2 * i] = ux[j];
                    U[2 * i + 1] = uy[j];
                }
            }
        }
    }
    for (int c0 = 0; c0 < (SLOT ? 0 : bk.nact); c0 += 4 * IB_T) {
        int nd[4];
#pragma unroll
        for (int j = 0; j < 4; j++) nd[j] = d.act[bk.act0 + min(c0 + j * IB_T + t, bk.nact - 1)];
        double ux[4], uy[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int r = nd[j] - bk.pad;
            support_u<T, PULL>(src, mac, bk.x0 + r % bk.w, bk.y0 + r / bk.w, nx, N, ux[j], uy[j], d.ref);
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int i = c0 + j * IB_T + t;
            if (i < bk.nact) {
                U[2 * i] = ux[j];
                U[2 * i + 1] = uy[j];
            }
        }
    }
    __syncthreads();
    // B. interpolation and direct forcing, one lane per marker
    for (int ml = t; ml < bk.M; ml += IB_T) {
        double ux0 = udx, uy0 = udy;
        if (ml != t) {
#pragma unroll
            for (int s = 0; s < 16; s++) {
                pi[s] = d.pidx[16 * (bk.m0 + ml) + s];
                pw[s] = d.w[16 * (bk.m0 + ml) + s];
            }
            ux0 = d.ud[2 * (bk.m0 + ml)];
            uy0 = d.ud[2 * (bk.m0 + ml) + 1];
        }
        double uix = 0, uiy = 0;
#pragma unroll
        for (int s = 0; s < 16; s++) {
            uix += (U[2 * pi[s]] * pw[s]);
            uiy += (U[2 * pi[s] + 1] * pw[s]);
        }
        if constexpr (sizeof(T) == 4) uix = uix + d.ref.U;  // FP32: deviations were interpolated
        const double Fx = 2 * (ux0 - uix), Fy = 2 * (uy0 - uiy);
        FB[2 * ml] = Fx;
        FB[2 * ml + 1] = Fy;
        if (bk.lead) {
            const int m = bk.m0 + ml;
            d.ui[2 * m] = uix;
            d.ui[2 * m + 1] = uiy;
            d.F[2 * m] = Fx;
            d.F[2 * m + 1] = Fy;
        }
    }
    __syncthreads();
    // C. spread (gather) for this block's slice of active nodes
    if (!gather && bk.a0 + t < bk.a1) {
        double fx = 0.0, fy = 0.0;
#pragma unroll
        for (int e = 0; e < E; e++) {
            if (e < cnt) {
                fx += (FB[2 * em[e]] * ew[e]);
                fy += (FB[2 * em[e] + 1] * ew[e]);
            }
        }
        d.patch_f[2 * pn] = fx;
        d.patch_f[2 * pn + 1] = fy;
    }
    if (bk.lead && d.hist) {
        // fixed-order tree sum of the marker forces (reuses U)
        __syncthreads();
        double sx = 0.0, sy = 0.0;
        for (int m = t; m < bk.M; m += IB_T) {
            sx += FB[2 * m];
            sy += FB[2 * m + 1];
        }
        U[2 * t] = sx;
        U[2 * t + 1] = sy;
        __syncthreads();
        for (int off = IB_T / 2; off > 0; off >>= 1) {
            if (t < off) {
                U[2 * t] += U[2 * (t + off)];
                U[2 * t + 1] += U[2 * (t + off) + 1];
            }
            __syncthreads();
        }
        if (t == 0) {
            d.hist[2 * bk.body] = -U[0];
            d.hist[2 * bk.body + 1] = -U[1];
        }
    }
}

// ------------------------------------------------- IB inside the step kernel
// The IB kernel's work (lead blocks, gather mode) moved into the FP32 step
// kernel, so that it overlaps the lattice update instead of preceding it.
// Grid row 0 holds one block per body, dispatched first; it pulls the
// post-streaming populations of the body's active nodes from the kernel's
// input (read-only here), computes the marker forces, stores them
// write-through at agent scope and then sets the body's flag to this step's
// sequence number. The lattice rows follow in rotated order, which places
// the band of rows whose nodes read marker forces (the bodies' patches) at a
// set point of the order; only those blocks wait for the flags, with a bounded
// wait (a timeout sets the divergence flag to 2 and the step is reported as
// failed, never left hanging). The arithmetic, the summation orders and the
// record's tree are those of k_ib, so the results are bit for bit those of
// the separate kernel.
constexpr int INK_MAXACT = 2048;     // active nodes per body, lattice-only kernel (16 KB of float velocities)
constexpr int INK_MAXACT_NN = IPPI_INK_MAXACT_NN;  // with the network (fewer resident blocks, more LDS each)
constexpr int INK_MAXACT64 = 1280;   // FP64 (20 KB of double velocities)
// wait limit in polls of the flag; each poll is a load from memory (~0.5-1
// us) and a short sleep, so the limit is some 30-70 ms: hundreds of times
// the longest step kernel, so only a lost flag reaches it
constexpr int INK_SPIN = 1 << 16;
// Position of the band in the row order, as a fraction of the rows. Timed on
// the 2048 x 1024 cylinder (fused step, one body) from 0.15 to 0.9 of the
// rows: 0.15 is fastest (26.3-26.8 us per step, 27.5 us at 0.9) - the
// blocks that wait do so while the rest of the lattice still has work. With
// the network (4096 x 2048) the step is flat from 0.3 on and 7% slower with
// the band first.
constexpr double INK_BAND_AT = 0.15, INK_BAND_AT_NN = 0.3;

template <typename T>
__device__ void ib_in_step(const StepArgs& a, int body, const T* __restrict__ src, T* U) {
#pragma clang fp contract(off)
    const IbBlk bk = a.ib_lead[body];
    const int t = threadIdx.x;
    const unsigned N = (unsigned)a.nx * a.ny;
    // A. velocity of every active node of the body (k_ib, pulling), NA nodes
    // in flight per lane (one in FP64, where the lattice path's registers
    // already set the kernel's budget)
    constexpr int NA = sizeof(T) == 8 ? 1 : 3;
    for (int c0 = 0; c0 < bk.nact; c0 += NA * BX) {
        int nd[NA];
#pragma unroll
        for (int j = 0; j < NA; j++) nd[j] = a.ib_act[bk.act0 + min(c0 + j * BX + t, bk.nact - 1)];
        T ux[NA], uy[NA];
#pragma unroll
        for (int j = 0; j < NA; j++) {
            const int r = nd[j] - bk.pad;
            double vx, vy;
            support_u<T, true>(src, nullptr, bk.x0 + r % bk.w, bk.y0 + r / bk.w, a.nx, N, vx, vy, a.ref);
            ux[j] = (T)vx;  // exact: the FP32 path's velocity deviations are FP32
            uy[j] = (T)vy;
        }
#pragma unroll
        for (int j = 0; j < NA; j++) {
            const int i = c0 + j * BX + t;
            if (i < bk.nact) {
                U[2 * i] = ux[j];
                U[2 * i + 1] = uy[j];
            }
        }
    }
    __syncthreads();
    // B. interpolation and direct forcing, markers t, t + BX, t + 2 BX, t + 3 BX
    double F[4][2];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int ml = t + r * BX;
        F[r][0] = 0.0;
        F[r][1] = 0.0;
        if (ml < bk.M) {
            const int m = bk.m0 + ml;
            double uix = 0, uiy = 0;
#pragma unroll
            for (int s4 = 0; s4 < 16; s4 += 4) {
                int pi[4];
                double pw[4];
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    pi[s] = a.ib_pidx[16 * m + s4 + s];
                    pw[s] = a.ib_w[16 * m + s4 + s];
                }
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    uix += ((double)U[2 * pi[s]] * pw[s]);
                    uiy += ((double)U[2 * pi[s] + 1] * pw[s]);
                }
            }
            if constexpr (sizeof(T) == 4) uix = uix + a.ref.U;  // FP32: deviations were interpolated
            const double Fx = 2 * (a.ib_ud[2 * m] - uix), Fy = 2 * (a.ib_ud[2 * m + 1] - uiy);
            a.ib_ui[2 * m] = uix;
            a.ib_ui[2 * m + 1] = uiy;
            double* Fp = const_cast<double*>(a.F) + 2 * m;
            __hip_atomic_store(Fp, Fx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(Fp + 1, Fy, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            F[r][0] = Fx;
            F[r][1] = Fy;
        }
    }
    // every force of the block is stored before the flag. Ordering: forces
    // and flag are agent-scope atomics, each coherent across the XCDs' L2s;
    // s_waitcnt(0) holds the flag store until every force store of the block
    // is acknowledged, and a reader issues its force loads (agent-scope
    // atomics too, patch_force<true>) only after a poll has returned the
    // flag. A release store and acquire loads would order the same but add
    // an L2 writeback and an L2 invalidate at agent scope, which the lattice
    // blocks' cached data would pay for.
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (t == 0) __hip_atomic_store(a.ib_flag + body, a.ib_seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // record: the 512-entry tree of k_ib (entry e = F[e] + F[e + 512]), whose
    // entries t and t + 256 this thread holds
    double v[2][2];
#pragma unroll
    for (int c = 0; c < 2; c++) {
        double e0 = 0.0, e1 = 0.0;
        if (t < bk.M) e0 += F[0][c];
        if (t + 2 * BX < bk.M) e0 += F[2][c];
        if (t + BX < bk.M) e1 += F[1][c];
        if (t + 3 * BX < bk.M) e1 += F[3][c];
        v[0][c] = e0;
        v[1][c] = e1;
    }
    double* R = reinterpret_cast<double*>(U);  // 2 x 256 doubles, after the barrier above
    R[2 * t] = v[0][0] + v[1][0];  // level 256
    R[2 * t + 1] = v[0][1] + v[1][1];
    __syncthreads();
    if (t < 128) {
        R[2 * t] += R[2 * (t + 128)];
        R[2 * t + 1] += R[2 * (t + 128) + 1];
    }
    __syncthreads();
    if (t < 64) {
        double sx = R[2 * t] + R[2 * (t + 64)], sy = R[2 * t + 1] + R[2 * (t + 64) + 1];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double ox = __shfl_down(sx, off, 64), oy = __shfl_down(sy, off, 64);
            if (t < off) {
                sx += ox;
                sy += oy;
            }
        }
        if (t == 0) {
            a.ib_hist[2 * bk.body] = -sx;
            a.ib_hist[2 * bk.body + 1] = -sy;
        }
    }
}

// Wait (every wave of a band block) until every body's forces of this step
// are stored: agent-coherent polls of the flags, bounded by INK_SPIN.
__device__ __forceinline__ void wait_ib_forces(const StepArgs& a) {
    for (int b = 0; b < a.nbody; b++) {
        int spins = 0;
        while (__hip_atomic_load(a.ib_flag + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != a.ib_wait) {
            if (++spins > INK_SPIN) {
                *a.flag = 2;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
    }
}

// NPX nodes per thread, BX columns apart: with the exact-weight network
// (issue-bound, four waves per SIMD) two nodes' loads go out together and
// their two networks run back to back on the data (2% faster at 4096 x
// 2048; the compensated network spills and is 37% slower, so it keeps one).
// Grid: ceil(nx / (NPX BX)) columns (the lattice's own width, so that the
// block order's XCD grouping holds); the body blocks fill the first
// a.brows rows, the lattice rows follow.
template <int COLL, int BC, int NN, bool NT, int NPX>
__global__ void __launch_bounds__(BX) __attribute__((amdgpu_waves_per_eu(NN ? IPPI_NN_WAVES : 8)))
k_step_f32_ib(const float* __restrict__ src, float* __restrict__ dst, const double* __restrict__ mac, StepArgs a) {
    // the IB block's velocities; the lattice blocks of the exact-weight
    // network keep the network's fragments in the same LDS
    __shared__ __attribute__((aligned(16))) float U[2 * (NN ? INK_MAXACT_NN : INK_MAXACT)];
    static_assert(NN != 3 || sizeof(U) >= 16 * MLP_FRAGE_CHUNKS * 64, "LDS for the network's fragments");
    __shared__ __attribute__((aligned(16))) __bf16 lds_mlp[NN ? BX * 16 : 1];
    const int nx = a.nx, ny = a.ny;
    if ((int)blockIdx.y < a.brows) {  // the body blocks' rows (block-uniform)
        const int j = blockIdx.y * gridDim.x + blockIdx.x;
        if (j < a.ink_bodies) ib_in_step(a, j, src, U);
        return;
    }
    int bxi = blockIdx.x, g = (int)blockIdx.y - a.brows;
    if (a.xcd) xcd_block(g * (int)gridDim.x + bxi, gridDim.x, ny, bxi, g);
    const int xs = bxi * BX * NPX;  // first column of this block
    int y = g + a.rot;  // rot < ny
    if (y >= ny) y -= ny;
    const unsigned N = (unsigned)nx * ny;
    u32x4* lds_frag = reinterpret_cast<u32x4*>(U);
    if (NN == 3) mlp_frage_to_lds<BX>(a.mlp, lds_frag);  // in flight with the population loads
    float cur[NPX][9];
#pragma unroll
    for (int p = 0; p < NPX; p++) {
        const int x = min(xs + p * BX + (int)threadIdx.x, nx - 1);
        load_raw<float, BC, true>(src, x, y, nx, ny, N, cur[p], a.ref.bb);
    }
    // the band waits with its populations in flight (the body blocks only
    // read src, so they are the same either way)
    if (y >= a.band0 && y < a.band1) wait_ib_forces(a);  // block-uniform
    if (NN == 3) __syncthreads();
    __bf16* lds = lds_mlp + (NN ? (threadIdx.x & ~63) * 16 : 0);
#pragma unroll
    for (int p = 0; p < NPX; p++) {
        const int x0 = xs + p * BX;
        const bool active = x0 + (int)threadIdx.x < nx;
        const int x = min(x0 + (int)threadIdx.x, nx - 1);
        node_f32<COLL, BC, true, NN, NT, true>(cur[p], x, y, N, src, mac, dst, a, active, lds, x0, lds_frag);
    }
}

// FP64 variant: NPT rows per thread as k_step_f64, the rotation and the
// band in units of those row groups.
template <int COLL, int BC, bool NT, int NPT>
__global__ void __launch_bounds__(BX) k_step_f64_ib(const double* __restrict__ src, double* __restrict__ dst, const double* __restrict__ mac, StepArgs a) {
    __shared__ double U[2 * INK_MAXACT64];
    const int nx = a.nx, ny = a.ny;
    if ((int)blockIdx.y < a.brows) {  // the body blocks' rows (block-uniform)
        const int j = blockIdx.y * gridDim.x + blockIdx.x;
        if (j < a.ink_bodies) ib_in_step(a, j, src, U);
        return;
    }
    const int G = (ny + NPT - 1) / NPT;
    int bxi = blockIdx.x, g = (int)blockIdx.y - a.brows;
    if (a.xcd) xcd_block(g * (int)gridDim.x + bxi, gridDim.x, G, bxi, g);
    int gi = g + a.rot;  // rot < G
    if (gi >= G) gi -= G;
    const int y0 = gi * NPT;
    const bool active = (int)(bxi * BX + threadIdx.x) < nx;
    const int x = min((int)(bxi * BX + threadIdx.x), nx - 1);
    const unsigned N = (unsigned)nx * ny;
    double cur[9];
    load_raw<double, BC, true>(src, x, y0, nx, ny, N, cur);
    if (y0 < a.band1 && y0 + NPT > a.band0) wait_ib_forces(a);  // block-uniform, first row's loads in flight
#pragma unroll
    for (int j = 0; j < NPT; j++) {
        const int y = y0 + j;
        if (y >= ny) break;
        const bool more = j + 1 < NPT && y + 1 < ny;
        double nxt[9];
        if (j + 1 < NPT) load_raw<double, BC, true>(src, x, min(y + 1, ny - 1), nx, ny, N, nxt);
        node_f64<COLL, BC, true, false, NT, true>(cur, x, y, N, src, mac, dst, a, active, bxi * BX);
        if (more) {
#pragma unroll
            for (int k = 0; k < 9; k++) cur[k] = nxt[k];
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

// ----------------------------------------------------------------- adjoint
// Discrete adjoint of the BGK + IB step with periodic x and bounce-back
// walls (node-local algebra in adjoint.hpp). The adjoint state is the
// adjoint of the stored post-collision populations; one reverse step runs
//   k_adj_act: force adjoint at every active (marker-support) node,
//   k_adj_mk:  marker transpose per body (force -> interpolation), centre
//              gradient, gather of the velocity adjoint back to the nodes,
//   k_adj_step: all nodes: reverse pull, collision / macro transpose, tau
//              gradient accumulation.
// A spatially constant adjoint (the same value for every population at
// every node) is mapped to itself by the transposed step (the adjoint of
// mass conservation) and contributes to no gradient. Every reverse step
// subtracts the adjoint of population 0 at node 0 from all populations as
// they are loaded, so no common offset builds up in the stored adjoint;
// the gradients are unchanged.

// Transpose of pull<T, IPPI_BC_PERIODIC_BB>: the adjoint of population k at
// node m is population k of the downstream node m + e_k, or, where m + e_k
// lies beyond a wall, the reversed population at m itself.
template <typename T>
__device__ __forceinline__ void pull_adj(const T* __restrict__ s, int x, int y, int nx, int ny, unsigned N, T* l) {
    const bool bot = y == 0, top = y == ny - 1;
    const unsigned xc[3] = {(unsigned)(x == 0 ? nx - 1 : x - 1), (unsigned)x, (unsigned)(x == nx - 1 ? 0 : x + 1)};
#pragma unroll
    for (int k = 0; k < 9; k++) {
        int pk = k, dy = EY[k], dx = EX[k];
        if ((EY[k] == -1 && bot) || (EY[k] == 1 && top)) {
            pk = OPP[k];
            dy = 0;
            dx = 0;
        }
        l[k] = ld(s + (unsigned)pk * N, (unsigned)(y + dy) * nx + xc[dx + 1]);
    }
}

struct AdjArgs {
    int nx, ny;
    RefState ref;           // FP32 storage offsets
    double tau0;
    int tau_field;
    const void* tau_ptr;
    const double* patch_f;  // IB force of the step [patch node][2]
    const double* patch_l;  // velocity adjoint from the interpolation [patch node][2]
    int nbody;
    int store;              // write the adjoint populations
    int xcd;                // XCD-grouped block order (xcd_block)
    // k_adj_fused: the bodies' blocks publish the velocity adjoint with a
    // per-body flag; lattice rows in [band0, band1) wait for the flags of
    // the bodies rowmask names
    const unsigned* rowmask;  // [ny] bit b: body b's patch covers the row
    unsigned* flags;          // [body] sequence number of the last published step
    unsigned* cnt;            // [body] blocks of the body done with the force adjoint this step
    unsigned seq;             // this reverse step's sequence number
    int band0, band1, rot;    // band rows; lattice row of the first lattice grid row (rows follow cyclically)
    int brows;                // grid rows of the bodies' blocks before the lattice rows
    int* err;                 // set to 2 when a wait times out
    Patch patch[IPPI_MAXB];
};

struct AdjBody {
    int m0, M, a0, nact;
    double wx, wy;  // objective weights of the body's force record
};

struct AdjIb {
    int nact, E;
    const int* act;       // active node -> patch node
    const int* act_node;  // active node -> y * nx + x
    double* lf;           // [active][2] force adjoint
    double* u;            // [active][2] velocity seen by the interpolation
    const int* pidx;      // [m][16] support -> active index within the body
    const double* w;      // [m][16] phi phi
    const double* wx;     // [m][16] phi'(x - X) phi(y - Y)
    const double* wy;     // [m][16] phi(x - X) phi'(y - Y)
    const double* F;      // [m][2] marker forces of the step
    const int* ecnt;
    const int* emk;
    const double* ecw;
    const AdjBody* body;
    double* patch_l;      // [patch node][2]
    double* gc;           // [body][2] centre gradient (accumulated)
    // k_adj_fused: the markers' velocity-adjoint weights (published for the
    // lattice rows' gather), the pairs' centre-gradient terms, patch node ->
    // active node (-1: none)
    double* lu;           // [marker][2]
    double* pc;           // [m][16][2]
    const int* pnr;       // [m][16][2] the pair's support node (y * nx + x) and patch node
    const int* pslot;
};

// agent-coherent load of a value another block of the same kernel stored
// with an agent-scope atomic (the XCDs' L2s are not coherent with each other)
__device__ __forceinline__ double ld_coh(const double* p) {
    return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_coh(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int adj_patch(const AdjArgs& a, int x, int y, int bx0, int bx1, int* body = nullptr) {
    for (int b = 0; b < a.nbody; ++b) {
        const Patch p = a.patch[b];
        if (bx1 < p.x0 || bx0 >= p.x0 + p.w || y < p.y0 || y >= p.y0 + p.h) continue;
        if (x >= p.x0 && x < p.x0 + p.w) {
            if (body) *body = b;
            return p.off + (y - p.y0) * p.w + (x - p.x0);
        }
    }
    return -1;
}

// Equilibrium at the corrected velocity (usx, usy) in the representation of
// the stored populations: full feq (FP64), or its deviation from c (FP32).
template <typename T, typename A>
__device__ __forceinline__ void adj_eq(T drho, T rho, A usx, A usy, const RefState& rs, T* q) {
    if constexpr (sizeof(T) == 8)
        ref::feq(rho, (double)usx, (double)usy, q);
    else if constexpr (sizeof(A) == 8)
        sh::geq_dev(drho, rho, (float)(usx - rs.U), (float)usy, rs.Uf, rs.cf, q);
    else
        sh::geq_dev(drho, rho, usx - rs.Uf, usy, rs.Uf, rs.cf, q);
}

// Forward state of the step at node (x, y): pulled populations (or the
// stored first state with the macro arrays), density and velocity.
template <typename T, bool PULL>
__device__ __forceinline__ void adj_fwd(const T* __restrict__ src, const double* __restrict__ mac, int x, int y,
                                        int nx, int ny, unsigned N, const RefState& rs, T* f, T& drho, T& rho, T& ux,
                                        T& uy) {
    load_raw<T, IPPI_BC_PERIODIC_BB, PULL>(src, x, y, nx, ny, N, f, rs.bb);
    const unsigned n = (unsigned)y * nx + x;
    if (PULL) {
        T s = (f[0] + (f[1] + f[3])) + ((f[2] + f[4]) + ((f[5] + f[7]) + (f[6] + f[8])));
        T jx = ((f[1] - f[3]) + (f[5] - f[7])) + (f[8] - f[6]);
        T jy = ((f[2] - f[4]) + (f[5] - f[7])) + (f[6] - f[8]);
        if constexpr (sizeof(T) == 8) {
            rho = s;
            drho = s - 1.0;
        } else {
            drho = s;
            rho = 1.0f + s;
        }
        if (rho < T(1e-6)) {
            ux = T(0);
            uy = T(0);
        } else if constexpr (sizeof(T) == 8) {
            ux = jx / rho;
            uy = jy / rho;
        } else {  // deviations of the momentum from (U, 0)
            ux = rs.Uf + (jx - rs.Uf * drho) / rho;
            uy = jy / rho;
        }
    } else {
        rho = (T)mac[n];
        drho = (T)(mac[n] - 1.0);
        ux = (T)mac[N + n];
        uy = (T)mac[2 * N + n];
    }
}

// Base tau - 1/2 of node n in the adjoint's precision: the FP64 state
// keeps tau, the FP32 state tau - 1/2.
template <typename S, typename A>
__device__ __forceinline__ A adj_tx(const AdjArgs& a, unsigned n) {
    if (!a.tau_field) return (A)(a.tau0 - 0.5);
    if constexpr (sizeof(S) == 8)
        return (A)(ld(static_cast<const double*>(a.tau_ptr), n) - 0.5);
    else
        return (A)ld(static_cast<const float*>(a.tau_ptr), n);
}
// S: type of the stored forward populations, A: adjoint arithmetic and
// storage.
// Force adjoint of active node i (its velocity kept beside it): FP64
// throughout (a few thousand nodes): the centre gradient sums the force
// adjoint and the velocity against the delta's derivative, whose weights
// sum to zero, so the rounding of FP32 arithmetic on these quantities would
// not cancel with the rest.
template <typename S, typename A, bool PULL>
__device__ __forceinline__ void adj_act_vals_at(int node, int r, const A* __restrict__ lin, const S* __restrict__ src,
                                                const double* __restrict__ mac, const AdjArgs& a, double& lfx,
                                                double& lfy, double& ux, double& uy) {
    const int nx = a.nx, ny = a.ny;
    const unsigned N = (unsigned)nx * ny;
    const int x = node % nx, y = node / nx;
    A la[9];
    pull_adj(lin, x, y, nx, ny, N, la);
    const A c0 = lin[0];
    double l[9];
#pragma unroll
    for (int k = 0; k < 9; k++) l[k] = (double)(la[k] - c0);
    double rho;
    if (PULL) {
        S fs[9];
        load_raw<S, IPPI_BC_PERIODIC_BB, true>(src, x, y, nx, ny, N, fs, a.ref.bb);
        double fd[9];
#pragma unroll
        for (int k = 0; k < 9; k++) fd[k] = sizeof(S) == 8 ? (double)fs[k] : (double)fs[k] + a.ref.c[k];
        ref::macro(fd, rho, ux, uy);
    } else {
        rho = mac[node];
        ux = mac[N + node];
        uy = mac[2 * N + node];
    }
    const adj::NodeFwd<double> s = adj::node_fwd<double>(rho, ux, uy, a.patch_f[2 * r], a.patch_f[2 * r + 1],
                                                         adj_tx<S, double>(a, (unsigned)node));
    const adj::WMom<double> m = adj::wmoments(l);
    const adj::UsAdj<double> ua = adj::us_adjoint(m, s, true);
    adj::force_adjoint(ua, s, lfx, lfy);
}
template <typename S, typename A, bool PULL>
__device__ __forceinline__ void adj_act_vals(int i, const A* __restrict__ lin, const S* __restrict__ src,
                                             const double* __restrict__ mac, const AdjIb& d, const AdjArgs& a,
                                             double& lfx, double& lfy, double& ux, double& uy) {
    adj_act_vals_at<S, A, PULL>(d.act_node[i], d.act[i], lin, src, mac, a, lfx, lfy, ux, uy);
}
template <typename S, typename A, bool PULL>
__device__ __forceinline__ void adj_act_node(int i, const A* __restrict__ lin, const S* __restrict__ src,
                                             const double* __restrict__ mac, const AdjIb& d, const AdjArgs& a) {
    double lfx, lfy, ux, uy;
    adj_act_vals<S, A, PULL>(i, lin, src, mac, d, a, lfx, lfy, ux, uy);
    d.lf[2 * i] = lfx;
    d.lf[2 * i + 1] = lfy;
    d.u[2 * i] = ux;
    d.u[2 * i + 1] = uy;
}

template <typename S, typename A, bool PULL>
__global__ void __launch_bounds__(256) k_adj_act(const A* __restrict__ lin, const S* __restrict__ src,
                                                 const double* __restrict__ mac, AdjIb d, AdjArgs a) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < d.nact) adj_act_node<S, A, PULL>(i, lin, src, mac, d, a);
}

// One block per body. cj = weight of this step's force record in the
// objective. One lane per (marker, support node) pair, in chunks of IB_T
// pairs (32 markers) held in flight together; the sum over a marker's 16
// supports is a shuffle tree inside its 16-lane group, the centre gradient
// a shuffle tree per wave and the wave sums in wave order, so the result is
// deterministic. The gather lists of each lane's first support node are
// loaded at the start (all E entries: padding entries have weight 0 and
// add nothing), so the gather at the end reads only LDS.
constexpr int ADJ_CH = 4;  // chunks in flight per lane (IB_T * ADJ_CH / 16 = 128 markers)
__global__ void __launch_bounds__(IB_T) k_adj_mk(AdjIb d, double cj) {
    __shared__ double LU[2 * IB_MAXM];
    __shared__ double red[2 * IB_T];
    const AdjBody b = d.body[blockIdx.x];
    const int t = threadIdx.x;
    // gather list of this lane's first support node (E = 8 lists only)
    int gmk[8];
    double gw[8];
    const bool pre = d.E == 8 && t < b.nact;
    if (pre) {
        const int i = b.a0 + t;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            gmk[e] = d.emk[8 * (size_t)i + e];
            gw[e] = d.ecw[8 * (size_t)i + e];
        }
    }
    double gx = 0.0, gy = 0.0;
    for (int c0 = 0; c0 < 16 * b.M; c0 += ADJ_CH * IB_T) {
        double w[ADJ_CH], wxs[ADJ_CH], wys[ADJ_CH], lfx[ADJ_CH], lfy[ADJ_CH], ux[ADJ_CH], uy[ADJ_CH];
        // all loads of the lane's pairs first (pairs past the body's end read pair 0, weight 0)
#pragma unroll
        for (int j = 0; j < ADJ_CH; j++) {
            const int item = c0 + j * IB_T + t;
            const bool ok = item < 16 * b.M;
            const int q = 16 * b.m0 + (ok ? item : 0);
            const int i = b.a0 + d.pidx[q];
            w[j] = ok ? d.w[q] : 0.0;
            wxs[j] = ok ? d.wx[q] : 0.0;
            wys[j] = ok ? d.wy[q] : 0.0;
            lfx[j] = d.lf[2 * i];
            lfy[j] = d.lf[2 * i + 1];
            ux[j] = d.u[2 * i];
            uy[j] = d.u[2 * i + 1];
        }
#pragma unroll
        for (int j = 0; j < ADJ_CH; j++) {
            const int item = c0 + j * IB_T + t;
            const int ml = min(item, 16 * b.M - 1) >> 4;  // marker of the pair (16-lane aligned)
            double ax = w[j] * lfx[j], ay = w[j] * lfy[j];
#pragma unroll
            for (int o = 8; o > 0; o >>= 1) {
                ax += __shfl_xor(ax, o, 16);
                ay += __shfl_xor(ay, o, 16);
            }
            // record = - sum of marker forces; F = 2 (U_d - U)
            const double lux = -2.0 * (ax - cj * b.wx), luy = -2.0 * (ay - cj * b.wy);
            if (item < 16 * b.M && (item & 15) == 0) {
                LU[2 * ml] = lux;
                LU[2 * ml + 1] = luy;
            }
            const int m = b.m0 + ml;
            const double Fx = d.F[2 * m], Fy = d.F[2 * m + 1];
            const double cc = lux * ux[j] + luy * uy[j] + lfx[j] * Fx + lfy[j] * Fy;
            gx += wxs[j] * cc;
            gy += wys[j] * cc;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        gx += __shfl_down(gx, off, 64);
        gy += __shfl_down(gy, off, 64);
    }
    if ((t & 63) == 0) {
        red[2 * (t >> 6)] = gx;
        red[2 * (t >> 6) + 1] = gy;
    }
    __syncthreads();  // also orders the LU stores before the gather
    if (t == 0) {
        double sx = 0.0, sy = 0.0;
        for (int w = 0; w < IB_T / 64; w++) {
            sx += red[2 * w];
            sy += red[2 * w + 1];
        }
        d.gc[2 * blockIdx.x] += sx;
        d.gc[2 * blockIdx.x + 1] += sy;
    }
    // velocity adjoint at the support nodes: the spread's gather lists
    if (pre) {
        double sx = 0.0, sy = 0.0;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            sx += gw[e] * LU[2 * gmk[e]];
            sy += gw[e] * LU[2 * gmk[e] + 1];
        }
        const int r = d.act[b.a0 + t];
        d.patch_l[2 * r] = sx;
        d.patch_l[2 * r + 1] = sy;
    }
    for (int il = pre ? t + IB_T : t; il < b.nact; il += IB_T) {
        const int i = b.a0 + il;
        const int cnt = d.ecnt[i];
        double sx = 0.0, sy = 0.0;
        for (int e = 0; e < cnt; e++) {
            const int mk = d.emk[(size_t)d.E * i + e];
            const double w = d.ecw[(size_t)d.E * i + e];
            sx += w * LU[2 * mk];
            sy += w * LU[2 * mk + 1];
        }
        const int r = d.act[i];
        d.patch_l[2 * r] = sx;
        d.patch_l[2 * r + 1] = sy;
    }
}

// Reverse step at node (x, y) of all-node block column bxi. COH: the
// velocity adjoint of the patch nodes was published by another block of the
// same kernel (k_adj_fused) and is read with agent-coherent loads.
template <typename S, typename A, bool PULL, bool COH>
__device__ __forceinline__ void adj_node(const A* __restrict__ lin, A* __restrict__ lout, const S* __restrict__ src,
                                         const double* __restrict__ mac, A* __restrict__ gacc, const AdjArgs& a,
                                         int x, int y, int bxi, const AdjIb* dg = nullptr) {
    const int nx = a.nx, ny = a.ny;
    const unsigned N = (unsigned)nx * ny, n = (unsigned)y * nx + x;
    // arithmetic in FP64 whatever the adjoint's storage: with FP32 arithmetic
    // the rounding of each reverse step accumulates into the centre gradient,
    // a sum that cancels by about 1e4 (C4: 1.7e-4 / 3.4e-4 off the FP64
    // adjoint, against 6e-6 / 7e-6 with FP64 arithmetic over the same FP32
    // storage; the kernel stays bound by its memory traffic)
    typedef double C;
    A la[9];
    S f[9], drho, rho, ux, uy;
    pull_adj(lin, x, y, nx, ny, N, la);
    const A c0 = lin[0];  // constant-mode removal
    C l[9];
#pragma unroll
    for (int k = 0; k < 9; k++) l[k] = (C)la[k] - (C)c0;
    adj_fwd<S, PULL>(src, mac, x, y, nx, ny, N, a.ref, f, drho, rho, ux, uy);
    C fx = C(0), fy = C(0), lix = C(0), liy = C(0);
    const int bx0 = bxi * BX;
    int body = 0;
    const int r = a.nbody > 0 ? adj_patch(a, x, y, bx0, bx0 + BX - 1, &body) : -1;
    const bool forced = r >= 0;
    if (forced) {
        fx = (C)a.patch_f[2 * r];
        fy = (C)a.patch_f[2 * r + 1];
        if (COH) {  // k_adj_mk's gather, from the weights the body's blocks published
            const int ia = dg->pslot[r];
            if (ia >= 0) {
                const int cnt = dg->ecnt[ia], m0 = a.patch[body].m0;
                double sx = 0.0, sy = 0.0;
                for (int e = 0; e < cnt; e++) {
                    const int mk = m0 + dg->emk[(size_t)dg->E * ia + e];
                    const double wg = dg->ecw[(size_t)dg->E * ia + e];
                    sx += wg * ld_coh(dg->lu + 2 * mk);
                    sy += wg * ld_coh(dg->lu + 2 * mk + 1);
                }
                lix = (C)sx;
                liy = (C)sy;
            }
        } else {
            lix = (C)a.patch_l[2 * r];
            liy = (C)a.patch_l[2 * r + 1];
        }
    }
    const adj::NodeFwd<C> s = adj::node_fwd<C>((C)rho, (C)ux, (C)uy, fx, fy, adj_tx<S, C>(a, n));
    const adj::WMom<C> m = adj::wmoments(l);
    const adj::UsAdj<C> ua = adj::us_adjoint(m, s, forced);
    // non-equilibrium part in the representation of the stored populations
    S q[9];
    adj_eq(drho, rho, ua.usx, ua.usy, a.ref, q);
    C dd[9], nu[9], dt;
#pragma unroll
    for (int k = 0; k < 9; k++) dd[k] = (C)(f[k] - q[k]);
    adj::node_adjoint(l, dd, m, ua, s, forced, lix, liy, nu, dt);
    if (a.store) {
#pragma unroll
        for (int k = 0; k < 9; k++) lout[k * N + n] = (A)nu[k];
    }
    gacc[n] += (A)dt;
}

template <typename S, typename A, bool PULL>
__global__ void __launch_bounds__(BX) k_adj_step(const A* __restrict__ lin, A* __restrict__ lout,
                                                 const S* __restrict__ src, const double* __restrict__ mac,
                                                 A* __restrict__ gacc, AdjArgs a) {
    int bxi = blockIdx.x, y = blockIdx.y;
    if (a.xcd) xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x, gridDim.y, bxi, y);
    const int x = bxi * BX + threadIdx.x;
    if (x >= a.nx) return;
    adj_node<S, A, PULL, false>(lin, lout, src, mac, gacc, a, x, y, bxi);
}

// The bodies' part of k_adj_fused. A block takes BX / 16 markers of one
// body, a thread one (marker, support) pair in k_adj_mk's numbering: it
// computes the force adjoint and velocity at its support node itself
// (k_adj_act's arithmetic; nodes shared by several markers are computed by
// each), the marker's velocity-adjoint weight from the 16-lane sum of
// k_adj_mk, and its pair's centre-gradient terms. Weights and terms are
// stored with agent-scope atomics; after the block's stores are
// acknowledged it counts itself done, and the body's last block sets the
// body's flag for the lattice rows, then sums the centre gradient over the
// pairs in k_adj_mk's lane, wave and block order (off the lattice rows'
// path), so the results are bit for bit k_adj_act's and k_adj_mk's.
constexpr int ADJ_MPB = BX / 16;  // markers per block
static_assert(IB_T % BX == 0, "k_adj_mk lanes per thread");
template <typename S, typename A, bool PULL>
__device__ __forceinline__ void adj_body_block(const A* __restrict__ lin, const S* __restrict__ src,
                                               const double* __restrict__ mac, const AdjIb& d, const AdjArgs& a,
                                               double cj, double* red) {
    int body = -1, mb0 = 0, nblk = 0, j = blockIdx.y * gridDim.x + blockIdx.x;
    for (int b = 0; b < a.nbody; b++) {
        const int n = (d.body[b].M + ADJ_MPB - 1) / ADJ_MPB;
        if (j < n) {
            body = b;
            mb0 = j * ADJ_MPB;
            nblk = n;
            break;
        }
        j -= n;
    }
    if (body < 0) return;  // block-uniform
    const AdjBody b = d.body[body];
    const int t = threadIdx.x;
    const int item = 16 * mb0 + t;  // the pair in k_adj_mk's numbering within the body
    const bool ok = item < 16 * b.M;
    const int qq = 16 * b.m0 + (ok ? item : 0);
    // the pair's node and patch node from the table (no dependent index load), its weights and
    // the marker's force requested with them
    const int ml = min(item, 16 * b.M - 1) >> 4;
    const int m = b.m0 + ml;
    const int pnode = d.pnr[2 * qq], pr = d.pnr[2 * qq + 1];
    const double w = ok ? d.w[qq] : 0.0, wxs = ok ? d.wx[qq] : 0.0, wys = ok ? d.wy[qq] : 0.0;
    const double Fx = d.F[2 * m], Fy = d.F[2 * m + 1];
    double lfx, lfy, ux, uy;
    adj_act_vals_at<S, A, PULL>(pnode, pr, lin, src, mac, a, lfx, lfy, ux, uy);
    double ax = w * lfx, ay = w * lfy;
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) {
        ax += __shfl_xor(ax, o, 16);
        ay += __shfl_xor(ay, o, 16);
    }
    // record = - sum of marker forces; F = 2 (U_d - U)
    const double lux = -2.0 * (ax - cj * b.wx), luy = -2.0 * (ay - cj * b.wy);
    if (ok && (t & 15) == 0) {
        st_coh(d.lu + 2 * m, lux);
        st_coh(d.lu + 2 * m + 1, luy);
    }
    const double cc = lux * ux + luy * uy + lfx * Fx + lfy * Fy;
    if (ok) {
        st_coh(d.pc + 2 * (size_t)qq, wxs * cc);
        st_coh(d.pc + 2 * (size_t)qq + 1, wys * cc);
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    __shared__ int last;
    if (t == 0)
        last = __hip_atomic_fetch_add(a.cnt + body, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
               (unsigned)(nblk - 1);
    __syncthreads();
    if (!last) return;  // block-uniform
    if (t == 0) {
        __hip_atomic_store(a.cnt + body, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(a.flags + body, a.seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    // centre gradient: lane L of k_adj_mk (thread t, L = q BX + t) sums the
    // terms of its pairs L, L + IB_T, ... in order; then k_adj_mk's trees
    constexpr int NV = IB_T / BX;
    double gx[NV], gy[NV];
#pragma unroll
    for (int q = 0; q < NV; q++) {
        gx[q] = 0.0;
        gy[q] = 0.0;
        for (int it = q * BX + t; it < 16 * b.M; it += IB_T) {
            gx[q] += ld_coh(d.pc + 2 * (size_t)(16 * b.m0 + it));
            gy[q] += ld_coh(d.pc + 2 * (size_t)(16 * b.m0 + it) + 1);
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            gx[q] += __shfl_down(gx[q], off, 64);
            gy[q] += __shfl_down(gy[q], off, 64);
        }
        if ((t & 63) == 0) {  // k_adj_mk's wave (q BX + t) / 64
            red[2 * ((q * BX + t) >> 6)] = gx[q];
            red[2 * ((q * BX + t) >> 6) + 1] = gy[q];
        }
    }
    __syncthreads();
    if (t == 0) {
        double sx = 0.0, sy = 0.0;
        for (int wv = 0; wv < IB_T / 64; wv++) {
            sx += red[2 * wv];
            sy += red[2 * wv + 1];
        }
        d.gc[2 * body] += sx;
        d.gc[2 * body + 1] += sy;
    }
}

// Position of the patch rows in the adjoint's row order, as a fraction of
// the rows. Timed on the 1024 x 512 cylinder (D = 40): FP64 25.9-26.0 us per
// reverse step with the patch rows last, 25.0-25.2 us at 0.7-0.8; FP32
// 18.6 us last, 18.4-18.5 us at 0.7-0.8 (the marker blocks' flags are set
// by then, and the rows after the band overlap the patch rows' gathers).
constexpr double ADJ_BAND_AT = 0.75;

// One reverse step in one launch: the first grid rows hold the bodies'
// blocks (adj_body_block), the lattice rows follow in rotated order with the
// bodies' patch rows at ADJ_BAND_AT of the order; a patch row waits for the flags of the bodies
// covering it (bounded: a time-out sets *err = 2 and the sweep reports the
// failure, never hangs) and its forced nodes gather their velocity adjoint
// from the published marker weights. Ordering as in the forward step's
// in-step IB: weights, terms and flag are agent-scope atomics, s_waitcnt(0)
// holds a block's count until every store of the block is acknowledged, and
// readers load the weights only after a poll returned the flag. Bit for bit
// the results of k_adj_act, k_adj_mk and k_adj_step in sequence.
template <typename S, typename A, bool PULL>
__global__ void __launch_bounds__(BX) k_adj_fused(const A* __restrict__ lin, A* __restrict__ lout,
                                                  const S* __restrict__ src, const double* __restrict__ mac,
                                                  A* __restrict__ gacc, AdjArgs a, AdjIb d, double cj) {
    __shared__ double red[2 * (IB_T / 64)];
    const int nx = a.nx, ny = a.ny;
    if ((int)blockIdx.y < a.brows) {  // the bodies' rows of blocks (block-uniform)
        adj_body_block<S, A, PULL>(lin, src, mac, d, a, cj, red);
        return;
    }
    // the lattice's own grid width, so that each row of blocks stays on one
    // XCD in the grouped order (the body rows shift every block's XCD alike)
    const int gx = (nx + BX - 1) / BX;
    int bxi = blockIdx.x, g = (int)blockIdx.y - a.brows;
    if (a.xcd) xcd_block(g * gx + bxi, gx, ny, bxi, g);
    int y = g + a.rot;
    if (y >= ny) y -= ny;
    if (y >= a.band0 && y < a.band1) {  // block-uniform: wait for the bodies covering the row
        unsigned m = __builtin_amdgcn_readfirstlane(a.rowmask[y]);
        while (m) {
            const int b = __builtin_ctz(m);
            m &= m - 1u;
            int spins = 0;
            while (__hip_atomic_load(a.flags + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != a.seq) {
                if (++spins > INK_SPIN) {
                    *a.err = 2;
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
    }
    const int x = bxi * BX + threadIdx.x;
    if (x >= nx) return;
    adj_node<S, A, PULL, true>(lin, lout, src, mac, gacc, a, x, y, bxi, &d);
}

// total += part; part = 0 (per-segment tau-gradient partial sums).
template <typename T>
__global__ void k_adj_flush(T* __restrict__ part, double* __restrict__ total, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    total[i] += (double)part[i];
    part[i] = T(0);
}

}  // namespace

// ====================================================================== host

struct ippi_handle {
    ippi_config c{};
    RefState ref{};  // FP32 storage offsets: s = f - feq(1, (u_in_x, 0))
    float qin[9] = {0};  // FP32 inlet equilibrium deviation (k_qin: the step kernels' own arithmetic)
    size_t N = 0;
    size_t esz = 8;
    hipStream_t s = nullptr;
    void* buf[2] = {nullptr, nullptr};
    int cur = 0;
    int phase = 0;  // 0: buf[cur] holds fnew + macro arrays; 1: buf[cur] holds fcol
    int64_t steps = 0;
    double* mac = nullptr;  // rho, ux, uy (3N)
    double* corner = nullptr;  // IPPI_BC_FLATPLATE corner populations [2][9]
    double* mean = nullptr;    // running means of rho, ux, uy (3N), allocated on first use
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
    int *d_pidx = nullptr, *d_ecnt = nullptr, *d_emk = nullptr, *d_act = nullptr;
    IbBlk* d_blk = nullptr;
    int ell = 8;
    double *d_w = nullptr, *d_ud = nullptr, *d_ui = nullptr, *d_F = nullptr, *d_ecw = nullptr, *d_pf = nullptr;
    int npn = 0, nact = 0, nblk_ib = 0;
    int* d_pslot = nullptr;  // [patch node] active node (all bodies' numbering), or -1
    int* d_scat = nullptr;   // [9][region node] active node pulling population k from it, or -1
    void* d_slot = nullptr;  // [9][active] populations the next IB kernel pulls
    int nreg = 0;            // region nodes (patches grown by one node)
    bool slots_ok = false;   // d_slot holds the stored state's populations
    bool prev_pull = false;  // buf[cur ^ 1] holds the state the last (pull) step read
    // IB inside the FP32 step kernel (k_step_f32_ib)
    IbBlk* d_lead = nullptr;      // [body] lead block entry
    unsigned* d_ibflag = nullptr;  // [body] sequence number of the last stored forces
    unsigned* d_rowmask = nullptr; // [ny] bodies whose patches cover each row
    unsigned ib_seq = 0;
    int band0 = 0, band1 = 0;     // rows that read marker forces (the patches)
    bool ink_ok = false;          // bodies fit the in-step IB
    bool ink_enabled = true;      // environment IPPI_IB_IN_STEP=0 turns it off
    bool ink_wait_skew = false;   // environment IPPI_IB_WAIT_SKEW=1: the band waits for a value never
                                  // stored, so its bounded wait times out (tests of that path only)
    int rows64 = 0;               // environment IPPI_F64_ROWS=1 or 4: rows per thread of the FP64 pull
                                  // steps without the network (0: by lattice size, rows64_for)
    int npx_nn = 2;               // nodes per thread of the in-step kernel with the exact-weight
                                  // network (environment IPPI_NN_NPX=1 for one)
    bool adj_overlap = true;      // environment IPPI_ADJ_OVERLAP=0: the adjoint recomputes each segment on
                                  // the handle's stream, between the sweeps (to time the sweep's kernels alone)
    bool bnn_mfma = true;         // environment IPPI_BNN_MFMA=0: the boundary networks as k_bnn's FMA
                                  // chains instead of k_bnn_mfma (the same results)
    bool adj_fused = true;        // environment IPPI_ADJ_FUSED=0: the reverse step as three kernels
                                  // (k_adj_act, k_adj_mk, k_adj_step) instead of k_adj_fused
    bool round32 = false;         // environment IPPI_FP64_ROUND_STORE=1: the FP64 path rounds each stored
                                  // population as the FP32 path stores it (accuracy diagnostic)
    std::vector<int> h_act;      // active node -> patch node (host copy)
    std::vector<int> h_pidx;     // [marker][16] support -> active index within its body (host copy)
    std::vector<int> body_act0;  // first active node of each body (+ total)
    double* hist = nullptr;
    int hist_cap = 0;
    int64_t nrec = 0;
    // relaxation-time network
    bool has_mlp = false;
    MlpFrag* d_frag = nullptr;
    u32x4* d_frage = nullptr;  // exact-weight network fragments, [MLP_FRAGE_CHUNKS][64]
    double* d_blob = nullptr;
    float mlp_mu[6] = {0}, mlp_isig[6] = {0}, mlp_isig_e[6] = {0}, mlp_b3 = 0.f;
    bool mlp_exact = false;  // W1 and W2 hold bf16 values only (the FP32 path's exact-weight network)
    // boundary-row / column networks (types 1, 3, 5, 7 -> index 0..3)
    float* d_bnn_w[4] = {nullptr, nullptr, nullptr, nullptr};
    float* d_bnn_wm[4] = {nullptr, nullptr, nullptr, nullptr};  // the MFMA kernel's layout
    BnnNode* d_bnn_nodes = nullptr;
    int bnn_nnodes = 0;
    int bnn_tend[4] = {0, 0, 0, 0};
    double* d_bnn_tau = nullptr;  // [2 nx + 2 ny]
    int bnn_on = 0;
    // optional per-launch timing (HIP events on this handle's stream)
    int timing = 0;                      // 1: events around each launch; 2: around each ippi_step call
    int64_t batch_steps = 0, batch_ib = 0;  // kernels launched inside the timed calls (mode 2)
    std::vector<hipEvent_t> ev_step, ev_ib, ev_adj;  // start/end pairs
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

// Network mode of the step kernels: 0 none, 1 compensated bf16 products,
// 2 plain bf16 products, 3 bf16-exact weights (2 and 3 FP32 only; FP64
// evaluates the network in FP64).
int nn_kind(const ippi_handle* h) {
    if (h->c.tau_mode == IPPI_TAU_NN) return h->esz == 4 && h->mlp_exact ? 3 : 1;
    if (h->c.tau_mode == IPPI_TAU_NN_BF16) return h->esz == 4 ? 2 : 1;
    return 0;
}

int check_flag(ippi_handle* h) {
    int v = 0;
    HIPC(h, hipMemcpyAsync(&v, h->flag, sizeof(int), hipMemcpyDeviceToHost, h->s));
    HIPC(h, hipStreamSynchronize(h->s));
    if (v == 2) return fail(h, IPPI_ERR_HIP, "in-kernel IB: marker forces or velocity adjoint of a step not stored in time");
    if (v) return fail(h, IPPI_ERR_DIVERGED, "density left [0, 10] or became NaN (ErrorCatch)");
    return IPPI_OK;
}

template <typename T>
using StepKernel = void (*)(const T*, T*, const double*, StepArgs);

// nt: non-temporal stores (pull steps of lattices whose two population
// buffers exceed 512 MB)
// NN: network mode (0 none, 1 compensated, 2 plain bf16; FP64 has one network)
// rows64: rows per thread of the FP64 kernels built for four (rows64_for)
template <typename T, bool PULL, int NN>
StepKernel<T> pick_step(int coll, int bc, bool nt = false, int rows64 = 4) {
#define IPPI_PICK(C, B)                                                                         \
    if (coll == C && bc == B) {                                                                 \
        constexpr int NPT_ = step_npt_bc<T, PULL, (NN != 0), B>();                            \
        if constexpr (sizeof(T) == 8) {                                                         \
            if constexpr (NPT_ == 4)                                                            \
                if (rows64 == 1)                                                                \
                    return PULL && nt ? (StepKernel<T>)k_step_f64<C, B, PULL, 1, (NN != 0), PULL> \
                                      : (StepKernel<T>)k_step_f64<C, B, PULL, 1, (NN != 0), false>; \
            if (PULL && nt) return (StepKernel<T>)k_step_f64<C, B, PULL, NPT_, (NN != 0), PULL>; \
            return (StepKernel<T>)k_step_f64<C, B, PULL, NPT_, (NN != 0), false>;               \
        } else {                                                                                \
            if (PULL && nt) return (StepKernel<T>)k_step_f32<C, B, PULL, NPT_, NN, PULL>;       \
            return (StepKernel<T>)k_step_f32<C, B, PULL, NPT_, NN, false>;                      \
        }                                                                                       \
    }
    IPPI_PICK(IPPI_COLL_BGK, IPPI_BC_PERIODIC_BB)
    IPPI_PICK(IPPI_COLL_MRT, IPPI_BC_PERIODIC_BB)
    IPPI_PICK(IPPI_COLL_RR, IPPI_BC_PERIODIC_BB)
    IPPI_PICK(IPPI_COLL_BGK, IPPI_BC_INFLOW)
    IPPI_PICK(IPPI_COLL_MRT, IPPI_BC_INFLOW)
    IPPI_PICK(IPPI_COLL_RR, IPPI_BC_INFLOW)
    IPPI_PICK(IPPI_COLL_BGK, IPPI_BC_FLATPLATE)
    IPPI_PICK(IPPI_COLL_MRT, IPPI_BC_FLATPLATE)
    IPPI_PICK(IPPI_COLL_RR, IPPI_BC_FLATPLATE)
#undef IPPI_PICK
    return nullptr;
}

template <int COLL, int BC, int NPX>
StepKernel<float> pick_step_ib_bc(int nn, bool nt) {
    if (nn == 3) return nt ? k_step_f32_ib<COLL, BC, 3, true, NPX> : k_step_f32_ib<COLL, BC, 3, false, NPX>;
    if (nn == 2) return nt ? k_step_f32_ib<COLL, BC, 2, true, NPX> : k_step_f32_ib<COLL, BC, 2, false, NPX>;
    if (nn == 1) return nt ? k_step_f32_ib<COLL, BC, 1, true, NPX> : k_step_f32_ib<COLL, BC, 1, false, NPX>;
    return nt ? k_step_f32_ib<COLL, BC, 0, true, 1> : k_step_f32_ib<COLL, BC, 0, false, 1>;
}

template <int COLL, int BC>
StepKernel<double> pick_step_ib64_bc(bool nt, int rows) {
    if (rows == 1) return nt ? k_step_f64_ib<COLL, BC, true, 1> : k_step_f64_ib<COLL, BC, false, 1>;
    return nt ? k_step_f64_ib<COLL, BC, true, 4> : k_step_f64_ib<COLL, BC, false, 4>;
}

// FP64 pull step with the IB inside (no network; periodic and inflow
// boundaries), rows (1 or 4) rows per thread
StepKernel<double> pick_step_ib64(int coll, int bc, bool nt, int rows) {
#define IPPI_PICK_IB(C, B) \
    if (coll == C && bc == B) return pick_step_ib64_bc<C, B>(nt, rows);
    IPPI_PICK_IB(IPPI_COLL_BGK, IPPI_BC_PERIODIC_BB)
    IPPI_PICK_IB(IPPI_COLL_MRT, IPPI_BC_PERIODIC_BB)
    IPPI_PICK_IB(IPPI_COLL_RR, IPPI_BC_PERIODIC_BB)
    IPPI_PICK_IB(IPPI_COLL_BGK, IPPI_BC_INFLOW)
    IPPI_PICK_IB(IPPI_COLL_MRT, IPPI_BC_INFLOW)
    IPPI_PICK_IB(IPPI_COLL_RR, IPPI_BC_INFLOW)
#undef IPPI_PICK_IB
    return nullptr;
}

// FP32 pull step with the IB inside (periodic and inflow boundaries); npx
// nodes per thread (2 with the network, else 1)
StepKernel<float> pick_step_ib(int coll, int bc, int nn, bool nt, int npx) {
#define IPPI_PICK_IB(C, B) \
    if (coll == C && bc == B) return npx == 2 ? pick_step_ib_bc<C, B, 2>(nn, nt) : pick_step_ib_bc<C, B, 1>(nn, nt);
    IPPI_PICK_IB(IPPI_COLL_BGK, IPPI_BC_PERIODIC_BB)
    IPPI_PICK_IB(IPPI_COLL_MRT, IPPI_BC_PERIODIC_BB)
    IPPI_PICK_IB(IPPI_COLL_RR, IPPI_BC_PERIODIC_BB)
    IPPI_PICK_IB(IPPI_COLL_BGK, IPPI_BC_INFLOW)
    IPPI_PICK_IB(IPPI_COLL_MRT, IPPI_BC_INFLOW)
    IPPI_PICK_IB(IPPI_COLL_RR, IPPI_BC_INFLOW)
#undef IPPI_PICK_IB
    return nullptr;
}

// FP32 inlet equilibrium deviation feq(1, u_in) - c (nee32's boundary term)
__global__ void k_qin(RefState r, float duinx, float uiny, float* q) {
    if (threadIdx.x == 0) {
        const float drb = 0.f;
        sh::geq_dev(drb, 1.0f + drb, duinx, uiny, r.Uf, r.cf, q);
    }
}

StepArgs make_args(ippi_handle* h) {
    StepArgs a{};
    a.nx = h->c.nx;
    a.ny = h->c.ny;
    a.ref = h->ref;
    a.duinxf = (float)(h->c.u_in_x - h->ref.U);
    for (int k = 0; k < 9; k++) a.qin[k] = h->qin[k];
    a.tau0 = h->c.tau0;
    a.cs = h->c.cs;
    a.uinx = h->c.u_in_x;
    a.uiny = h->c.u_in_y;
    a.tx0f = (float)(h->c.tau0 - 0.5);
    a.csf = (float)h->c.cs;
    a.uinxf = (float)h->c.u_in_x;
    a.uinyf = (float)h->c.u_in_y;
    a.les = h->c.les;
    a.tau_field = h->c.tau_mode == IPPI_TAU_FIELD;
    a.tau_ptr = h->tau;
    a.patch_f = h->d_pf;
    a.flag = h->flag;
    a.nbody = (int)h->bodies.size();
    a.nt = 2 * 9 * h->N * h->esz > ((size_t)512 << 20);
    a.round32 = h->round32 && h->esz == 8;
    a.nn_max = h->c.tau_nn_max > 0 ? h->c.tau_nn_max : 1e30;
    a.nn_maxtxf = (float)(a.nn_max - 0.5);
    a.mlp.frag = h->d_frag;
    a.mlp.frage = h->d_frage;
    a.mlp.blob = h->d_blob;
    for (int i = 0; i < 6; i++) {
        a.mlp.mu[i] = h->mlp_mu[i];
        a.mlp.isig[i] = h->mlp_isig[i];
        a.mlp.isig_e[i] = h->mlp_isig_e[i];
    }
    a.mlp.b3 = h->mlp_b3;
    a.plate_x = h->c.plate_x;
    a.corner = h->corner;
    a.bnn = h->d_bnn_tau;
    a.bnn_on = h->bnn_on;
    a.mac = h->mac;
    for (size_t b = 0; b < h->bodies.size(); b++) a.patch[b] = h->bodies[b].p;
    a.rowmask = h->d_rowmask;
    a.ink_bodies = 0;
    a.brows = 0;
    return a;
}

// Upload marker weights, the active patch nodes with their gather lists and
// the IB block table after bodies were added.
void free_ib(ippi_handle* h) {
    dfree(h->d_pidx); dfree(h->d_ecnt); dfree(h->d_emk); dfree(h->d_act); dfree(h->d_blk);
    dfree(h->d_w); dfree(h->d_ud); dfree(h->d_ui); dfree(h->d_F); dfree(h->d_ecw); dfree(h->d_pf);
    dfree(h->hist);
    dfree(h->d_scat);
    dfree(h->d_slot);
    dfree(h->d_pslot);
    dfree(h->d_lead);
    dfree(h->d_rowmask);
    dfree(h->d_ibflag);
    h->ink_ok = false;
    h->slots_ok = false;
}

int upload_ib(ippi_handle* h) {
    if (!h->ib_dirty) return IPPI_OK;
    const int nm = (int)(h->mk_ud.size() / 2);
    const int nb = (int)h->bodies.size();
    int off = 0;
    for (auto& b : h->bodies) {
        b.p.off = off;
        b.p.m0 = b.first;
        off += b.p.w * b.p.h;
    }
    h->npn = off;
    // gather lists per patch node, markers in increasing order
    std::vector<std::vector<std::pair<int, double>>> rows(off);
    size_t emax = 0;
    for (int bi = 0; bi < nb; bi++) {
        const auto& b = h->bodies[bi];
        for (int m = b.first; m < b.first + b.count; m++)
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++) {
                    const int X = h->mk_base[2 * m] + i, Y = h->mk_base[2 * m + 1] + j;
                    const int r = b.p.off + (Y - b.p.y0) * b.p.w + (X - b.p.x0);
                    rows[r].push_back({m - b.first, h->mk_w[16 * m + 4 * i + j]});
                    emax = std::max(emax, rows[r].size());
                }
    }
    if (emax > 16) return fail(h, IPPI_ERR_ARG, "more than 16 markers reach one node (markers too dense)");
    const int E = emax <= 8 ? 8 : 16;
    h->ell = E;
    std::vector<int> act, ecnt, emk, pidx(16 * (size_t)nm), slot(off, -1), gslot(off, -1);
    std::vector<double> ecw;
    std::vector<IbBlk> blk;
    std::vector<int> act0s;
    for (int bi = 0; bi < nb; bi++) {
        const auto& b = h->bodies[bi];
        const int a0 = (int)act.size();
        act0s.push_back(a0);
        for (int r = b.p.off; r < b.p.off + b.p.w * b.p.h; r++) {
            if (rows[r].empty()) continue;
            slot[r] = (int)act.size() - a0;
            gslot[r] = (int)act.size();
            act.push_back(r);
            ecnt.push_back((int)rows[r].size());
            for (int e = 0; e < E; e++) {
                const bool v = e < (int)rows[r].size();
                emk.push_back(v ? rows[r][e].first : 0);
                ecw.push_back(v ? rows[r][e].second : 0.0);
            }
        }
        const int a1 = (int)act.size();
        if (a1 - a0 > IB_MAXACT) return fail(h, IPPI_ERR_ARG, "body support too large (active nodes > 6144)");
        if (bi == nb - 1) act0s.push_back(a1);
        for (int m = b.first; m < b.first + b.count; m++)
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++) {
                    const int X = h->mk_base[2 * m] + i, Y = h->mk_base[2 * m + 1] + j;
                    pidx[16 * m + 4 * i + j] = slot[b.p.off + (Y - b.p.y0) * b.p.w + (X - b.p.x0)];
                }
        for (int s = a0; s < a1; s += IB_T)
            blk.push_back(IbBlk{bi, s, std::min(a1, s + IB_T), b.first, b.count, a0, a1 - a0, s == a0 ? 1 : 0,
                                b.p.x0, b.p.y0, b.p.w, b.p.off});
    }
    h->nact = (int)act.size();
    h->nblk_ib = (int)blk.size();
    h->h_act = act;
    h->h_pidx = pidx;
    h->body_act0 = act0s;
    free_ib(h);
    HIPC(h, hipMalloc(&h->d_pidx, sizeof(int) * 16 * nm));
    HIPC(h, hipMalloc(&h->d_w, sizeof(double) * 16 * nm));
    HIPC(h, hipMalloc(&h->d_ud, sizeof(double) * 2 * nm));
    HIPC(h, hipMalloc(&h->d_ui, sizeof(double) * 2 * nm));
    HIPC(h, hipMalloc(&h->d_F, sizeof(double) * 2 * nm));
    HIPC(h, hipMalloc(&h->d_act, sizeof(int) * act.size()));
    HIPC(h, hipMalloc(&h->d_ecnt, sizeof(int) * ecnt.size()));
    HIPC(h, hipMalloc(&h->d_emk, sizeof(int) * emk.size()));
    HIPC(h, hipMalloc(&h->d_ecw, sizeof(double) * ecw.size()));
    HIPC(h, hipMalloc(&h->d_pf, sizeof(double) * 2 * off));
    HIPC(h, hipMalloc(&h->d_blk, sizeof(IbBlk) * blk.size()));
    h->hist_cap = h->c.force_history > 0 ? h->c.force_history : 65536;
    HIPC(h, hipMalloc(&h->hist, sizeof(double) * 2 * (size_t)nb * h->hist_cap));
    HIPC(h, hipMemcpy(h->d_pidx, pidx.data(), sizeof(int) * 16 * nm, hipMemcpyHostToDevice));
    HIPC(h, hipMemcpy(h->d_w, h->mk_w.data(), sizeof(double) * 16 * nm, hipMemcpyHostToDevice));
    HIPC(h, hipMemcpy(h->d_ud, h->mk_ud.data(), sizeof(double) * 2 * nm, hipMemcpyHostToDevice));
    HIPC(h, hipMemset(h->d_ui, 0, sizeof(double) * 2 * nm));
    HIPC(h, hipMemset(h->d_F, 0, sizeof(double) * 2 * nm));
    HIPC(h, hipMemset(h->d_pf, 0, sizeof(double) * 2 * off));  // inactive patch nodes stay zero
    HIPC(h, hipMemcpy(h->d_act, act.data(), sizeof(int) * act.size(), hipMemcpyHostToDevice));
    HIPC(h, hipMemcpy(h->d_ecnt, ecnt.data(), sizeof(int) * ecnt.size(), hipMemcpyHostToDevice));
    HIPC(h, hipMemcpy(h->d_emk, emk.data(), sizeof(int) * emk.size(), hipMemcpyHostToDevice));
    HIPC(h, hipMemcpy(h->d_ecw, ecw.data(), sizeof(double) * ecw.size(), hipMemcpyHostToDevice));
    HIPC(h, hipMemcpy(h->d_blk, blk.data(), sizeof(IbBlk) * blk.size(), hipMemcpyHostToDevice));
    // slot map: active node D pulls population k from D - e_k, a node of D's
    // body region (patch grown by one node)
    {
        int roff = 0;
        for (auto& b : h->bodies) {
            b.p.roff = roff;
            roff += (b.p.w + 2) * (b.p.h + 2);
        }
        h->nreg = roff;
        std::vector<int> scat(9 * (size_t)roff, -1);
        for (int bi = 0; bi < nb; bi++) {
            const Patch& p = h->bodies[bi].p;
            for (int i = act0s[bi]; i < act0s[bi + 1]; i++) {
                const int r = act[i] - p.off, X = p.x0 + r % p.w, Y = p.y0 + r / p.w;
                for (int k = 0; k < 9; k++) {
                    const int sx = X - EX[k], sy = Y - EY[k];
                    const int ri = p.roff + (sy - p.y0 + 1) * (p.w + 2) + (sx - p.x0 + 1);
                    scat[(size_t)k * roff + ri] = i;
                }
            }
        }
        HIPC(h, hipMalloc(&h->d_scat, sizeof(int) * scat.size()));
        HIPC(h, hipMemcpy(h->d_scat, scat.data(), sizeof(int) * scat.size(), hipMemcpyHostToDevice));
        HIPC(h, hipMalloc(&h->d_slot, h->esz * 9 * act.size()));
        HIPC(h, hipMalloc(&h->d_pslot, sizeof(int) * off));
        HIPC(h, hipMemcpy(h->d_pslot, gslot.data(), sizeof(int) * off, hipMemcpyHostToDevice));
        h->slots_ok = false;
    }
    // in-step IB: one lead entry per body, the band of rows that read marker
    // forces (the patches), and whether the bodies fit
    {
        std::vector<IbBlk> lead;
        for (const auto& e : blk)
            if (e.lead) lead.push_back(e);
        int b0 = h->c.ny, b1 = 0;
        const bool nnm = nn_kind(h) != 0;
        bool fit = h->c.bc != IPPI_BC_FLATPLATE && nb <= (h->c.nx + BX - 1) / BX && (int)lead.size() == nb &&
                   (h->esz == 4 || !nnm);
        const int cap = h->esz == 8 ? INK_MAXACT64 : (nnm ? INK_MAXACT_NN : INK_MAXACT);
        for (int bi = 0; bi < nb; bi++) {
            const Patch& p = h->bodies[bi].p;
            b0 = std::min(b0, p.y0);
            b1 = std::max(b1, p.y0 + p.h);
            fit = fit && act0s[bi + 1] - act0s[bi] <= cap;
        }
        h->band0 = std::max(b0, 0);
        h->band1 = std::min(b1, h->c.ny);
        // per-row body masks for the step kernels' patch test
        std::vector<unsigned> rm(h->c.ny, 0u);
        for (int bi = 0; bi < nb; bi++) {
            const Patch& p = h->bodies[bi].p;
            for (int y = std::max(p.y0, 0); y < std::min(p.y0 + p.h, h->c.ny); y++) rm[y] |= 1u << bi;
        }
        HIPC(h, hipMalloc(&h->d_rowmask, sizeof(unsigned) * rm.size()));
        HIPC(h, hipMemcpy(h->d_rowmask, rm.data(), sizeof(unsigned) * rm.size(), hipMemcpyHostToDevice));
        // worth it only while the band is a minority of the rows
        h->ink_ok = fit && h->band1 > h->band0 && 2 * (h->band1 - h->band0) <= h->c.ny;
        if (h->ink_ok) {
            HIPC(h, hipMalloc(&h->d_lead, sizeof(IbBlk) * nb));
            HIPC(h, hipMemcpy(h->d_lead, lead.data(), sizeof(IbBlk) * nb, hipMemcpyHostToDevice));
            HIPC(h, hipMalloc(&h->d_ibflag, sizeof(unsigned) * nb));
            HIPC(h, hipMemset(h->d_ibflag, 0, sizeof(unsigned) * nb));
            h->ib_seq = 0;
        }
    }
    h->nrec = 0;
    h->ib_dirty = false;
    return IPPI_OK;
}

// IB kernel on the handle's stream; slot: read the pulled populations from
// the slots the previous step kernel wrote (pull steps only).
template <typename T>
void launch_ib_to(ippi_handle* h, const T* src, bool pull, const IbDev& d, bool slot = false, bool gather = false,
                  hipStream_t st = nullptr) {
    dim3 g(h->nblk_ib), b(IB_T);
    if (!st) st = h->s;
    const int nx = h->c.nx, ny = h->c.ny, na = h->nact, ga = gather ? 1 : 0;
    const T* sl = (const T*)h->d_slot;
#define IPPI_IB_LAUNCH(E)                                                                                     \
    if (slot)                                                                                                 \
        hipLaunchKernelGGL((k_ib<T, true, E, true>), g, b, 0, st, src, h->mac, d, nx, ny, sl, na, ga);          \
    else if (pull)                                                                                            \
        hipLaunchKernelGGL((k_ib<T, true, E, false>), g, b, 0, st, src, h->mac, d, nx, ny, sl, na, ga);         \
    else                                                                                                      \
        hipLaunchKernelGGL((k_ib<T, false, E, false>), g, b, 0, st, src, h->mac, d, nx, ny, sl, na, ga);
    if (h->ell == 8) {
        IPPI_IB_LAUNCH(8)
    } else {
        IPPI_IB_LAUNCH(16)
    }
#undef IPPI_IB_LAUNCH
}

template <typename T>
int launch_ib(ippi_handle* h, bool pull) {
    const int nb = (int)h->bodies.size();
    IbDev d{h->d_w, h->d_pidx, h->d_ud, h->d_ui, h->d_F, h->d_blk, h->d_act, h->d_ecnt, h->d_emk, h->d_ecw,
            h->d_pf, h->hist + 2 * (size_t)nb * (size_t)(h->nrec % h->hist_cap), h->ref};
    const T* src = (const T*)h->buf[h->cur];
    launch_ib_to<T>(h, src, pull, d, pull && h->slots_ok, true);
    HIPC(h, hipGetLastError());
    h->nrec++;
    return IPPI_OK;
}

// Timing events only: no system-scope fence (cache writeback / invalidate)
// when they are recorded, which would slow the kernels being timed; the
// stream is synchronised before the elapsed times are read.
hipEvent_t new_event(ippi_handle* h) {
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
    (void)hipEventRecord(e, h->s);
    return e;
}

void clear_events(std::vector<hipEvent_t>& v) {
    for (auto e : v)
        if (e) (void)hipEventDestroy(e);
    v.clear();
}

// Total time of the (start, end) event pairs in v; clears v.
int sum_events(ippi_handle* h, std::vector<hipEvent_t>& v, double* ms, int64_t* cnt) {
    double t = 0.0;
    int64_t c = 0;
    for (size_t i = 0; i + 1 < v.size(); i += 2) {
        if (!v[i] || !v[i + 1]) {
            clear_events(v);
            return fail(h, IPPI_ERR_HIP, "event creation failed");
        }
        float e = 0.f;
        hipError_t err = hipEventElapsedTime(&e, v[i], v[i + 1]);
        if (err != hipSuccess) {
            clear_events(v);
            return fail(h, IPPI_ERR_HIP, std::string("hipEventElapsedTime: ") + hipGetErrorString(err));
        }
        t += e;
        c++;
    }
    if (ms) *ms = t;
    if (cnt) *cnt = c;
    clear_events(v);
    return IPPI_OK;
}

// Timing mode 2: one event pair around the launches of one ippi_step call
// (also when the call returns early with an error).
struct CallTimer {
    ippi_handle* h;
    explicit CallTimer(ippi_handle* hh) : h(hh) {
        if (h->timing == 2) h->ev_step.push_back(new_event(h));
    }
    ~CallTimer() {
        if (h->timing == 2) h->ev_step.push_back(new_event(h));
    }
};

// Boundary-network tau of the state buf[cur] holds (before the step kernel
// that reads it, or before a tau read-back).
template <typename T>
int launch_bnn(ippi_handle* h, const StepArgs& a, bool pull) {
    BnnDev d{h->d_bnn_nodes, h->bnn_nnodes, {h->d_bnn_w[0], h->d_bnn_w[1], h->d_bnn_w[2], h->d_bnn_w[3]},
             h->d_bnn_tau, {h->bnn_tend[0], h->bnn_tend[1], h->bnn_tend[2], h->bnn_tend[3]},
             {h->d_bnn_wm[0], h->d_bnn_wm[1], h->d_bnn_wm[2], h->d_bnn_wm[3]}};
    const dim3 g((unsigned)((h->bnn_nnodes + BNN_NODES - 1) / BNN_NODES));
    const T* src = (const T*)h->buf[h->cur];
    const int p = pull ? 1 : 0;
    if (h->bnn_mfma) {  // the MFMA networks (bitwise k_bnn's results, see k_bnn_mfma)
        const dim3 gm((unsigned)((h->bnn_nnodes + BNM_NODES - 1) / BNM_NODES));
        if (h->c.bc == IPPI_BC_PERIODIC_BB)
            hipLaunchKernelGGL((k_bnn_mfma<T, IPPI_BC_PERIODIC_BB>), gm, dim3(BNM_T), 0, h->s, src, h->mac, h->corner,
                               p, a, d);
        else if (h->c.bc == IPPI_BC_INFLOW)
            hipLaunchKernelGGL((k_bnn_mfma<T, IPPI_BC_INFLOW>), gm, dim3(BNM_T), 0, h->s, src, h->mac, h->corner, p,
                               a, d);
        else
            hipLaunchKernelGGL((k_bnn_mfma<T, IPPI_BC_FLATPLATE>), gm, dim3(BNM_T), 0, h->s, src, h->mac, h->corner,
                               p, a, d);
        HIPC(h, hipGetLastError());
        return IPPI_OK;
    }
    if (h->c.bc == IPPI_BC_PERIODIC_BB)
        hipLaunchKernelGGL((k_bnn<T, IPPI_BC_PERIODIC_BB>), g, dim3(BNN_T), 0, h->s, src, h->mac, h->corner, p, a, d);
    else if (h->c.bc == IPPI_BC_INFLOW)
        hipLaunchKernelGGL((k_bnn<T, IPPI_BC_INFLOW>), g, dim3(BNN_T), 0, h->s, src, h->mac, h->corner, p, a, d);
    else
        hipLaunchKernelGGL((k_bnn<T, IPPI_BC_FLATPLATE>), g, dim3(BNN_T), 0, h->s, src, h->mac, h->corner, p, a, d);
    HIPC(h, hipGetLastError());
    return IPPI_OK;
}

template <typename T>
int run_steps(ippi_handle* h, int nsteps) {
    CallTimer call_timer(h);
    const int nx = h->c.nx, ny = h->c.ny;
    const bool nn_mode = nn_kind(h) != 0;
    int npt = h->c.bc == IPPI_BC_FLATPLATE
                  ? (nn_mode ? step_npt_bc<T, true, true, IPPI_BC_FLATPLATE>()
                             : step_npt_bc<T, true, false, IPPI_BC_FLATPLATE>())
                  : (nn_mode ? step_npt<T, true, true>() : step_npt<T, true, false>());
    const int rows64 = h->rows64 ? h->rows64 : rows64_for(nx, ny);
    if (npt == 4) npt = rows64;  // FP64 kernels built for four rows: per lattice
    const dim3 grid_pull((nx + BX - 1) / BX, (ny + npt - 1) / npt), grid_first((nx + BX - 1) / BX, ny);
    const bool ib = !h->bodies.empty();
    if (ib) {
        int r = upload_ib(h);
        if (r) return r;
    }
    StepArgs a = make_args(h);
    if (ib) {  // patch forces gathered from the marker forces the IB kernel leaves
        a.ib_gather = 1;
        a.ell = h->ell;
        a.pslot = h->d_pslot;
        a.ecnt = h->d_ecnt;
        a.emk = h->d_emk;
        a.ecw = h->d_ecw;
        a.F = h->d_F;
    }
    if (ib && sizeof(T) == 4 && !nn_mode) {  // FP32 step kernels without the network
                                                                 // copy what the next IB kernel pulls
        a.scat = h->d_scat;
        a.slot = h->d_slot;
        a.nreg = h->nreg;
        a.nact = h->nact;
    }
    for (int t = 0; t < nsteps; t++) {
        const bool pull = h->phase == 1;
        if (!pull && h->c.bc == IPPI_BC_FLATPLATE) {
            hipLaunchKernelGGL(k_corner<T>, dim3(1), dim3(64), 0, h->s, (const T*)h->buf[h->cur], h->corner, nx,
                               (unsigned)h->N, h->ref);
            HIPC(h, hipGetLastError());
        }
        const int nk = nn_kind(h);
        const bool nn = nk != 0;
        if (h->bnn_on) {
            int r = launch_bnn<T>(h, a, pull);
            if (r) return r;
        }
        // IB inside the step kernel (pull steps; FP64 without the network)
        if (ib && pull && h->ink_ok && h->ink_enabled && (sizeof(T) == 4 || !nn)) {
            const int nb = (int)h->bodies.size();
            StepArgs ai = a;
            ai.scat = nullptr;  // the in-step IB pulls: no slots
            ai.ib_lead = h->d_lead;
            ai.ib_act = h->d_act;
            ai.ib_pidx = h->d_pidx;
            ai.ib_w = h->d_w;
            ai.ib_ud = h->d_ud;
            ai.ib_ui = h->d_ui;
            ai.ib_hist = h->hist + 2 * (size_t)nb * (size_t)(h->nrec % h->hist_cap);
            ai.ib_flag = h->d_ibflag;
            ai.ib_seq = ++h->ib_seq;
            if (ai.ib_seq == 0) ai.ib_seq = ++h->ib_seq;  // 0 is the flags' initial value
            ai.ib_wait = ai.ib_seq + (h->ink_wait_skew ? 1u : 0u);
            ai.band0 = h->band0;
            ai.band1 = h->band1;
            ai.ink_bodies = nb;
            const int at = (int)((nn ? INK_BAND_AT_NN : INK_BAND_AT) * ny);
            if (h->timing == 1) h->ev_step.push_back(new_event(h));
            if (sizeof(T) == 4) {
                const int npx = nk == 3 ? h->npx_nn : 1;  // the compensated network needs the registers
                StepKernel<float> k = pick_step_ib(h->c.collision, h->c.bc, nk, a.nt, npx);
                if (!k) return fail(h, IPPI_ERR_ARG, "unsupported collision / boundary combination");
                ai.rot = ((h->band0 - at) % ny + ny) % ny;
                ai.xcd = !a.nt;
                // the lattice's width; the bodies take as many leading rows as they need
                const int gx = (nx + npx * BX - 1) / (npx * BX);
                ai.brows = (nb + gx - 1) / gx;
                hipLaunchKernelGGL(k, dim3(gx, ny + ai.brows), dim3(BX), 0, h->s,
                                   (const float*)h->buf[h->cur], (float*)h->buf[h->cur ^ 1], h->mac, ai);
            } else {
                const int NPT = h->rows64 ? h->rows64 : rows64_for(nx, ny);
                StepKernel<double> k = pick_step_ib64(h->c.collision, h->c.bc, a.nt, NPT);
                if (!k) return fail(h, IPPI_ERR_ARG, "unsupported collision / boundary combination");
                const int G = (ny + NPT - 1) / NPT;
                ai.rot = ((h->band0 / NPT - at / NPT) % G + G) % G;
                ai.xcd = !a.nt;
                const int gx = (nx + BX - 1) / BX;
                ai.brows = (nb + gx - 1) / gx;
                hipLaunchKernelGGL(k, dim3(gx, G + ai.brows), dim3(BX), 0, h->s,
                                   (const double*)h->buf[h->cur], (double*)h->buf[h->cur ^ 1], h->mac, ai);
            }
            HIPC(h, hipGetLastError());
            if (h->timing == 1) h->ev_step.push_back(new_event(h));
            h->batch_steps++;
            h->nrec++;
            h->slots_ok = false;
            h->prev_pull = true;
            h->cur ^= 1;
            h->steps++;
            continue;
        }
        if (ib) {
            if (h->timing == 1) h->ev_ib.push_back(new_event(h));
            int r = launch_ib<T>(h, pull);
            if (r) return r;
            if (h->timing == 1) h->ev_ib.push_back(new_event(h));
            h->batch_ib++;
        }
        if constexpr (sizeof(T) == 4) {
            // exact-weight network on a pull step without in-step bodies: the
            // in-step kernel's two nodes per thread, with no body blocks (the
            // forces, if any, come from the k_ib launch above)
            StepKernel<float> k2 = pull && nk == 3 ? pick_step_ib(h->c.collision, h->c.bc, 3, a.nt, h->npx_nn) : nullptr;
            if (k2) {
                StepArgs ai = a;
                ai.ink_bodies = 0;
                ai.band0 = ai.band1 = 0;
                ai.rot = 0;
                ai.xcd = !a.nt;
                const int gx = (nx + h->npx_nn * BX - 1) / (h->npx_nn * BX);
                if (h->timing == 1) h->ev_step.push_back(new_event(h));
                ai.brows = 0;
                hipLaunchKernelGGL(k2, dim3(gx, ny), dim3(BX), 0, h->s, (const float*)h->buf[h->cur],
                                   (float*)h->buf[h->cur ^ 1], h->mac, ai);
                HIPC(h, hipGetLastError());
                if (h->timing == 1) h->ev_step.push_back(new_event(h));
                h->batch_steps++;
                h->slots_ok = false;
                h->prev_pull = true;
                h->cur ^= 1;
                h->phase = 1;
                h->steps++;
                continue;
            }
        }
        StepKernel<T> k = pull ? (nk == 3   ? pick_step<T, true, 3>(h->c.collision, h->c.bc, a.nt)
                                  : nk == 2 ? pick_step<T, true, 2>(h->c.collision, h->c.bc, a.nt)
                                  : nk == 1 ? pick_step<T, true, 1>(h->c.collision, h->c.bc, a.nt)
                                            : pick_step<T, true, 0>(h->c.collision, h->c.bc, a.nt, rows64))
                               : (nk == 3   ? pick_step<T, false, 3>(h->c.collision, h->c.bc)
                                  : nk == 2 ? pick_step<T, false, 2>(h->c.collision, h->c.bc)
                                  : nk == 1 ? pick_step<T, false, 1>(h->c.collision, h->c.bc)
                                            : pick_step<T, false, 0>(h->c.collision, h->c.bc));
        if (!k) return fail(h, IPPI_ERR_ARG, "unsupported collision / boundary combination");
        // XCD-grouped block order while the stores stay temporal
        a.xcd = !(pull && a.nt);
        if (h->timing == 1 && pull) h->ev_step.push_back(new_event(h));
        hipLaunchKernelGGL(k, pull ? grid_pull : grid_first, dim3(BX), 0, h->s, (const T*)h->buf[h->cur],
                           (T*)h->buf[h->cur ^ 1], h->mac, a);
        HIPC(h, hipGetLastError());
        if (h->timing == 1 && pull) h->ev_step.push_back(new_event(h));
        h->batch_steps++;
        h->slots_ok = a.scat != nullptr;
        h->prev_pull = pull;
        h->cur ^= 1;
        h->phase = 1;
        h->steps++;
    }
    return IPPI_OK;
}

template <typename T>
int run_macro(ippi_handle* h, double* f_out, double* mean = nullptr, double kk = 0.0) {
    if (h->phase == 0 && mean) {  // before the first step: the stored initial state
        hipLaunchKernelGGL(k_mean_phase0, dim3((unsigned)((h->N + 255) / 256)), dim3(256), 0, h->s, h->mac, mean,
                           h->N, kk);
        HIPC(h, hipGetLastError());
        return IPPI_OK;
    }
    if (h->phase == 0) {
        if (f_out) {  // fnew state stored directly
            std::vector<T> tmp(9 * h->N);
            HIPC(h, hipMemcpyAsync(tmp.data(), h->buf[h->cur], sizeof(T) * 9 * h->N, hipMemcpyDeviceToHost, h->s));
            HIPC(h, hipStreamSynchronize(h->s));
            for (size_t k = 0; k < 9; k++)
                for (size_t n = 0; n < h->N; n++)
                    f_out[k * h->N + n] =
                        sizeof(T) == 8 ? (double)tmp[k * h->N + n] : (double)tmp[k * h->N + n] + h->ref.c[k];
        }
        return IPPI_OK;
    }
    double* d_f = nullptr;
    if (f_out) HIPC(h, hipMalloc(&d_f, sizeof(double) * 9 * h->N));
    dim3 grid((h->c.nx + BX - 1) / BX, h->c.ny);
    const T* src = (const T*)h->buf[h->cur];
    const int nx = h->c.nx, ny = h->c.ny, px = h->c.plate_x;
    const double ux = h->c.u_in_x, uy = h->c.u_in_y;
    if (h->c.bc == IPPI_BC_PERIODIC_BB)
        hipLaunchKernelGGL((k_macro<T, IPPI_BC_PERIODIC_BB>), grid, dim3(BX), 0, h->s, src, h->mac, d_f, nx, ny, ux, uy,
                           px, h->corner, mean, kk, h->ref);
    else if (h->c.bc == IPPI_BC_INFLOW)
        hipLaunchKernelGGL((k_macro<T, IPPI_BC_INFLOW>), grid, dim3(BX), 0, h->s, src, h->mac, d_f, nx, ny, ux, uy,
                           px, h->corner, mean, kk, h->ref);
    else
        hipLaunchKernelGGL((k_macro<T, IPPI_BC_FLATPLATE>), grid, dim3(BX), 0, h->s, src, h->mac, d_f, nx, ny, ux, uy,
                           px, h->corner, mean, kk, h->ref);
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

// ------------------------------------------------------------ adjoint run
// Scratch of one adjoint run, freed when it ends.
struct AdjBufs {
    std::vector<void*> ptrs;
    ~AdjBufs() { release(); }
    void release() {
        for (void* p : ptrs) (void)hipFree(p);
        ptrs.clear();
    }
    template <typename P>
    hipError_t alloc(P** p, size_t bytes) {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, bytes > 0 ? bytes : 16);
        if (e == hipSuccess) ptrs.push_back(q);
        *p = (P*)q;
        return e;
    }
};


template <typename T, typename A>
int adjoint_run(ippi_handle* h, int nsteps, int avg_from, int ckpt, const double* bw, double* J, double* g_tau,
                double* g_c) {
    const int nx = h->c.nx, ny = h->c.ny, nb = (int)h->bodies.size();
    const size_t N = h->N, SZ = 9 * N;
    int r = upload_ib(h);
    if (r) return r;
    const int nwin = nsteps - avg_from + 1;
    if (nwin > h->hist_cap) return fail(h, IPPI_ERR_ARG, "averaging window longer than the force history");
    const int nm = (int)(h->mk_ud.size() / 2);
    // marker weight derivatives and the per-body tables
    std::vector<double> wx(16 * (size_t)nm), wy(16 * (size_t)nm);
    for (int m = 0; m < nm; m++)
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) {
                const double dx = h->mk_xy[2 * m] - (h->mk_base[2 * m] + i);
                const double dy = h->mk_xy[2 * m + 1] - (h->mk_base[2 * m + 1] + j);
                wx[16 * m + 4 * i + j] = adj::ddelta4(dx) * delta4(dy);
                wy[16 * m + 4 * i + j] = delta4(dx) * adj::ddelta4(dy);
            }
    std::vector<int> act_node(h->nact);
    std::vector<AdjBody> tab(nb);
    for (int b = 0; b < nb; b++) {
        const Patch& p = h->bodies[b].p;
        for (int i = h->body_act0[b]; i < h->body_act0[b + 1]; i++) {
            const int q = h->h_act[i] - p.off;
            act_node[i] = (p.y0 + q / p.w) * nx + p.x0 + q % p.w;
        }
        tab[b] = AdjBody{h->bodies[b].first, h->bodies[b].count, h->body_act0[b],
                         h->body_act0[b + 1] - h->body_act0[b], bw ? bw[2 * b] : (b == 0 ? 1.0 : 0.0),
                         bw ? bw[2 * b + 1] : 0.0};
    }
    const int K = std::min(ckpt, nsteps);
    const int nseg = (nsteps + K - 1) / K;
    AdjBufs mem;
    T *ck = nullptr, *seg = nullptr;
    A *lam = nullptr, *gacc = nullptr;
    double *pf = nullptr, *Fs = nullptr, *ui = nullptr, *gtot = nullptr, *lf = nullptr, *uact = nullptr;
    double *pl = nullptr, *gc = nullptr, *dwx = nullptr, *dwy = nullptr;
    int* dan = nullptr;
    AdjBody* dtab = nullptr;
    HIPC(h, mem.alloc(&ck, sizeof(T) * SZ * nseg));
    HIPC(h, mem.alloc(&seg, sizeof(T) * SZ * (K - 1)));
    HIPC(h, mem.alloc(&lam, sizeof(A) * SZ * 2));
    HIPC(h, mem.alloc(&gacc, sizeof(A) * N));
    HIPC(h, mem.alloc(&gtot, sizeof(double) * N));
    HIPC(h, mem.alloc(&pf, sizeof(double) * 2 * (size_t)h->npn * K));
    HIPC(h, mem.alloc(&Fs, sizeof(double) * 2 * (size_t)nm * K));
    HIPC(h, mem.alloc(&ui, sizeof(double) * 2 * (size_t)nm));
    HIPC(h, mem.alloc(&lf, sizeof(double) * 2 * (size_t)h->nact));
    HIPC(h, mem.alloc(&uact, sizeof(double) * 2 * (size_t)h->nact));
    HIPC(h, mem.alloc(&pl, sizeof(double) * 2 * (size_t)h->npn));
    HIPC(h, mem.alloc(&gc, sizeof(double) * 2 * nb));
    HIPC(h, mem.alloc(&dwx, sizeof(double) * 16 * (size_t)nm));
    HIPC(h, mem.alloc(&dwy, sizeof(double) * 16 * (size_t)nm));
    HIPC(h, mem.alloc(&dan, sizeof(int) * (size_t)h->nact));
    HIPC(h, mem.alloc(&dtab, sizeof(AdjBody) * nb));
    HIPC(h, hipMemcpyAsync(dwx, wx.data(), sizeof(double) * 16 * nm, hipMemcpyHostToDevice, h->s));
    HIPC(h, hipMemcpyAsync(dwy, wy.data(), sizeof(double) * 16 * nm, hipMemcpyHostToDevice, h->s));
    HIPC(h, hipMemcpyAsync(dan, act_node.data(), sizeof(int) * h->nact, hipMemcpyHostToDevice, h->s));
    HIPC(h, hipMemcpyAsync(dtab, tab.data(), sizeof(AdjBody) * nb, hipMemcpyHostToDevice, h->s));
    HIPC(h, hipMemsetAsync(pf, 0, sizeof(double) * 2 * (size_t)h->npn * K, h->s));  // inactive patch nodes
    HIPC(h, hipMemsetAsync(pl, 0, sizeof(double) * 2 * (size_t)h->npn, h->s));
    HIPC(h, hipMemsetAsync(gc, 0, sizeof(double) * 2 * nb, h->s));
    HIPC(h, hipMemsetAsync(lam, 0, sizeof(A) * SZ, h->s));  // adjoint after the last step
    HIPC(h, hipMemsetAsync(gacc, 0, sizeof(A) * N, h->s));
    HIPC(h, hipMemsetAsync(gtot, 0, sizeof(double) * N, h->s));

    // forward pass, checkpointing the stored state every K steps
    const int phase0 = h->phase;
    const int64_t rec0 = h->nrec;
    for (int sg = 0; sg < nseg; sg++) {
        HIPC(h, hipMemcpyAsync(ck + SZ * sg, h->buf[h->cur], sizeof(T) * SZ, hipMemcpyDeviceToDevice, h->s));
        r = run_steps<T>(h, std::min(K, nsteps - sg * K));
        if (r) return r;
    }
    r = check_flag(h);
    if (r) return r;
    // objective from the force records of the window (a contiguous range of
    // the record ring, copied in at most two pieces)
    {
        const size_t rb = 2 * (size_t)nb;
        std::vector<double> rec(rb * (size_t)nwin);
        const int64_t s0 = (rec0 + avg_from - 1) % h->hist_cap;
        const int64_t n1 = std::min<int64_t>(nwin, h->hist_cap - s0);
        HIPC(h, hipMemcpy(rec.data(), h->hist + rb * s0, sizeof(double) * rb * n1, hipMemcpyDeviceToHost));
        if (n1 < nwin)
            HIPC(h, hipMemcpy(rec.data() + rb * n1, h->hist, sizeof(double) * rb * (nwin - n1),
                              hipMemcpyDeviceToHost));
        double acc = 0.0;
        for (int t = 0; t < nwin; t++)
            for (int b = 0; b < nb; b++) acc += tab[b].wx * rec[rb * t + 2 * b] + tab[b].wy * rec[rb * t + 2 * b + 1];
        if (J) *J = acc / nwin;
    }

    // reverse sweep
    StepArgs sa = make_args(h);
    AdjArgs aa{};
    aa.nx = nx;
    aa.ny = ny;
    aa.tau0 = h->c.tau0;
    aa.ref = h->ref;
    aa.tau_field = h->c.tau_mode == IPPI_TAU_FIELD;
    aa.tau_ptr = h->tau;
    aa.patch_l = pl;
    aa.nbody = nb;
    // XCD-grouped rows while the two adjoint buffers stay within 512 MB (temporal stores)
    aa.xcd = 2 * 9 * h->N * sizeof(A) <= ((size_t)512 << 20);
    for (int b = 0; b < nb; b++) aa.patch[b] = h->bodies[b].p;
    AdjIb ad{h->nact, h->ell, h->d_act, dan, lf, uact, h->d_pidx, h->d_w, dwx, dwy, nullptr,
             h->d_ecnt, h->d_emk, h->d_ecw, dtab, pl, gc};
    // one launch per reverse step (k_adj_fused) unless IPPI_ADJ_FUSED=0: the
    // bodies' blocks publish through per-body flags, sequence numbers from 1
    bool fused = h->adj_fused && nb > 0 && h->d_rowmask;
    for (int b = 0; b < nb; b++) fused = fused && tab[b].M > 0;  // every body publishes its flag
    unsigned* flags = nullptr;
    unsigned seq = 0;
    int nab = 0;  // row-0 blocks: ceil(markers / ADJ_MPB) per body
    for (int b = 0; b < nb; b++) nab += (tab[b].M + ADJ_MPB - 1) / ADJ_MPB;
    std::vector<int> pnr;  // the body blocks' pairs: support node and patch node
    if (fused) {
        pnr.resize(32 * (size_t)nm);
        for (int b = 0; b < nb; b++)
            for (int mk = tab[b].m0; mk < tab[b].m0 + tab[b].M; mk++)
                for (int sp = 0; sp < 16; sp++) {
                    const int i = tab[b].a0 + h->h_pidx[16 * (size_t)mk + sp];
                    pnr[2 * (16 * (size_t)mk + sp)] = act_node[i];
                    pnr[2 * (16 * (size_t)mk + sp) + 1] = h->h_act[i];
                }
        int* dpnr = nullptr;
        HIPC(h, mem.alloc(&dpnr, sizeof(int) * 32 * (size_t)nm));
        HIPC(h, hipMemcpyAsync(dpnr, pnr.data(), sizeof(int) * 32 * (size_t)nm, hipMemcpyHostToDevice, h->s));
        ad.pnr = dpnr;
        HIPC(h, mem.alloc(&flags, sizeof(unsigned) * 2 * nb));
        HIPC(h, hipMemsetAsync(flags, 0, sizeof(unsigned) * 2 * nb, h->s));
        HIPC(h, mem.alloc(&ad.lu, sizeof(double) * 2 * (size_t)nm));
        HIPC(h, mem.alloc(&ad.pc, sizeof(double) * 32 * (size_t)nm));
        ad.pslot = h->d_pslot;
        aa.cnt = flags + nb;
        aa.rowmask = h->d_rowmask;
        aa.flags = flags;
        aa.err = h->flag;
        aa.band0 = h->band0;
        aa.band1 = h->band1;
        // the band wholly inside the order (a long band ends the order)
        const int at = std::min((int)(ADJ_BAND_AT * ny), ny - (h->band1 - h->band0));
        aa.rot = ((h->band0 - at) % ny + ny) % ny;  // lattice row of the first lattice grid row
    }
    const int gxa = (nx + BX - 1) / BX;
    aa.brows = (nab + gxa - 1) / gxa;
    const dim3 grid_fused(gxa, ny + aa.brows);
    const int rows64 = h->rows64 ? h->rows64 : rows64_for(nx, ny);
    const int npt = step_npt<T, true, false>() == 4 ? rows64 : step_npt<T, true, false>();
    const dim3 grid_pull((nx + BX - 1) / BX, (ny + npt - 1) / npt), grid_first((nx + BX - 1) / BX, ny);
    const dim3 grid_adj((nx + BX - 1) / BX, ny), grid_act((h->nact + 255) / 256);
    const StepKernel<T> kpull = pick_step<T, true, false>(IPPI_COLL_BGK, IPPI_BC_PERIODIC_BB, false, rows64);
    const StepKernel<T> kfirst = pick_step<T, false, false>(IPPI_COLL_BGK, IPPI_BC_PERIODIC_BB);
    // Two buffer sets for the recomputed segments (states, patch forces,
    // marker forces): while the reverse sweep of segment sg runs on the
    // handle's stream, segment sg - 1 is recomputed into the other set on a
    // second stream (default priority). The sweep's IB kernels are small and
    // leave most of the GPU idle; the recompute fills it. Used while the
    // second set of states fits in 8 GB; otherwise, or with one segment, the
    // two phases alternate on the handle's stream. The results are the same
    // either way: the same kernels read the same inputs.
    AdjBufs mem2;  // the second set (declared before the stream, so freed after it is synchronised)
    struct Overlap {
        hipStream_t s = nullptr;
        hipEvent_t rec[2] = {nullptr, nullptr}, rev[2] = {nullptr, nullptr};
        ~Overlap() {
            if (s) {
                (void)hipStreamSynchronize(s);
                (void)hipStreamDestroy(s);
            }
            for (hipEvent_t e : rec)
                if (e) (void)hipEventDestroy(e);
            for (hipEvent_t e : rev)
                if (e) (void)hipEventDestroy(e);
        }
    } ov;
    // the second set of segment buffers: at most 8 GB, only while it leaves
    // 1 GB of the device free, and without it (one stream) if allocating fails
    const size_t ov_bytes = sizeof(T) * SZ * (K - 1) + sizeof(double) * 2 * ((size_t)h->npn + (size_t)nm) * K;
    bool overlap = h->adj_overlap && nseg > 1 && ov_bytes <= ((size_t)8 << 30);
    if (overlap) {
        size_t fr = 0, tot = 0;
        overlap = hipMemGetInfo(&fr, &tot) == hipSuccess && fr >= ov_bytes + ((size_t)1 << 30);
    }
    T* segs[2] = {seg, seg};
    double* pfs[2] = {pf, pf};
    double* Fss[2] = {Fs, Fs};
    if (overlap && (mem2.alloc(&segs[1], sizeof(T) * SZ * (K - 1)) != hipSuccess ||
                    mem2.alloc(&pfs[1], sizeof(double) * 2 * (size_t)h->npn * K) != hipSuccess ||
                    mem2.alloc(&Fss[1], sizeof(double) * 2 * (size_t)nm * K) != hipSuccess)) {
        (void)hipGetLastError();  // clear the allocation error; run on one stream
        mem2.release();
        segs[1] = seg;
        pfs[1] = pf;
        Fss[1] = Fs;
        overlap = false;
    }
    if (overlap) {
        HIPC(h, hipMemsetAsync(pfs[1], 0, sizeof(double) * 2 * (size_t)h->npn * K, h->s));  // inactive patch nodes
        HIPC(h, hipStreamCreateWithFlags(&ov.s, hipStreamNonBlocking));
        for (int b = 0; b < 2; b++) {
            HIPC(h, hipEventCreateWithFlags(&ov.rec[b], hipEventDisableTiming));
            HIPC(h, hipEventCreateWithFlags(&ov.rev[b], hipEventDisableTiming));
        }
        // the second stream starts after the forward pass and the clears
        HIPC(h, hipEventRecord(ov.rev[0], h->s));
        HIPC(h, hipStreamWaitEvent(ov.s, ov.rev[0], 0));
    }
    // state i of segment sg: its checkpoint, then the recomputed states
    auto state = [&](int sg, int i) { return i == 0 ? ck + SZ * sg : segs[sg & 1] + SZ * (i - 1); };
    // recompute segment sg on stream st: states, patch forces and marker forces
    auto recompute = [&](int sg, hipStream_t st) {
        const int t0 = sg * K, Ks = std::min(K, nsteps - t0);
        double* pfb = pfs[sg & 1];
        double* Fsb = Fss[sg & 1];
        StepArgs sr = sa;
        for (int i = 0; i < Ks; i++) {
            const bool pull = !(t0 + i == 0 && phase0 == 0);
            IbDev d{h->d_w, h->d_pidx, h->d_ud, ui, Fsb + 2 * (size_t)nm * i, h->d_blk, h->d_act, h->d_ecnt,
                    h->d_emk, h->d_ecw, pfb + 2 * (size_t)h->npn * i, nullptr, h->ref};
            launch_ib_to<T>(h, state(sg, i), pull, d, false, false, st);
            if (i + 1 < Ks) {
                sr.patch_f = pfb + 2 * (size_t)h->npn * i;
                hipLaunchKernelGGL(pull ? kpull : kfirst, pull ? grid_pull : grid_first, dim3(BX), 0, st,
                                   (const T*)state(sg, i), state(sg, i + 1), h->mac, sr);
            }
        }
    };
    int cl = 0;
    if (overlap) {
        recompute(nseg - 1, ov.s);
        HIPC(h, hipEventRecord(ov.rec[(nseg - 1) & 1], ov.s));
    }
    for (int sg = nseg - 1; sg >= 0; sg--) {
        const int t0 = sg * K, Ks = std::min(K, nsteps - t0);
        if (overlap) {
            if (sg > 0) {
                // set (sg - 1) & 1 was last read by the sweep of segment sg + 1
                if (sg + 1 < nseg) HIPC(h, hipStreamWaitEvent(ov.s, ov.rev[(sg + 1) & 1], 0));
                recompute(sg - 1, ov.s);
                HIPC(h, hipEventRecord(ov.rec[(sg - 1) & 1], ov.s));
            }
            HIPC(h, hipStreamWaitEvent(h->s, ov.rec[sg & 1], 0));
        } else {
            recompute(sg, h->s);
        }
        HIPC(h, hipGetLastError());
        const double* pfb = pfs[sg & 1];
        const double* Fsb = Fss[sg & 1];
        for (int i = Ks - 1; i >= 0; i--) {
            const int t = t0 + i;
            const bool pull = !(t == 0 && phase0 == 0);
            const double cj = t + 1 >= avg_from ? 1.0 / nwin : 0.0;
            const A* lin = lam + SZ * cl;
            A* lout = lam + SZ * (cl ^ 1);
            aa.patch_f = pfb + 2 * (size_t)h->npn * i;
            aa.store = t > 0;
            ad.F = Fsb + 2 * (size_t)nm * i;
            if (fused) {  // one launch: the bodies' blocks first, the lattice rows after
                aa.seq = ++seq;
                if (h->timing) h->ev_adj.push_back(new_event(h));
                if (pull)
                    hipLaunchKernelGGL((k_adj_fused<T, A, true>), grid_fused, dim3(BX), 0, h->s, lin, lout,
                                       state(sg, i), h->mac, gacc, aa, ad, cj);
                else
                    hipLaunchKernelGGL((k_adj_fused<T, A, false>), grid_fused, dim3(BX), 0, h->s, lin, lout,
                                       state(sg, i), h->mac, gacc, aa, ad, cj);
                if (h->timing) h->ev_adj.push_back(new_event(h));
                cl ^= 1;
                continue;
            }
            if (pull)
                hipLaunchKernelGGL((k_adj_act<T, A, true>), grid_act, dim3(256), 0, h->s, lin, state(sg, i), h->mac,
                                   ad, aa);
            else
                hipLaunchKernelGGL((k_adj_act<T, A, false>), grid_act, dim3(256), 0, h->s, lin, state(sg, i), h->mac,
                                   ad, aa);
            hipLaunchKernelGGL(k_adj_mk, dim3(nb), dim3(IB_T), 0, h->s, ad, cj);
            if (h->timing) h->ev_adj.push_back(new_event(h));
            if (pull)
                hipLaunchKernelGGL((k_adj_step<T, A, true>), grid_adj, dim3(BX), 0, h->s, lin, lout, state(sg, i),
                                   h->mac, gacc, aa);
            else
                hipLaunchKernelGGL((k_adj_step<T, A, false>), grid_adj, dim3(BX), 0, h->s, lin, lout, state(sg, i),
                                   h->mac, gacc, aa);
            if (h->timing) h->ev_adj.push_back(new_event(h));
            cl ^= 1;
        }
        hipLaunchKernelGGL(k_adj_flush<A>, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, h->s, gacc, gtot, N);
        if (overlap) HIPC(h, hipEventRecord(ov.rev[sg & 1], h->s));
        HIPC(h, hipGetLastError());
    }
    if (g_tau) HIPC(h, hipMemcpyAsync(g_tau, gtot, sizeof(double) * N, hipMemcpyDeviceToHost, h->s));
    if (g_c) HIPC(h, hipMemcpyAsync(g_c, gc, sizeof(double) * 2 * nb, hipMemcpyDeviceToHost, h->s));
    return check_flag(h);  // synchronises; reports a k_adj_fused wait that timed out
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
        cfg->precision > 1 || cfg->bc < 0 || cfg->bc > 2 || cfg->tau_mode < 0 || cfg->tau_mode > 3 ||
        !(cfg->tau0 > 0.5) || cfg->tau_nn_max < 0)
        return IPPI_ERR_ARG;
    if (9 * (uint64_t)cfg->nx * (uint64_t)cfg->ny >= ((uint64_t)1 << 32)) return IPPI_ERR_ARG;
    ippi_handle* h = new ippi_handle();
    h->c = *cfg;
    if (const char* e = std::getenv("IPPI_IB_IN_STEP")) h->ink_enabled = std::strcmp(e, "0") != 0;
    if (const char* e = std::getenv("IPPI_IB_WAIT_SKEW")) h->ink_wait_skew = std::strcmp(e, "1") == 0;
    if (const char* e = std::getenv("IPPI_FP64_ROUND_STORE")) h->round32 = std::strcmp(e, "1") == 0;
    if (const char* e = std::getenv("IPPI_ADJ_OVERLAP")) h->adj_overlap = std::strcmp(e, "0") != 0;
    if (const char* e = std::getenv("IPPI_ADJ_FUSED")) h->adj_fused = std::strcmp(e, "0") != 0;
    if (const char* e = std::getenv("IPPI_BNN_MFMA")) h->bnn_mfma = std::strcmp(e, "0") != 0;
    if (const char* e = std::getenv("IPPI_NN_NPX")) h->npx_nn = std::strcmp(e, "1") == 0 ? 1 : 2;
    if (const char* e = std::getenv("IPPI_F64_ROWS")) h->rows64 = std::strcmp(e, "1") == 0 ? 1 : std::strcmp(e, "4") == 0 ? 4 : 0;
    h->N = (size_t)cfg->nx * cfg->ny;
    h->esz = cfg->precision == IPPI_FP64 ? 8 : 4;
    // reference state of the FP32 storage: the free stream (u_in_x, 0)
    h->ref.U = cfg->u_in_x;
    h->ref.Uf = (float)cfg->u_in_x;
    ref::feq(1.0, h->ref.U, 0.0, h->ref.c);
    for (int k = 0; k < 9; k++) h->ref.cf[k] = (float)h->ref.c[k];
    h->ref.bb = (float)(h->ref.c[7] - h->ref.c[5]);
    auto bail = [&](int code) {
        ippi_destroy(h);
        return code;
    };
    if (hipSetDevice(cfg->device) != hipSuccess) return bail(IPPI_ERR_HIP);
    if (hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking) != hipSuccess) return bail(IPPI_ERR_HIP);
    for (int i = 0; i < 2; i++)
        if (hipMalloc(&h->buf[i], h->esz * 9 * h->N) != hipSuccess) return bail(IPPI_ERR_NOMEM);
    if (hipMalloc(&h->mac, sizeof(double) * 3 * h->N) != hipSuccess) return bail(IPPI_ERR_NOMEM);
    if (hipMalloc(&h->corner, sizeof(double) * 18) != hipSuccess) return bail(IPPI_ERR_NOMEM);
    if (hipMemset(h->corner, 0, sizeof(double) * 18) != hipSuccess) return bail(IPPI_ERR_HIP);
    if (hipMalloc(&h->flag, sizeof(int)) != hipSuccess) return bail(IPPI_ERR_NOMEM);
    if (hipMalloc(&h->part, sizeof(double) * 1024) != hipSuccess) return bail(IPPI_ERR_NOMEM);
    if (hipMemset(h->flag, 0, sizeof(int)) != hipSuccess) return bail(IPPI_ERR_HIP);
    // the inlet's FP32 equilibrium deviation, evaluated once by the device's
    // arithmetic (the step kernels then read it from their arguments)
    k_qin<<<1, 64, 0, h->s>>>(h->ref, (float)(cfg->u_in_x - h->ref.U), (float)cfg->u_in_y,
                              reinterpret_cast<float*>(h->part));
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(h->qin, h->part, sizeof(h->qin), hipMemcpyDeviceToHost, h->s) != hipSuccess ||
        hipStreamSynchronize(h->s) != hipSuccess)
        return bail(IPPI_ERR_HIP);
    *out = h;
    return IPPI_OK;
}

int ippi_destroy(ippi_handle* h) {
    if (!h) return IPPI_OK;
    if (h->s) (void)hipStreamSynchronize(h->s);
    clear_events(h->ev_step);
    clear_events(h->ev_ib);
    clear_events(h->ev_adj);
    dfree(h->buf[0]); dfree(h->buf[1]); dfree(h->mac); dfree(h->corner); dfree(h->mean); dfree(h->tau); dfree(h->flag); dfree(h->part);
    free_ib(h);
    dfree(h->d_frag);
    dfree(h->d_frage);
    for (auto& w : h->d_bnn_w) dfree(w);
    for (auto& w : h->d_bnn_wm) dfree(w);
    dfree(h->d_bnn_nodes);
    dfree(h->d_bnn_tau);
    dfree(h->d_blob);
    if (h->s) (void)hipStreamDestroy(h->s);
    delete h;
    return IPPI_OK;
}

static int init_common(ippi_handle* h, int mode, uint64_t seed, double rho0, double ux0, double uy0, double amp) {
    dim3 g((unsigned)((h->N + 255) / 256));
    if (h->esz == 8)
        hipLaunchKernelGGL(k_init<double>, g, dim3(256), 0, h->s, (double*)h->buf[0], h->mac, h->N, mode,
                           (unsigned long long)seed, rho0, ux0, uy0, amp, h->ref);
    else
        hipLaunchKernelGGL(k_init<float>, g, dim3(256), 0, h->s, (float*)h->buf[0], h->mac, h->N, mode,
                           (unsigned long long)seed, rho0, ux0, uy0, amp, h->ref);
    HIPC(h, hipGetLastError());
    HIPC(h, hipMemsetAsync(h->flag, 0, sizeof(int), h->s));
    h->cur = 0;
    h->phase = 0;
    h->slots_ok = false;
    h->prev_pull = false;
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


// Restriction of the current post-streaming state to the grid of half the
// resolution (k_restrict): fc (9 Nc, skipped when null) and mc (3 Nc).
static hipError_t launch_restrict(ippi_handle* h, double sc, double* f_out, double* m_out) {
    const int nxc = h->c.nx / 2, nyc = h->c.ny / 2;
    dim3 grid((nxc + BX - 1) / BX, nyc);
    const bool pb = h->c.bc == IPPI_BC_PERIODIC_BB;
    if (h->esz == 8) {
        const double* src = (const double*)h->buf[h->cur];
        if (pb)
            hipLaunchKernelGGL((k_restrict<double, IPPI_BC_PERIODIC_BB>), grid, dim3(BX), 0, h->s, src, f_out, m_out,
                               h->c.nx, h->c.ny, h->c.u_in_x, h->c.u_in_y, sc, f_out ? 1 : 0, h->ref);
        else
            hipLaunchKernelGGL((k_restrict<double, IPPI_BC_INFLOW>), grid, dim3(BX), 0, h->s, src, f_out, m_out,
                               h->c.nx, h->c.ny, h->c.u_in_x, h->c.u_in_y, sc, f_out ? 1 : 0, h->ref);
    } else {
        const float* src = (const float*)h->buf[h->cur];
        if (pb)
            hipLaunchKernelGGL((k_restrict<float, IPPI_BC_PERIODIC_BB>), grid, dim3(BX), 0, h->s, src, f_out, m_out,
                               h->c.nx, h->c.ny, h->c.u_in_x, h->c.u_in_y, sc, f_out ? 1 : 0, h->ref);
        else
            hipLaunchKernelGGL((k_restrict<float, IPPI_BC_INFLOW>), grid, dim3(BX), 0, h->s, src, f_out, m_out,
                               h->c.nx, h->c.ny, h->c.u_in_x, h->c.u_in_y, sc, f_out ? 1 : 0, h->ref);
    }
    return hipGetLastError();
}

// Coarse cells covered by a body's marker support, plus a margin of 2.
static std::vector<int4> pair_exclusions(const ippi_handle* h) {
    std::vector<int4> ex;
    for (const auto& b : h->bodies)
        ex.push_back(make_int4(b.p.x0 / 2 - 2, b.p.y0 / 2 - 2, (b.p.x0 + b.p.w) / 2 + 2, (b.p.y0 + b.p.h) / 2 + 2));
    return ex;
}

// The current state restricted to the grid of half the resolution
// (k_restrict): coarse populations fc [9][ny/2][nx/2] (skipped when null)
// and rho, ux, uy [3][ny/2][nx/2].
int ippi_restrict(ippi_handle* h, double tau_coarse, double* fc, double* mc) {
    if (!h || !mc) return IPPI_ERR_ARG;
    if (h->c.nx % 2 || h->c.ny % 2 || h->c.nx < 16 || h->c.ny < 16) return fail(h, IPPI_ERR_ARG, "grid must be even and >= 16");
    if (h->c.bc == IPPI_BC_FLATPLATE)
        return fail(h, IPPI_ERR_ARG, "the restriction is built for the periodic and inflow boundary kinds");
    if (!(tau_coarse > 0.5)) return IPPI_ERR_ARG;
    if (h->phase == 0) {
        int r = ippi_step(h, 1);
        if (r) return r;
    }
    const size_t Nc = (size_t)(h->c.nx / 2) * (h->c.ny / 2);
    double *dfc = nullptr, *dmc = nullptr;
    hipError_t e = hipMalloc(&dmc, sizeof(double) * 3 * Nc);
    if (e == hipSuccess && fc) e = hipMalloc(&dfc, sizeof(double) * 9 * Nc);
    if (e == hipSuccess) e = launch_restrict(h, 2.0 * tau_coarse / h->c.tau0, dfc, dmc);
    if (e == hipSuccess && fc) e = hipMemcpyAsync(fc, dfc, sizeof(double) * 9 * Nc, hipMemcpyDeviceToHost, h->s);
    if (e == hipSuccess) e = hipMemcpyAsync(mc, dmc, sizeof(double) * 3 * Nc, hipMemcpyDeviceToHost, h->s);
    if (e == hipSuccess) e = hipStreamSynchronize(h->s);
    if (dfc) (void)hipFree(dfc);
    if (dmc) (void)hipFree(dmc);
    if (e != hipSuccess) return fail(h, IPPI_ERR_HIP, std::string("ippi_restrict: ") + hipGetErrorString(e));
    return check_flag(h);
}

// Training pairs at half this solver's resolution (see k_restrict / k_labels).
int ippi_coarse_pairs(ippi_handle* h, double tau_coarse, int coarse_collision, float* feat, float* label,
                      float* weight) {
    if (!h || !feat || !label || !weight) return IPPI_ERR_ARG;
    if (h->c.nx % 2 || h->c.ny % 2 || h->c.nx < 16 || h->c.ny < 16) return fail(h, IPPI_ERR_ARG, "grid must be even and >= 16");
    if (coarse_collision != IPPI_COLL_BGK && coarse_collision != IPPI_COLL_RR)
        return fail(h, IPPI_ERR_ARG, "labels are defined for BGK and RR coarse collisions");
    if (h->c.bc == IPPI_BC_FLATPLATE)
        return fail(h, IPPI_ERR_ARG, "training pairs are built for the periodic and inflow boundary kinds");
    if (!(tau_coarse > 0.5)) return IPPI_ERR_ARG;
    if (h->phase == 0) {
        int r = ippi_step(h, 1);
        if (r) return r;
    }
    const int nxc = h->c.nx / 2, nyc = h->c.ny / 2;
    const size_t Nc = (size_t)nxc * nyc;
    double *fc = nullptr, *mc = nullptr, *tg = nullptr;
    float *df = nullptr, *dl = nullptr, *dw = nullptr;
    int4* dex = nullptr;
    const std::vector<int4> ex = pair_exclusions(h);
    int rc = IPPI_OK;
    auto cleanup = [&]() {
        for (void* q : {(void*)fc, (void*)mc, (void*)tg, (void*)df, (void*)dl, (void*)dw, (void*)dex})
            if (q) (void)hipFree(q);
    };
#define PCK(expr)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) {                                                                \
            cleanup();                                                                         \
            return fail(h, IPPI_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));   \
        }                                                                                      \
    } while (0)
    PCK(hipMalloc(&fc, sizeof(double) * 9 * Nc));
    PCK(hipMalloc(&mc, sizeof(double) * 3 * Nc));
    PCK(hipMalloc(&tg, sizeof(double) * 3 * Nc));
    PCK(hipMalloc(&df, sizeof(float) * 6 * Nc));
    PCK(hipMalloc(&dl, sizeof(float) * Nc));
    PCK(hipMalloc(&dw, sizeof(float) * Nc));
    PCK(hipMalloc(&dex, sizeof(int4) * std::max<size_t>(1, ex.size())));
    if (!ex.empty()) PCK(hipMemcpy(dex, ex.data(), sizeof(int4) * ex.size(), hipMemcpyHostToDevice));
    const double sc = 2.0 * tau_coarse / h->c.tau0;
    dim3 grid((nxc + BX - 1) / BX, nyc);
    auto restrict_ = [&](double* f_out, double* m_out) { return launch_restrict(h, sc, f_out, m_out); };
    PCK(restrict_(fc, mc));
    rc = ippi_step(h, 2);  // one coarse step = two fine steps
    if (rc) {
        cleanup();
        return rc;
    }
    PCK(restrict_(nullptr, tg));
    hipLaunchKernelGGL(k_labels, grid, dim3(BX), 0, h->s, (const double*)fc, (const double*)mc, (const double*)tg, nxc,
                       nyc, coarse_collision == IPPI_COLL_RR ? 1 : 0, (const int4*)dex, (int)ex.size(), df, dl, dw);
    PCK(hipGetLastError());
    PCK(hipMemcpyAsync(feat, df, sizeof(float) * 6 * Nc, hipMemcpyDeviceToHost, h->s));
    PCK(hipMemcpyAsync(label, dl, sizeof(float) * Nc, hipMemcpyDeviceToHost, h->s));
    PCK(hipMemcpyAsync(weight, dw, sizeof(float) * Nc, hipMemcpyDeviceToHost, h->s));
    PCK(hipStreamSynchronize(h->s));
#undef PCK
    cleanup();
    return check_flag(h);
}

// Populations of the post-streaming state, [k][y][x] (FP32 path: stored as
// f - w); moments from CalMacro (NodeAndLattice.cpp:196-226).
int ippi_set_populations(ippi_handle* h, const double* f) {
    if (!h || !f) return IPPI_ERR_ARG;
    std::vector<double> m(3 * h->N);
    for (size_t n = 0; n < h->N; n++) {
        double q[9], r, ux, uy;
        for (int k = 0; k < 9; k++) q[k] = f[k * h->N + n];
        ref::macro(q, r, ux, uy);
        m[n] = r;
        m[h->N + n] = ux;
        m[2 * h->N + n] = uy;
    }
    HIPC(h, hipStreamSynchronize(h->s));
    if (h->esz == 8) {
        HIPC(h, hipMemcpy(h->buf[0], f, sizeof(double) * 9 * h->N, hipMemcpyHostToDevice));
    } else {
        std::vector<float> g(9 * h->N);
        for (size_t k = 0; k < 9; k++)
            for (size_t n = 0; n < h->N; n++) g[k * h->N + n] = (float)(f[k * h->N + n] - h->ref.c[k]);
        HIPC(h, hipMemcpy(h->buf[0], g.data(), sizeof(float) * 9 * h->N, hipMemcpyHostToDevice));
    }
    HIPC(h, hipMemcpy(h->mac, m.data(), sizeof(double) * 3 * h->N, hipMemcpyHostToDevice));
    HIPC(h, hipMemset(h->flag, 0, sizeof(int)));
    h->cur = 0;
    h->phase = 0;
    h->slots_ok = false;
    h->prev_pull = false;
    h->steps = 0;
    h->nrec = 0;
    return IPPI_OK;
}

int ippi_add_body(ippi_handle* h, double cx, double cy, double diameter, int M, double udx, double udy, int* body) {
    if (!h || M <= 0 || !(diameter > 0)) return IPPI_ERR_ARG;
    if (M > IB_MAXM) return fail(h, IPPI_ERR_ARG, "at most 1024 markers per body");
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
    b.p = Patch{x0, y0, x1 - x0 + 1, y1 - y0 + 1, 0, 0, 0};
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
        for (size_t n = 0; n < h->N; n++) t[n] = (float)(tau[n] - 0.5);  // FP32 path keeps tau - 1/2
        HIPC(h, hipMemcpy(h->tau, t.data(), sizeof(float) * h->N, hipMemcpyHostToDevice));
    }
    return IPPI_OK;
}

int ippi_set_mlp(ippi_handle* h, const double* blob, int n) {
    if (!h || !blob) return IPPI_ERR_ARG;
    if (n != MLP_BLOB) return fail(h, IPPI_ERR_ARG, "MLP blob must hold 1325 values");
    for (int i = 0; i < n; i++)
        if (!std::isfinite(blob[i])) return fail(h, IPPI_ERR_ARG, "MLP blob holds a non-finite value");
    // per-lane MFMA fragments of the FP32 path (layouts described in mlp.hpp)
    // constants folded as described at MlpFrag
    std::vector<MlpFrag> fr(64);
    const double L = 1.4426950408889634, LN2 = 0.6931471805599453;
    bool exact = true;  // every W1 and W2 entry a bf16 value
    for (int i = OFF_W1; i < OFF_B1; i++) exact = exact && (double)(float)(__bf16)(float)blob[i] == blob[i];
    for (int i = OFF_W2; i < OFF_B2; i++) exact = exact && (double)(float)(__bf16)(float)blob[i] == blob[i];
    auto hi = [](double v) { return (__bf16)(float)v; };
    auto lo = [](double v) { return (__bf16)((float)v - (float)(__bf16)(float)v); };
    for (int l = 0; l < 64; l++) {
        const int g = l >> 4, c = l & 15;
        MlpFrag& F = fr[l];
        for (int mb = 0; mb < 2; mb++) {
            const int o = mb * 16 + c;
            for (int j = 0; j < 4; j++) {
                const int k = 4 * g + j;
                const double w = k < MLP_NIN ? -L * blob[OFF_W1 + o * MLP_NIN + k] : 0.0;
                F.a1h[mb][j] = hi(w);
                F.a1l[mb][j] = lo(w);
            }
            // [W1, p0, p1, W1, p2, 0] for the exact-weight network: the bias
            // -L b1 as three bf16 parts (remainders taken in FP64, so their
            // sum is -L b1 to ~2^-27 relative) against the activations' 1.0
            // entries
            double bp[3];
            {
                double r = -L * blob[OFF_B1 + o];
                for (int q = 0; q < 3; q++) {
                    bp[q] = (double)(float)(__bf16)(float)r;
                    r -= bp[q];
                }
            }
            for (int j = 0; j < 4; j++) {
                const int k = 4 * g + j;
                const double v = k == 6 ? bp[0] : k == 7 ? bp[1] : k == 14 ? bp[2]
                                 : (k % 8) < MLP_NIN ? blob[OFF_W1 + o * MLP_NIN + k % 8] : 0.0;
                F.a1e[mb][j] = (__bf16)(float)v;
            }
            for (int j = 0; j < 8; j++) {
                const int k = (j / 4) * 16 + 4 * g + (j % 4);  // order of the layer-1 outputs in the lanes
                F.a2h[mb][j] = hi(blob[OFF_W2 + o * MLP_NH + k]);
                F.a2l[mb][j] = lo(blob[OFF_W2 + o * MLP_NH + k]);
            }
            for (int r = 0; r < 4; r++) {
                const int oo = mb * 16 + 4 * g + r;
                F.b1[4 * mb + r] = (float)(-L * blob[OFF_B1 + oo]);
                F.b2[4 * mb + r] = (float)(-L * blob[OFF_B2 + oo]);
                F.w3[4 * mb + r] = (float)(-LN2 * blob[OFF_W3 + oo]);
            }
        }
    }
    // chunk-major copy of the exact-weight network's per-lane fragments
    std::vector<u32x4> fe(MLP_FRAGE_CHUNKS * 64);
    for (int l = 0; l < 64; l++) {
        struct {
            bf16x4 a1e[2];
            bf16x8 a2h[2];
            float b2[8], w3[8];
        } e;
        static_assert(sizeof(e) == 16 * MLP_FRAGE_CHUNKS, "fragment layout");
        memcpy(e.a1e, fr[l].a1e, sizeof(e.a1e));
        memcpy(e.a2h, fr[l].a2h, sizeof(e.a2h));
        memcpy(e.b2, fr[l].b2, sizeof(e.b2));
        memcpy(e.w3, fr[l].w3, sizeof(e.w3));
        for (int c = 0; c < MLP_FRAGE_CHUNKS; c++) memcpy(&fe[c * 64 + l], reinterpret_cast<const char*>(&e) + 16 * c, 16);
    }
    if (!h->d_frag) HIPC(h, hipMalloc(&h->d_frag, sizeof(MlpFrag) * 64));
    if (!h->d_frage) HIPC(h, hipMalloc(&h->d_frage, sizeof(u32x4) * fe.size()));
    if (!h->d_blob) HIPC(h, hipMalloc(&h->d_blob, sizeof(double) * MLP_BLOB));
    HIPC(h, hipStreamSynchronize(h->s));
    HIPC(h, hipMemcpy(h->d_frag, fr.data(), sizeof(MlpFrag) * 64, hipMemcpyHostToDevice));
    HIPC(h, hipMemcpy(h->d_frage, fe.data(), sizeof(u32x4) * fe.size(), hipMemcpyHostToDevice));
    HIPC(h, hipMemcpy(h->d_blob, blob, sizeof(double) * MLP_BLOB, hipMemcpyHostToDevice));
    for (int i = 0; i < 6; i++) {
        // the FP32 kernels pass rho - 1 and u - (U, 0) as the first three features
        const double shift = i == 0 ? 1.0 : (i == 1 ? h->ref.U : 0.0);
        h->mlp_mu[i] = (float)(blob[OFF_MU + i] - shift);
        h->mlp_isig[i] = (float)blob[OFF_IS + i];
        h->mlp_isig_e[i] = (float)(-L * blob[OFF_IS + i]);
    }
    h->mlp_b3 = (float)blob[OFF_B3];
    h->mlp_exact = exact;
    h->has_mlp = true;
    return IPPI_OK;
}

int ippi_set_boundary_mlp(ippi_handle* h, int type, const double* blob, int n) {
    if (!h || !blob) return IPPI_ERR_ARG;
    const int t = type == 1 ? 0 : type == 3 ? 1 : type == 5 ? 2 : type == 7 ? 3 : -1;
    if (t < 0) return fail(h, IPPI_ERR_ARG, "boundary network type must be 1, 3, 5 or 7");
    if (n != BNN_NPARAM) return fail(h, IPPI_ERR_ARG, "boundary network blob must hold 101249 values");
    if (h->c.nx < 3 || h->c.ny < 4) return fail(h, IPPI_ERR_ARG, "lattice too small for the boundary networks");
    for (int i = 0; i < n; i++)
        if (!std::isfinite(blob[i])) return fail(h, IPPI_ERR_ARG, "boundary network blob holds a non-finite value");
    // device layout: each weight matrix input-major (W[o][k] -> Wt[k][o]), biases as given
    std::vector<float> w(BNN_NPARAM);
    size_t off = 0;
    auto layer = [&](int nin, int nout) {
        for (int o = 0; o < nout; o++)
            for (int k = 0; k < nin; k++) w[off + (size_t)k * nout + o] = (float)blob[off + (size_t)o * nin + k];
        off += (size_t)nin * nout;
        for (int o = 0; o < nout; o++) w[off + o] = (float)blob[off + o];
        off += nout;
    };
    layer(BNN_NIN, BNN_NH);
    for (int l = 0; l < BNN_NHID; l++) layer(BNN_NH, BNN_NH);
    for (int k = 0; k <= BNN_NH; k++) w[off + k] = (float)blob[off + k];  // output row and bias
    HIPC(h, hipStreamSynchronize(h->s));
    if (!h->d_bnn_w[t]) HIPC(h, hipMalloc(&h->d_bnn_w[t], sizeof(float) * BNN_NPARAM));
    HIPC(h, hipMemcpy(h->d_bnn_w[t], w.data(), sizeof(float) * BNN_NPARAM, hipMemcpyHostToDevice));
    // the MFMA kernel's copy (layout at BNM_OFF_*), from the input-major one
    std::vector<float> wm(BNM_NPARAM, 0.0f);
    const size_t L0 = (size_t)BNN_NIN * BNN_NH + BNN_NH, LH = (size_t)BNN_NH * BNN_NH + BNN_NH;
    for (int wv = 0; wv < 8; wv++)
        for (int ln = 0; ln < 64; ln++) {
            const int kq = ln / 16, o = 16 * wv + ln % 16;
            for (int j = 0; j < 4; j++) {
                const int k = 4 * j + kq;
                wm[4 * (64 * wv + ln) + j] = k < BNN_NIN ? w[(size_t)k * BNN_NH + o] : 0.0f;
            }
            for (int hl = 1; hl <= BNN_NHID; hl++)
                for (int j = 0; j < 32; j++)
                    wm[BNM_OFF_H1 + (size_t)(hl - 1) * BNM_LAYER_H + 4 * ((8 * wv + j / 4) * 64 + ln) + j % 4] =
                        w[L0 + (size_t)(hl - 1) * LH + (size_t)(4 * j + kq) * BNN_NH + o];
        }
    for (int o = 0; o < BNN_NH; o++) {
        wm[BNM_OFF_B0 + o] = w[(size_t)BNN_NIN * BNN_NH + o];
        for (int hl = 1; hl <= BNN_NHID; hl++)
            wm[BNM_OFF_H1 + (size_t)hl * BNM_LAYER_H - BNN_NH + o] = w[L0 + (size_t)(hl - 1) * LH + (size_t)BNN_NH * BNN_NH + o];
    }
    for (int k = 0; k <= BNN_NH; k++) wm[BNM_OFF_OUT + k] = w[L0 + (size_t)BNN_NHID * LH + k];
    if (!h->d_bnn_wm[t]) HIPC(h, hipMalloc(&h->d_bnn_wm[t], sizeof(float) * BNM_NPARAM));
    HIPC(h, hipMemcpy(h->d_bnn_wm[t], wm.data(), sizeof(float) * BNM_NPARAM, hipMemcpyHostToDevice));
    h->bnn_on |= 1 << t;
    // node list of every type set, grouped by type
    const int nx = h->c.nx, ny = h->c.ny;
    std::vector<BnnNode> nodes;
    for (int tt = 0; tt < 4; tt++) {
        if (h->bnn_on & (1 << tt)) {
            if (tt == 0 || tt == 2)
                for (int y = 1; y <= ny - 2; y++)
                    nodes.push_back(BnnNode{tt == 0 ? 0 : nx - 1, y, 2 * nx + (tt == 2 ? ny : 0) + y, tt});
            else
                for (int x = 0; x < nx; x++) nodes.push_back(BnnNode{x, tt == 3 ? 1 : ny - 2, (tt == 1 ? nx : 0) + x, tt});
            BnnNode last = nodes.back();
            last.slot = -1;
            while (nodes.size() % BNN_PAD) nodes.push_back(last);
        }
        h->bnn_tend[tt] = (int)nodes.size();
    }
    dfree(h->d_bnn_nodes);
    HIPC(h, hipMalloc(&h->d_bnn_nodes, sizeof(BnnNode) * nodes.size()));
    HIPC(h, hipMemcpy(h->d_bnn_nodes, nodes.data(), sizeof(BnnNode) * nodes.size(), hipMemcpyHostToDevice));
    h->bnn_nnodes = (int)nodes.size();
    if (!h->d_bnn_tau) HIPC(h, hipMalloc(&h->d_bnn_tau, sizeof(double) * (2 * (size_t)nx + 2 * (size_t)ny)));
    HIPC(h, hipMemset(h->d_bnn_tau, 0, sizeof(double) * (2 * (size_t)nx + 2 * (size_t)ny)));
    return IPPI_OK;
}

int ippi_step(ippi_handle* h, int nsteps) {
    if (!h || nsteps < 0) return IPPI_ERR_ARG;
    if (h->c.tau_mode == IPPI_TAU_FIELD && !h->tau) return fail(h, IPPI_ERR_STATE, "tau field not set");
    if (nn_kind(h) && !h->has_mlp) return fail(h, IPPI_ERR_STATE, "relaxation-time MLP not set");
    return h->esz == 8 ? run_steps<double>(h, nsteps) : run_steps<float>(h, nsteps);
}

int ippi_set_timing(ippi_handle* h, int enable) {
    if (!h) return IPPI_ERR_ARG;
    HIPC(h, hipStreamSynchronize(h->s));
    clear_events(h->ev_step);
    clear_events(h->ev_ib);
    clear_events(h->ev_adj);
    if (enable < 0 || enable > 2) return IPPI_ERR_ARG;
    h->timing = enable;
    h->batch_steps = 0;
    h->batch_ib = 0;
    return IPPI_OK;
}

int ippi_get_timing(ippi_handle* h, double* step_ms, int64_t* step_launches, double* ib_ms, int64_t* ib_launches) {
    if (!h) return IPPI_ERR_ARG;
    HIPC(h, hipStreamSynchronize(h->s));
    int r = sum_events(h, h->ev_step, step_ms, step_launches);
    if (r) return r;
    r = sum_events(h, h->ev_ib, ib_ms, ib_launches);
    if (r) return r;
    if (h->timing == 2) {  // the calls' time over the step kernels they launched
        if (step_launches) *step_launches = h->batch_steps;
        if (ib_launches) *ib_launches = h->batch_ib;
        h->batch_steps = 0;
        h->batch_ib = 0;
    }
    return IPPI_OK;
}

int ippi_pattern_probe(int device, int nx, int ny, int precision, int mode, int reps, double* us_per_launch) {
    if (nx < 4 || ny < 4 || precision < 0 || precision > 1 || mode < 0 || mode > 3 ||
        reps < 1 || !us_per_launch)
        return IPPI_ERR_ARG;
    if (9 * (uint64_t)nx * (uint64_t)ny >= ((uint64_t)1 << 32)) return IPPI_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return IPPI_ERR_HIP;
    const size_t N = (size_t)nx * ny, esz = precision == IPPI_FP64 ? 8 : 4;
    const int nt = 2 * 9 * N * esz > ((size_t)512 << 20);
    void* b[2] = {nullptr, nullptr};
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = IPPI_OK;
    auto chk = [&](hipError_t e, int code) {
        if (e != hipSuccess && rc == IPPI_OK) rc = code;
        return e == hipSuccess;
    };
    if (chk(hipMalloc(&b[0], 9 * N * esz), IPPI_ERR_NOMEM) && chk(hipMalloc(&b[1], 9 * N * esz), IPPI_ERR_NOMEM) &&
        chk(hipMemset(b[0], 0, 9 * N * esz), IPPI_ERR_HIP) && chk(hipMemset(b[1], 0, 9 * N * esz), IPPI_ERR_HIP) &&
        chk(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), IPPI_ERR_HIP) &&
        chk(hipEventCreateWithFlags(&e0, hipEventDisableSystemFence), IPPI_ERR_HIP) &&
        chk(hipEventCreateWithFlags(&e1, hipEventDisableSystemFence), IPPI_ERR_HIP)) {
        const dim3 g((nx + BX - 1) / BX, ny);
        auto launch = [&](int i) {
            const void* s0 = b[i & 1];
            void* d0 = b[(i & 1) ^ 1];
#define IPPI_PROBE(TT, M, NTV) \
    hipLaunchKernelGGL((k_probe<TT, M, NTV>), g, dim3(BX), 0, st, (const TT*)s0, (TT*)d0, nx, ny)
#define IPPI_PROBE_T(TT)                               \
    if (mode == 2 && nt) IPPI_PROBE(TT, 2, true);      \
    else if (mode == 2) IPPI_PROBE(TT, 2, false);      \
    else if (mode == 1 && nt) IPPI_PROBE(TT, 1, true); \
    else if (mode == 1) IPPI_PROBE(TT, 1, false);      \
    else if (nt) IPPI_PROBE(TT, 0, true);              \
    else IPPI_PROBE(TT, 0, false);
            if (mode == 3) {  // 16 bytes per lane, one flat array of 9 N elements
                const size_t n4 = 9 * N * esz / 16;
                const dim3 gc((unsigned)std::min<size_t>((n4 + 255) / 256, 65536));
                if (nt)
                    hipLaunchKernelGGL(k_copy16<true>, gc, dim3(256), 0, st, (const f32x4v*)s0, (f32x4v*)d0, n4);
                else
                    hipLaunchKernelGGL(k_copy16<false>, gc, dim3(256), 0, st, (const f32x4v*)s0, (f32x4v*)d0, n4);
            } else if (esz == 8) {
                IPPI_PROBE_T(double)
            } else {
                IPPI_PROBE_T(float)
            }
#undef IPPI_PROBE_T
#undef IPPI_PROBE
        };
        for (int i = 0; i < 10; i++) launch(i);  // warm-up
        chk(hipEventRecord(e0, st), IPPI_ERR_HIP);
        for (int i = 0; i < reps; i++) launch(i);
        chk(hipEventRecord(e1, st), IPPI_ERR_HIP);
        chk(hipGetLastError(), IPPI_ERR_HIP);
        float ms = 0.f;
        if (chk(hipEventSynchronize(e1), IPPI_ERR_HIP) && chk(hipEventElapsedTime(&ms, e0, e1), IPPI_ERR_HIP))
            *us_per_launch = 1e3 * (double)ms / reps;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    for (void* q : b)
        if (q) (void)hipFree(q);
    return rc;
}

int ippi_get_adjoint_timing(ippi_handle* h, double* ms, int64_t* launches) {
    if (!h) return IPPI_ERR_ARG;
    HIPC(h, hipStreamSynchronize(h->s));
    return sum_events(h, h->ev_adj, ms, launches);
}

int ippi_adjoint_drag(ippi_handle* h, int nsteps, int avg_from, int ckpt, int flags, const double* body_w,
                      double* J, double* g_tau, double* g_center) {
    if (!h || nsteps < 1 || avg_from < 1 || avg_from > nsteps || ckpt < 1 || (flags & ~IPPI_ADJ_FP64))
        return IPPI_ERR_ARG;
    if (h->c.collision != IPPI_COLL_BGK || h->c.bc != IPPI_BC_PERIODIC_BB || h->c.les || nn_kind(h) || h->bnn_on)
        return fail(h, IPPI_ERR_ARG, "the adjoint covers BGK, periodic x / bounce-back walls, no LES, "
                                     "constant or per-node tau");
    if (h->bodies.empty()) return fail(h, IPPI_ERR_STATE, "the drag objective needs an immersed body");
    if (h->c.tau_mode == IPPI_TAU_FIELD && !h->tau) return fail(h, IPPI_ERR_STATE, "tau field not set");
    int r;
    if (h->esz == 8)
        r = adjoint_run<double, double>(h, nsteps, avg_from, ckpt, body_w, J, g_tau, g_center);
    else if (flags & IPPI_ADJ_FP64)
        r = adjoint_run<float, double>(h, nsteps, avg_from, ckpt, body_w, J, g_tau, g_center);
    else
        r = adjoint_run<float, float>(h, nsteps, avg_from, ckpt, body_w, J, g_tau, g_center);
    h->slots_ok = false;  // the recomputation ran step kernels that write no slots
    h->prev_pull = false;
    return r;
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

int ippi_accumulate_mean(ippi_handle* h, int64_t k) {
    if (!h || k < 0) return IPPI_ERR_ARG;
    if (!h->mean) {
        HIPC(h, hipMalloc(&h->mean, sizeof(double) * 3 * h->N));
        HIPC(h, hipMemsetAsync(h->mean, 0, sizeof(double) * 3 * h->N, h->s));
    }
    int r = h->esz == 8 ? run_macro<double>(h, nullptr, h->mean, (double)k)
                        : run_macro<float>(h, nullptr, h->mean, (double)k);
    if (r) return r;
    return IPPI_OK;
}

int ippi_get_mean(ippi_handle* h, double* rho, double* ux, double* uy) {
    if (!h || !rho || !ux || !uy) return IPPI_ERR_ARG;
    if (!h->mean) return fail(h, IPPI_ERR_STATE, "no running mean accumulated");
    const size_t B = sizeof(double) * h->N;
    HIPC(h, hipMemcpyAsync(rho, h->mean, B, hipMemcpyDeviceToHost, h->s));
    HIPC(h, hipMemcpyAsync(ux, h->mean + h->N, B, hipMemcpyDeviceToHost, h->s));
    HIPC(h, hipMemcpyAsync(uy, h->mean + 2 * h->N, B, hipMemcpyDeviceToHost, h->s));
    HIPC(h, hipStreamSynchronize(h->s));
    return check_flag(h);
}

int ippi_get_stencil(ippi_handle* h, int n, const int* xy, double* fnew, double* feq, double* tau) {
    if (!h || n < 0 || n > INT_MAX / 9 || (n > 0 && (!xy || !fnew || !feq || !tau))) return IPPI_ERR_ARG;
    if (nn_kind(h))
        return fail(h, IPPI_ERR_ARG, "the stencil covers constant or per-node tau (with or without LES)");
    if (h->c.tau_mode == IPPI_TAU_FIELD && !h->tau) return fail(h, IPPI_ERR_STATE, "tau field not set");
    if (!h->prev_pull) return fail(h, IPPI_ERR_STATE, "no pull step since the state was set");
    for (int i = 0; i < n; i++)
        if (xy[2 * i] < 2 || xy[2 * i] > h->c.nx - 3 || xy[2 * i + 1] < 2 || xy[2 * i + 1] > h->c.ny - 3)
            return fail(h, IPPI_ERR_ARG, "stencil nodes must lie two or more nodes inside the domain");
    if (n == 0) return IPPI_OK;
    StepArgs a = make_args(h);
    int* d_xy = nullptr;
    double* d_o = nullptr;
    const size_t s9 = sizeof(double) * 9 * (size_t)n;
    hipError_t e = hipMalloc(&d_xy, sizeof(int) * 2 * (size_t)n);
    if (e == hipSuccess) e = hipMalloc(&d_o, 3 * s9);
    if (e == hipSuccess) e = hipMemcpyAsync(d_xy, xy, sizeof(int) * 2 * (size_t)n, hipMemcpyHostToDevice, h->s);
    if (e == hipSuccess) {
        const dim3 g((unsigned)((9 * (size_t)n + 255) / 256));
        double* o = d_o;
        if (h->esz == 8)
            hipLaunchKernelGGL(k_stencil<double>, g, dim3(256), 0, h->s, (const double*)h->buf[h->cur ^ 1], d_xy, n,
                               a, o, o + 9 * (size_t)n, o + 18 * (size_t)n);
        else
            hipLaunchKernelGGL(k_stencil<float>, g, dim3(256), 0, h->s, (const float*)h->buf[h->cur ^ 1], d_xy, n, a,
                               o, o + 9 * (size_t)n, o + 18 * (size_t)n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(fnew, d_o, s9, hipMemcpyDeviceToHost, h->s);
    if (e == hipSuccess) e = hipMemcpyAsync(feq, d_o + 9 * (size_t)n, s9, hipMemcpyDeviceToHost, h->s);
    if (e == hipSuccess) e = hipMemcpyAsync(tau, d_o + 18 * (size_t)n, s9, hipMemcpyDeviceToHost, h->s);
    if (e == hipSuccess) e = hipStreamSynchronize(h->s);
    dfree(d_xy);
    dfree(d_o);
    if (e != hipSuccess) return fail(h, IPPI_ERR_HIP, std::string("ippi_get_stencil: ") + hipGetErrorString(e));
    return IPPI_OK;
}

// The relaxation time of the next collision at every node into d_out (device,
// N doubles), queued on the solver's stream.
static int launch_tau(ippi_handle* h, double* d_out) {
    if (h->phase == 0) return fail(h, IPPI_ERR_STATE, "no step taken yet");
    if (h->c.tau_mode == IPPI_TAU_FIELD && !h->tau) return fail(h, IPPI_ERR_STATE, "tau field not set");
    if (nn_kind(h) && !h->has_mlp) return fail(h, IPPI_ERR_STATE, "relaxation-time MLP not set");
    StepArgs a = make_args(h);
    if (h->bnn_on) {
        int r = h->esz == 8 ? launch_bnn<double>(h, a, true) : launch_bnn<float>(h, a, true);
        if (r) return r;
    }
    dim3 grid((h->c.nx + BX - 1) / BX, h->c.ny);
    const int nk = nn_kind(h);
    const bool nn = nk != 0;
    const int bc = h->c.bc;
#define IPPI_TAU_LAUNCH(K, T, B)                                                                          \
    if (bc == B) {                                                                                        \
        if (nn)                                                                                           \
            hipLaunchKernelGGL((K<B, true>), grid, dim3(BX), 0, h->s, (const T*)h->buf[h->cur], d_out, a);  \
        else                                                                                              \
            hipLaunchKernelGGL((K<B, false>), grid, dim3(BX), 0, h->s, (const T*)h->buf[h->cur], d_out, a); \
    }
#define IPPI_TAU32_LAUNCH(B)                                                                               \
    if (bc == B) {                                                                                         \
        if (nk == 3)                                                                                       \
            hipLaunchKernelGGL((k_tau32<B, 3>), grid, dim3(BX), 0, h->s, (const float*)h->buf[h->cur], d_out, a); \
        else if (nk == 2)                                                                                  \
            hipLaunchKernelGGL((k_tau32<B, 2>), grid, dim3(BX), 0, h->s, (const float*)h->buf[h->cur], d_out, a); \
        else if (nk == 1)                                                                                  \
            hipLaunchKernelGGL((k_tau32<B, 1>), grid, dim3(BX), 0, h->s, (const float*)h->buf[h->cur], d_out, a); \
        else                                                                                               \
            hipLaunchKernelGGL((k_tau32<B, 0>), grid, dim3(BX), 0, h->s, (const float*)h->buf[h->cur], d_out, a); \
    }
    if (h->esz == 4) {
        IPPI_TAU32_LAUNCH(IPPI_BC_PERIODIC_BB)
        IPPI_TAU32_LAUNCH(IPPI_BC_INFLOW)
        IPPI_TAU32_LAUNCH(IPPI_BC_FLATPLATE)
#undef IPPI_TAU32_LAUNCH
    } else {
        IPPI_TAU_LAUNCH(k_tau64, double, IPPI_BC_PERIODIC_BB)
        IPPI_TAU_LAUNCH(k_tau64, double, IPPI_BC_INFLOW)
        IPPI_TAU_LAUNCH(k_tau64, double, IPPI_BC_FLATPLATE)
    }
#undef IPPI_TAU_LAUNCH
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(h, IPPI_ERR_HIP, std::string("relaxation-time kernel: ") + hipGetErrorString(e));
    return IPPI_OK;
}

int ippi_get_tau(ippi_handle* h, double* tau) {
    if (!h || !tau) return IPPI_ERR_ARG;
    double* d_out = nullptr;
    HIPC(h, hipMalloc(&d_out, sizeof(double) * h->N));
    int r = launch_tau(h, d_out);
    if (r) {
        (void)hipFree(d_out);
        return r;
    }
    hipError_t e = hipMemcpyAsync(tau, d_out, sizeof(double) * h->N, hipMemcpyDeviceToHost, h->s);
    if (e == hipSuccess) e = hipStreamSynchronize(h->s);
    (void)hipFree(d_out);
    if (e != hipSuccess) return fail(h, IPPI_ERR_HIP, std::string("ippi_get_tau: ") + hipGetErrorString(e));
    return check_flag(h);
}

// Effective-relaxation-time pairs at half this solver's resolution (see
// k_stress_labels); the solver does not advance.
int ippi_coarse_stress_pairs(ippi_handle* h, double tau_coarse, float* feat, float* label, float* weight) {
    if (!h || !feat || !label || !weight) return IPPI_ERR_ARG;
    if (h->c.nx % 2 || h->c.ny % 2 || h->c.nx < 16 || h->c.ny < 16) return fail(h, IPPI_ERR_ARG, "grid must be even and >= 16");
    if (h->c.bc == IPPI_BC_FLATPLATE)
        return fail(h, IPPI_ERR_ARG, "training pairs are built for the periodic and inflow boundary kinds");
    if (!(tau_coarse > 0.5)) return IPPI_ERR_ARG;
    if (h->phase == 0) {
        int r = ippi_step(h, 1);
        if (r) return r;
    }
    const int nxc = h->c.nx / 2, nyc = h->c.ny / 2;
    const size_t Nc = (size_t)nxc * nyc;
    double *fc = nullptr, *mc = nullptr, *tf = nullptr;
    float *df = nullptr, *dl = nullptr, *dw = nullptr;
    int4* dex = nullptr;
    const std::vector<int4> ex = pair_exclusions(h);
    auto cleanup = [&]() {
        for (void* q : {(void*)fc, (void*)mc, (void*)tf, (void*)df, (void*)dl, (void*)dw, (void*)dex})
            if (q) (void)hipFree(q);
    };
#define PCK(expr)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) {                                                                \
            cleanup();                                                                         \
            return fail(h, IPPI_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));   \
        }                                                                                      \
    } while (0)
    PCK(hipMalloc(&fc, sizeof(double) * 9 * Nc));
    PCK(hipMalloc(&mc, sizeof(double) * 3 * Nc));
    PCK(hipMalloc(&tf, sizeof(double) * h->N));
    PCK(hipMalloc(&df, sizeof(float) * 6 * Nc));
    PCK(hipMalloc(&dl, sizeof(float) * Nc));
    PCK(hipMalloc(&dw, sizeof(float) * Nc));
    PCK(hipMalloc(&dex, sizeof(int4) * std::max<size_t>(1, ex.size())));
    if (!ex.empty()) PCK(hipMemcpy(dex, ex.data(), sizeof(int4) * ex.size(), hipMemcpyHostToDevice));
    PCK(launch_restrict(h, 2.0 * tau_coarse / h->c.tau0, fc, mc));
    int rc = launch_tau(h, tf);
    if (rc) {
        cleanup();
        return rc;
    }
    dim3 grid((nxc + BX - 1) / BX, nyc);
    const bool pb = h->c.bc == IPPI_BC_PERIODIC_BB;
#define IPPI_SL_LAUNCH(T, B)                                                                                   \
    hipLaunchKernelGGL((k_stress_labels<T, B>), grid, dim3(BX), 0, h->s, (const T*)h->buf[h->cur],             \
                       (const double*)tf, (const double*)fc, (const double*)mc, h->c.nx, h->c.ny, h->c.u_in_x, \
                       h->c.u_in_y, h->ref, (const int4*)dex, (int)ex.size(), df, dl, dw)
    if (h->esz == 8) {
        if (pb)
            IPPI_SL_LAUNCH(double, IPPI_BC_PERIODIC_BB);
        else
            IPPI_SL_LAUNCH(double, IPPI_BC_INFLOW);
    } else {
        if (pb)
            IPPI_SL_LAUNCH(float, IPPI_BC_PERIODIC_BB);
        else
            IPPI_SL_LAUNCH(float, IPPI_BC_INFLOW);
    }
#undef IPPI_SL_LAUNCH
    PCK(hipGetLastError());
    PCK(hipMemcpyAsync(feat, df, sizeof(float) * 6 * Nc, hipMemcpyDeviceToHost, h->s));
    PCK(hipMemcpyAsync(label, dl, sizeof(float) * Nc, hipMemcpyDeviceToHost, h->s));
    PCK(hipMemcpyAsync(weight, dw, sizeof(float) * Nc, hipMemcpyDeviceToHost, h->s));
    PCK(hipStreamSynchronize(h->s));
#undef PCK
    cleanup();
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
    const size_t need = sizeof(SaveHdr) + h->esz * 9 * h->N + sizeof(double) * (3 * h->N + 18);
    if (!buf) {
        *size = need;
        return IPPI_OK;
    }
    if (*size < need) return fail(h, IPPI_ERR_ARG, "checkpoint buffer too small");
    SaveHdr hd{0x4950504932ull, h->c.nx, h->c.ny, h->c.precision, h->phase, h->steps, h->nrec};
    char* p = (char*)buf;
    std::memcpy(p, &hd, sizeof(hd));
    HIPC(h, hipStreamSynchronize(h->s));
    HIPC(h, hipMemcpy(p + sizeof(hd), h->buf[h->cur], h->esz * 9 * h->N, hipMemcpyDeviceToHost));
    HIPC(h, hipMemcpy(p + sizeof(hd) + h->esz * 9 * h->N, h->mac, sizeof(double) * 3 * h->N, hipMemcpyDeviceToHost));
    HIPC(h, hipMemcpy(p + sizeof(hd) + h->esz * 9 * h->N + sizeof(double) * 3 * h->N, h->corner, sizeof(double) * 18,
                      hipMemcpyDeviceToHost));
    *size = need;
    return IPPI_OK;
}

int ippi_load(ippi_handle* h, const void* buf, size_t size) {
    if (!h || !buf) return IPPI_ERR_ARG;
    const size_t need = sizeof(SaveHdr) + h->esz * 9 * h->N + sizeof(double) * (3 * h->N + 18);
    if (size < need) return fail(h, IPPI_ERR_ARG, "checkpoint too small");
    SaveHdr hd;
    std::memcpy(&hd, buf, sizeof(hd));
    if (hd.magic != 0x4950504932ull || hd.nx != h->c.nx || hd.ny != h->c.ny || hd.precision != h->c.precision)
        return fail(h, IPPI_ERR_ARG, "checkpoint does not match this configuration");
    const char* p = (const char*)buf;
    HIPC(h, hipStreamSynchronize(h->s));
    h->cur = 0;
    HIPC(h, hipMemcpy(h->buf[0], p + sizeof(hd), h->esz * 9 * h->N, hipMemcpyHostToDevice));
    HIPC(h, hipMemcpy(h->mac, p + sizeof(hd) + h->esz * 9 * h->N, sizeof(double) * 3 * h->N, hipMemcpyHostToDevice));
    HIPC(h, hipMemcpy(h->corner, p + sizeof(hd) + h->esz * 9 * h->N + sizeof(double) * 3 * h->N, sizeof(double) * 18,
                      hipMemcpyHostToDevice));
    HIPC(h, hipMemset(h->flag, 0, sizeof(int)));
    h->phase = hd.phase;
    h->steps = hd.steps;
    h->slots_ok = false;
    h->prev_pull = false;
    return IPPI_OK;
}

}  // extern "C"
