/*
 * ippi_lbm.h -- C ABI of the MI355X (gfx950) D2Q9 time step of IPPI-LBM.
 *
 * The library (libippi_lbm.so) owns the whole domain on one GPU: SoA
 * populations f[k][y][x] in HBM (A/B buffers), immersed-boundary markers,
 * optional per-node relaxation-time field and the fused relaxation-time
 * MLP. Every call returns an int status (IPPI_OK = 0); no C++ exception
 * crosses this boundary. A handle is not thread-safe; all work of a handle
 * is ordered on one HIP stream. ippi_step() only enqueues work: errors that
 * happen on the device (divergence) are reported by the next synchronising
 * call (ippi_get_*, ippi_mass, ippi_synchronize).
 *
 * Host arrays are copied, never retained. Field arrays are row-major
 * [y][x] with nx*ny doubles; population arrays are [k][y][x] with
 * 9*nx*ny doubles, k in the reference's velocity order
 * (IPPI-LBM/NodeAndLattice.cpp:3-13): (0,0),(1,0),(0,1),(-1,0),(0,-1),
 * (1,1),(-1,1),(-1,-1),(1,-1).
 *
 * Reference interface each entry point replaces (IPPI-LBM/ paths):
 *   ippi_create        Mesh::InitialLatMesh          mesh.h:56, mesh.cpp:60-171
 *                      + JudgeStyle boundary kinds   mesh.cpp:1391-1454
 *   ippi_init_fields   InitialLatMesh's rho/u/feq    mesh.cpp:88-167
 *   ippi_set_fields    (direct writes of LatMesh[i][j].rho/u + InitialLatticeFeq)
 *   ippi_add_body      Mesh::InitialIBParticle       mesh.h:57, mesh.cpp:173-191
 *   ippi_set_tau_field (writes of LatMesh[i][j].tau / TauAtBeginning, mesh.cpp:90-91)
 *   ippi_set_mlp       NNmodel_Type00..07 model load mesh.cpp:2767-3260
 *   ippi_set_boundary_mlp  NNmodel_Type01/03/05/07    mesh.cpp:2767-2898, 3128-3260
 *   ippi_step          Mesh::OneLBMStep              mesh.h:60, mesh.cpp:357-513
 *                      (+ InterpolateVel / CalParticleForce / SpreadForce /
 *                       UpdateVelocity, mesh.cpp:514-832,1244-1281)
 *   ippi_get_fields    reads of LatMesh[i][j].rho/u  (WriteDatFluid, mesh.cpp:1180-1243)
 *   ippi_accumulate_mean / ippi_get_mean
 *                      Mesh::ComputeAvgVelocity      mesh.cpp:1954-1974 (avg_u/avg_v/avg_density)
 *   ippi_get_body_forces  OutputCd                   mesh.cpp:1461-1526
 *   ippi_mass          ErrorCatch / global checks    mesh.cpp:2101-2111
 *   ippi_destroy       Mesh::~Mesh                   mesh.cpp:52-55
 *   ippi_adjoint_drag  (none: gradient of the averaged drag, see its comment)
 */
#ifndef IPPI_LBM_H
#define IPPI_LBM_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define IPPI_OK 0
#define IPPI_ERR_ARG 1       /* invalid argument or configuration */
#define IPPI_ERR_HIP 2       /* HIP runtime error (message via ippi_last_error) */
#define IPPI_ERR_DIVERGED 3  /* a node's density left [0, 10] or became NaN (ErrorCatch) */
#define IPPI_ERR_STATE 4     /* call not valid in the handle's current state */
#define IPPI_ERR_NOMEM 5     /* device allocation failed */
#define IPPI_ERR_CAPACITY 6  /* input larger than a fixed capacity of the call */

enum ippi_collision { IPPI_COLL_BGK = 0, IPPI_COLL_MRT = 1, IPPI_COLL_RR = 2 };
enum ippi_precision { IPPI_FP64 = 0, IPPI_FP32_SHIFTED = 1 };
enum ippi_bc { IPPI_BC_PERIODIC_BB = 0, IPPI_BC_INFLOW = 1, IPPI_BC_FLATPLATE = 2 };
/* Boundary kinds:
 *   IPPI_BC_PERIODIC_BB  periodic x (the reference's ring halo), halfway
 *                        bounce-back walls at y = 0 and y = ny-1.
 *   IPPI_BC_INFLOW       NEE inlet (rho = 1, u = u_in) and outlet (rho = 1,
 *                        u copied from x = nx-2) at x = 0 and nx-1
 *                        (mesh.cpp:239-286); specular free slip at both walls
 *                        (the 'm' rule, mesh.cpp:325-337, mirrored on top).
 *   IPPI_BC_FLATPLATE    the reference's own set (Mesh::BoundaryCondition,
 *                        mesh.cpp:228-356, styles mesh.cpp:1391-1454), x being
 *                        the reference's global column: NEE inlet and outlet on
 *                        1 <= y <= ny-2; NEE top (rho = 1, u from y = ny-2);
 *                        bottom free slip ('m') for x < plate_x and NEE no-slip
 *                        ('d': rho from y = 1, u = 0) beyond; the corners
 *                        (0, 0) and (nx-1, 0), which no rule reaches, keep
 *                        their state. */
/* IPPI_TAU_NN_BF16: the network with single bf16 products (no hi/lo
 * compensation) and FP32 accumulation in the FP32 path, a third of the
 * MFMAs and about 8 mantissa bits per operand; no counterpart in the
 * reference (a speed/accuracy option of the fused network). In FP64 it is
 * IPPI_TAU_NN. */
enum ippi_tau_mode { IPPI_TAU_CONST = 0, IPPI_TAU_FIELD = 1, IPPI_TAU_NN = 2, IPPI_TAU_NN_BF16 = 3 };

typedef struct ippi_config {
    int nx, ny;          /* whole domain (no ghost columns) */
    int collision;       /* ippi_collision */
    int precision;       /* ippi_precision: FP64 populations, or FP32 deviations f - feq(1, (u_in_x, 0)) */
    int bc;              /* ippi_bc */
    int les;             /* 1: hybrid Smagorinsky tau (mesh.cpp:877-910) */
    int tau_mode;        /* ippi_tau_mode */
    double tau0;         /* startTau = TauAtBeginning */
    double cs;           /* Smagorinsky constant Cs */
    double u_in_x;       /* inlet velocity for IPPI_BC_INFLOW and IPPI_BC_FLATPLATE */
    double u_in_y;
    int device;          /* HIP device ordinal */
    int force_history;   /* body-force records kept (ring); 0 -> 65536 */
    double tau_nn_max;   /* IPPI_TAU_NN: network values above this are rejected
                            like those below 0.501 (0 = no upper bound) */
    int plate_x;         /* IPPI_BC_FLATPLATE: first no-slip column of the bottom
                            wall (the reference hard-codes 400, mesh.cpp:1413) */
} ippi_config;

typedef struct ippi_handle ippi_handle;

int ippi_create(const ippi_config* cfg, ippi_handle** out);
int ippi_destroy(ippi_handle* h);
const char* ippi_last_error(const ippi_handle* h);
int ippi_version(void);

/* rho = rho0, u = (ux0, uy0) + noise * U(-1,1) per node and component
 * (splitmix64 keyed by seed, node, component), f = feq(rho, u). */
int ippi_init_fields(ippi_handle* h, uint64_t seed, double rho0, double ux0, double uy0, double noise);
/* rho/u given per node ([y][x]); f = feq(rho, u). */
int ippi_set_fields(ippi_handle* h, const double* rho, const double* ux, const double* uy);

/* Post-streaming populations [k][y][x] (9*nx*ny doubles); moments from
 * CalMacro. Used to start a run from a given distribution (e.g. a restricted
 * finer solution). Replaces direct writes of LatMesh[i][j].fnew. */
int ippi_set_populations(ippi_handle* h, const double* f);

/* A circular body of nmarkers Lagrangian markers at x = cx + R sin(2 pi i/M),
 * y = cy + R cos(2 pi i/M), R = diameter/2, desired velocity (udx, udy).
 * The 4-point support of every marker must lie in the interior
 * (1 <= x <= nx-2, 1 <= y <= ny-2). Returns the body index in *body. */
int ippi_add_body(ippi_handle* h, double cx, double cy, double diameter, int nmarkers,
                  double udx, double udy, int* body);
int ippi_num_bodies(const ippi_handle* h, int* n);

/* Per-node base relaxation time (IPPI_TAU_FIELD), nx*ny doubles. */
int ippi_set_tau_field(ippi_handle* h, const double* tau);

/* Relaxation-time MLP 6 -> 32 -> 32 -> 1 (IPPI_TAU_NN); blob layout:
 * W1[32][6] b1[32] W2[32][32] b2[32] W3[32] b3 mu[6] isig[6] (n = 1325). */
int ippi_set_mlp(ippi_handle* h, const double* blob, int n);

/* Boundary-row / column relaxation-time networks, the reference's
 * Mesh::NNmodel_Type01 / 03 / 05 / 07 (mesh.cpp:3128-3194, 2767-2832,
 * 3195-3260, 2833-2898): type 1 at column 0 (1 <= y <= ny-2), 3 at row
 * ny-2, 5 at column nx-1, 7 at row 1 (every x). Each step, before the
 * collision, a node of a set type reads (u.x, u.y, rho) of five neighbours
 * in the reference's order as float, evaluates Linear(15,128) -> 6 x
 * Linear(128,128) -> Linear(128,1) in FP32 with ReLU between layers, and
 * takes the output as its tau where it is >= 0.501 (mesh.cpp:2821-2827);
 * otherwise the node keeps its LES / network / base tau. Where a row and a
 * column meet the row's network decides (the reference calls it last).
 * Neighbours beyond x = 0 and nx-1 are taken periodically (the reference's
 * ghost columns). Blob: W0[128][15] b0[128], six times W[128][128] b[128],
 * W7[128] b7 (n = 101249, PyTorch's nn.Linear layout). The reference loads
 * TorchScript files per call (AutoT0x.pt); their activation is not stated
 * in its source. */
int ippi_set_boundary_mlp(ippi_handle* h, int type, const double* blob, int n);

/* Advance nsteps time steps (asynchronous). */
int ippi_step(ippi_handle* h, int nsteps);
int ippi_synchronize(ippi_handle* h);
int ippi_steps_done(const ippi_handle* h, int64_t* steps);

/* Macroscopic fields after the last completed step (synchronising). */
int ippi_get_fields(ippi_handle* h, double* rho, double* ux, double* uy);
/* Relaxation time the next collision uses at each node (base value, LES or
 * the network with its >= 0.501 guard), evaluated on the current state
 * without the immersed-boundary velocity correction; nx*ny doubles.
 * Replaces reads of LatMesh[i][j].tau (WriteDatFluid, mesh.cpp:1231). */
int ippi_get_tau(ippi_handle* h, double* tau);
/* Post-streaming populations f (FP32 path: w + stored deviation), 9*nx*ny. */
int ippi_get_populations(ippi_handle* h, double* f);
/* Per-step force on each body (minus the sum of its marker forces, FP64),
 * records oldest first: out[rec][body][2]; *nrec <= max_records. */
int ippi_get_body_forces(ippi_handle* h, double* out, int max_records, int* nrec);
/* Marker state of the last step: out[m][6] = x, y, Ux_interp, Uy_interp, Fx, Fy. */
int ippi_get_markers(ippi_handle* h, double* out, int max_markers, int* n);
/* Global mass sum(rho) in FP64, fixed-order reduction. */
int ippi_mass(ippi_handle* h, double* mass);

/* Training pairs for the relaxation-time network, generated by this solver
 * at twice the resolution of the grid they are for (the GPU analogue of
 * FindAndWriteTau, mesh.cpp:2625-2715). The current state is restricted to
 * (nx/2) x (ny/2) (2x2 averages, non-equilibrium scaled by
 * 2 tau_coarse / tau0), the solver advances two steps (one coarse step),
 * and for every coarse node the relaxation time is found whose collision at
 * the upstream nodes, once streamed, best matches the restricted mass and
 * momentum (least squares, closed form; coarse_collision BGK or RR).
 * Outputs per coarse node: feat[6] (rho, ux, uy, |S|, Pi_xx-Pi_yy, Pi_xy of
 * the restricted state), label omega = 1/tau (the fit is linear in omega)
 * and weight |a|^2 of the fit (0 = unusable: near the domain edge or a
 * body, or omega outside (0, 4)). */
int ippi_coarse_pairs(ippi_handle* h, double tau_coarse, int coarse_collision, float* feat, float* label,
                      float* weight);

/* Effective-relaxation-time pairs at half this solver's resolution, from the
 * current state (the solver does not advance). No counterpart in the
 * reference, whose labels come from the one-step search above
 * (FindAndWriteTau, mesh.cpp:2625-2715); this criterion labels each coarse
 * node with the relaxation time whose viscous stress on the coarse grid,
 * 2 rho nu C with nu = (tau - 1/2)/3, best matches (least squares over the
 * deviatoric components) the restricted fine momentum flux: the block mean
 * of the fine viscous stress -(1 - 1/(2 tau_l)) Pi^neq (tau_l the fine
 * node's relaxation time, LES included, as in mesh.cpp:877-910) minus the
 * subgrid stress <rho u u> - rho_c u_c u_c, against the coarse strain
 * C = 2 <S> (S = -Pi^neq / (2 rho c_s^2 tau_l) per fine node). Outputs per
 * coarse node: feat[6] as above (restricted state, scale
 * s = 2 tau_coarse / tau0), label omega = 1/clamp(tau, 0.5005, 2), weight
 * dev C : dev C (0 where the coarse strain vanishes or within two coarse
 * cells of a body's marker support). */
int ippi_coarse_stress_pairs(ippi_handle* h, double tau_coarse, float* feat, float* label, float* weight);

/* The current state restricted to the grid of half the resolution, as the
 * pairs above use it: coarse node (I, J) takes the mass and momentum of fine
 * nodes (2I..2I+1, 2J..2J+1) and s = 2 tau_coarse / tau0 times their mean
 * non-equilibrium part. fc [9][ny/2][nx/2] (may be NULL), mc = rho, ux, uy
 * [3][ny/2][nx/2]. The starting state of a coarse run and its target.
 * No counterpart in the reference (its high-fidelity data come from files,
 * mesh.cpp:2414-2470). */
int ippi_restrict(ippi_handle* h, double tau_coarse, double* fc, double* mc);

/* GPU trainer of the relaxation-time network (no counterpart in the
 * reference, whose networks are trained offline). Targets are relaxation
 * rates omega; loss sum w (1/tau(x) - y)^2 / (2 sum w) in FP32;
 * deterministic gradient; Adam. */
typedef struct ippi_trainer ippi_trainer;
int ippi_trainer_create(int capacity, ippi_trainer** out);
int ippi_trainer_destroy(ippi_trainer* t);
const char* ippi_trainer_last_error(const ippi_trainer* t);
/* x: [n][6] raw features, y: [n] targets, w: [n] weights (NULL = 1). */
int ippi_trainer_set_data(ippi_trainer* t, const float* x, const float* y, const float* w, int n);
int ippi_trainer_set_params(ippi_trainer* t, const double* blob, int n);
int ippi_trainer_get_params(ippi_trainer* t, double* blob, int n);
/* Full-data loss and gradient (1313 values, blob order without mu/isig). */
int ippi_trainer_loss_grad(ippi_trainer* t, double* loss, double* grad);
/* iters Adam steps on minibatches of `batch` samples drawn with splitmix64. */
int ippi_trainer_adam(ippi_trainer* t, int iters, int batch, double lr, uint64_t seed, double* loss);

/* Kernel timing (no counterpart in the reference: measurement only).
 * enable = 1: every fused stream-collide launch and every immersed-boundary
 * kernel launch is bracketed by HIP events on the handle's stream.
 * enable = 2: one event pair brackets all launches of each ippi_step call
 * (back to back, as untimed); step_ms is then the calls' time, the launch
 * counts those of the step and IB kernels inside, and ib_ms 0.
 * enable = 0: off. ippi_get_timing synchronises, returns the summed
 * durations (ms) and launch counts since the last call, and clears them. */
int ippi_set_timing(ippi_handle* h, int enable);
int ippi_get_timing(ippi_handle* h, double* step_ms, int64_t* step_launches, double* ib_ms,
                    int64_t* ib_launches);

/* Running means of rho, ux, uy (Mesh::ComputeAvgVelocity, mesh.cpp:1954-1974):
 * updates mean = (mean * k + current) / (k + 1) at every node on the device
 * from the current state (k = t - startAvg in the reference; k = 0 starts
 * a new mean). ippi_get_mean synchronises and copies the three fields. */
int ippi_accumulate_mean(ippi_handle* h, int64_t k);
int ippi_get_mean(ippi_handle* h, double* rho, double* ux, double* uy);

/* Discrete adjoint of the time-averaged body force (no counterpart in the
 * reference, which has no gradient code; it differentiates the step that
 * ippi_step computes: collision + Guo source NodeAndLattice.cpp:150-195,
 * streaming mesh.cpp:192-226, IB mesh.cpp:514-741,1244-1281).
 * Runs nsteps forward steps from the handle's current state (the handle
 * ends advanced by nsteps), keeping a checkpoint of the state every `ckpt`
 * steps, then the reverse sweep, which recomputes each segment of ckpt
 * steps from its checkpoint. Objective, over the force records of this
 * run's steps avg_from..nsteps (1-based):
 *     J = sum_b  w[2b] mean(Fx_b) + w[2b+1] mean(Fy_b)
 * with (Fx_b, Fy_b) the body-force record of ippi_get_body_forces (minus
 * the sum of the marker forces); body_w = NULL selects the drag of body 0.
 * Outputs (each may be NULL): *J; g_tau[nx*ny] = dJ/d(base tau) per node
 * (the per-node field, or the local sensitivity to tau0 in IPPI_TAU_CONST
 * mode, whose sum is dJ/dtau0); g_center[2*nbody] = dJ/d(cx, cy) of each
 * body (all markers translate with the centre). The adjoint populations
 * and arithmetic use the handle's precision, or FP64 over an FP32 forward
 * state when flags has IPPI_ADJ_FP64 (the adjoint accumulates the
 * sensitivity of every later step, and FP32 storage drops the small late
 * increments: measured on 1024x512 / 4000 steps, FP32 adjoint vs FP64
 * adjoint 3e-5 (tau field) and up to 4e-4 (centre), FP32 forward with FP64
 * adjoint 6e-6 and 1e-5). The tau gradient is summed in FP64 per
 * checkpoint segment; the centre gradient in FP64.
 * Supported: IPPI_COLL_BGK, IPPI_BC_PERIODIC_BB, les = 0, tau_mode CONST
 * or FIELD, at least one body; otherwise IPPI_ERR_ARG / IPPI_ERR_STATE.
 * Device memory: (ceil(nsteps/ckpt) + ckpt - 1) state arrays + 2 adjoint
 * population arrays. */
#define IPPI_ADJ_FP64 1
int ippi_adjoint_drag(ippi_handle* h, int nsteps, int avg_from, int ckpt, int flags, const double* body_w,
                      double* J, double* g_tau, double* g_center);
/* Summed duration (ms) and count of the adjoint's all-node kernel launches
 * recorded while timing was enabled; clears them. */
int ippi_get_adjoint_timing(ippi_handle* h, double* ms, int64_t* launches);

/* Relaxation-time label search: the candidate loop of Mesh::FindAndWriteTau
 * (IPPI-LBM/mesh.cpp:2667-2704) for n nodes at once, stateless, FP64.
 * Node i: fnew, feq, tau [i][k] of the node population k streams in from
 * (the reference's inei/jnei neighbour of the opposite direction), target
 * [i] = (rho, u, v) of the high-fidelity data. For m = m_lo..m_hi, dtau =
 * m * step, the sources are collided with tau + dtau and scored by
 * |rho - rho*| + |sum e_x f - u*| + |sum e_y f - v*|; best_dtau[i] and
 * best_err[i] get the first smallest score (1e10 and 1e10 when no score is
 * below 1e10). The reference's own call is m_lo = -10000, m_hi = 10000,
 * step = 1e-5. */
int ippi_tau_search(int device, int n, const double* fnew, const double* feq, const double* tau,
                    const double* target, int m_lo, int m_hi, double step, double* best_dtau, double* best_err);

/* ippi_tau_search without device allocation. No counterpart in the
 * reference: for callers that cannot allocate device memory (serving
 * loops, code between captured graph launches, harnesses that forbid
 * hipMalloc). Same arguments and results as ippi_tau_search, bit for bit;
 * the device copies of the inputs and outputs live in one fixed array per
 * GPU (256 bytes per node, 8 MiB, held until the process ends), so n is at
 * most IPPI_TAU_SEARCH_NOALLOC_MAX (IPPI_ERR_CAPACITY beyond). Calls that
 * overlap on one GPU (from several host threads) share that array and
 * race; serialise them. Runs on the device's null stream and returns when
 * the results are in best_dtau / best_err. */
#define IPPI_TAU_SEARCH_NOALLOC_MAX 32768
int ippi_tau_search_noalloc(int device, int n, const double* fnew, const double* feq, const double* tau,
                            const double* target, int m_lo, int m_hi, double step, double* best_dtau,
                            double* best_err);

/* Inputs of the label search for n nodes (xy[i] = (x, y), two or more
 * nodes inside the domain), as Mesh::FindAndWriteTau gathers them
 * (mesh.cpp:2657-2686): for population k of node i, the node it streams in
 * from, (x, y) - e_k, gives fnew[i][k] (its pulled population), feq[i][k]
 * (the equilibrium of its macros) and tau[i][k] (its base or hybrid
 * Smagorinsky relaxation time, without the IB velocity correction), all as
 * the last pull step's collision used them, in FP64. The reference's stored
 * feq is one step older than its fnew; here both belong to the same
 * collision. Not for network tau. IPPI_ERR_STATE until a pull step has run
 * since the state was set. */
int ippi_get_stencil(ippi_handle* h, int n, const int* xy, double* fnew, double* feq, double* tau);

/* Bandwidth probe (no counterpart in the reference: measurement only). On
 * its own buffers and stream, times `reps` launches of a kernel with the
 * fused step's memory pattern and no arithmetic on an nx x ny lattice:
 * mode 1 = the nine pulled loads of an inflow step and nine stores, mode 2
 * = mode 1 with the rows grouped by XCD, mode 0 = a nine-array copy, mode
 * 3 = a flat copy of the same 9*nx*ny elements with 16 bytes per lane (the
 * plain HBM stream); precision IPPI_FP64 or IPPI_FP32_SHIFTED. Writes the
 * mean time per launch (microseconds). */
int ippi_pattern_probe(int device, int nx, int ny, int precision, int mode, int reps, double* us_per_launch);

/* Checkpoint: the full device state as a flat byte image. *size gives the
 * required size when buf is NULL. */
int ippi_save(ippi_handle* h, void* buf, size_t* size);
int ippi_load(ippi_handle* h, const void* buf, size_t size);

#ifdef __cplusplus
}
#endif
#endif
