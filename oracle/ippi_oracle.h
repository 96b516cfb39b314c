/*
 * IPPI-LBM D2Q9 time step -- CPU oracle (TEST INFRASTRUCTURE ONLY).
 *
 * This is a plain-C, FP64 restatement of the reference solver's per-node
 * kernels and of its time step, used only as the checker for the HIP
 * library: tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py may load it. Nothing in the product path links or calls it.
 *
 * Reference files (paths relative to the reference repository):
 *   IPPI-LBM/NodeAndLattice.cpp   per-node feq / BGK / MRT / RR / macros / Guo
 *   IPPI-LBM/mesh.cpp             streaming, boundary conditions, LES tau,
 *                                 immersed boundary, step order
 *
 * Velocity ordering (NodeAndLattice.cpp:3-13):
 *   k : 0  1  2  3  4  5  6  7  8
 *   ex: 0  1  0 -1  0  1 -1 -1  1
 *   ey: 0  0  1  0 -1  1  1 -1 -1
 * Population layout: SoA, f[k*N + y*nx + x], N = nx*ny.
 */
#ifndef IPPI_ORACLE_H
#define IPPI_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum { OR_COLL_BGK = 0, OR_COLL_MRT = 1, OR_COLL_RR = 2 };
enum { OR_BC_PERIODIC_BB = 0, OR_BC_INFLOW = 1, OR_BC_FLATPLATE = 2 };
enum { OR_TAU_CONST = 0, OR_TAU_FIELD = 1, OR_TAU_NN = 2 };

typedef struct {
    int nx, ny;
    int collision;   /* OR_COLL_* */
    int bc;          /* OR_BC_* */
    int les;         /* 1: hybrid Smagorinsky tau (mesh.cpp:877-910) */
    int tau_mode;    /* OR_TAU_* */
    double tau0;     /* startTau / TauAtBeginning */
    double cs;       /* Smagorinsky constant */
    double u_in_x;   /* inlet velocity for OR_BC_INFLOW */
    double u_in_y;
    double tau_nn_max;  /* network values above this fall back too (0 = none) */
    int plate_x;        /* OR_BC_FLATPLATE: bottom free-slip for x < plate_x, no-slip beyond */
} or_config;

/* ---- per-node kernels (NodeAndLattice.cpp, mesh.cpp) ---- */
void   or_feq(double rho, double ux, double uy, double feq[9]);
void   or_feq_init(double rho, double ux, double uy, double f[9]);
void   or_macro(const double f[9], double out[3]);
void   or_collide_bgk(const double f[9], double rho, double ux, double uy, double tau, double out[9]);
void   or_collide_mrt(const double f[9], double rho, double ux, double uy, double tau, double out[9]);
void   or_collide_rr(const double f[9], double rho, double ux, double uy, double tau, double out[9]);
void   or_guo(double ux, double uy, double fx, double fy, double tau_begin, double out[9]);
double or_les_tau(const double f[9], double rho, double ux, double uy, double tau0, double cs);
double or_delta(double r);
double or_noise(uint64_t seed, int64_t node, int comp);

/* ---- MLP relaxation-time model, FP64 (6 -> 32 -> 32 -> 1) ---- */
#define OR_MLP_NIN 6
#define OR_MLP_NH 32
#define OR_MLP_NPARAM (OR_MLP_NIN * OR_MLP_NH + OR_MLP_NH + OR_MLP_NH * OR_MLP_NH + OR_MLP_NH + OR_MLP_NH + 1)
void   or_mlp_features(const double f[9], double rho, double ux, double uy, double feat[6]);
double or_mlp_forward(const double* params, const double in[6]);
/* loss = mean_i 0.5*(1/tau(x_i) - y_i)^2 (y: relaxation rates); returns loss, writes dL/dparams */
double or_mlp_loss_grad(const double* params, const double* x, const double* y, int n, double* grad);

/* ---- boundary-row / column networks (Mesh::NNmodel_Type01/03/05/07) ---- */
/* params (FP64 values, used as FP32): W0[128][15] b0[128], 6 x (W[128][128]
 * b[128]), W7[128] b7; ReLU between layers. */
#define OR_BNN_NIN 15
#define OR_BNN_NH 128
#define OR_BNN_NHID 6
#define OR_BNN_NPARAM (OR_BNN_NIN * OR_BNN_NH + OR_BNN_NH + OR_BNN_NHID * (OR_BNN_NH * OR_BNN_NH + OR_BNN_NH) + OR_BNN_NH + 1)
float  or_bnn_forward(const float* params, const float in[OR_BNN_NIN]);

/* ---- whole-domain solver ---- */
typedef struct or_sim or_sim;
or_sim* or_create(const or_config* cfg);
void    or_destroy(or_sim* s);
void    or_set_threads(int n);
void    or_init_fields(or_sim* s, uint64_t seed, double rho0, double ux0, double uy0, double noise);
void    or_set_state(or_sim* s, const double* fnew, const double* rho, const double* ux, const double* uy);
int     or_add_body(or_sim* s, double cx, double cy, double diameter, int nmarkers, double udx, double udy);
void    or_set_tau_field(or_sim* s, const double* tau);
void    or_set_mlp(or_sim* s, const double* params);
/* type 1, 3, 5 or 7: the reference's network type; returns 0, or -1 for another type */
int     or_set_bnn(or_sim* s, int type, const double* params);
int     or_step(or_sim* s, int nsteps);
void    or_get_fields(const or_sim* s, double* rho, double* ux, double* uy);
void    or_get_fnew(const or_sim* s, double* fnew);
void    or_get_fcol(const or_sim* s, double* fcol);
void    or_get_tau(const or_sim* s, double* tau);
int     or_num_force_records(const or_sim* s);
void    or_get_forces(const or_sim* s, double* out); /* [rec][body][fx,fy] */
int     or_num_markers(const or_sim* s);
void    or_get_markers(const or_sim* s, double* out); /* [m][x,y,ux_interp,uy_interp,Fx,Fy] */
double  or_mass(const or_sim* s);
void    or_get_force_field(const or_sim* s, double* fx, double* fy);

/* ---- discrete adjoint of the BGK + IB step (FP64, all states kept) ---- */
/* J = mean of the body's drag record over steps avg_from..nsteps (1-based);
 * gradients w.r.t. the per-node base tau (g_tau, N values) and the body
 * centre (g_c, 2 values). Periodic x, bounce-back walls, one body. */
int    or_adjoint_drag(int nx, int ny, uint64_t seed, double ux0, double uy0, double noise, const double* tau,
                       double cx, double cy, double diam, int M, int nsteps, int avg_from, double* J,
                       double* g_tau, double* g_c);
double or_drag_objective(int nx, int ny, uint64_t seed, double ux0, double uy0, double noise, const double* tau,
                         double cx, double cy, double diam, int M, int nsteps, int avg_from);

/* ---- tau-label search (FindAndWriteTau, IPPI-LBM/mesh.cpp:2667-2704) ---- */
/* n nodes; node i gives, for each population k, the post-streaming value
 * fnew[i][k], the equilibrium feq[i][k] and the relaxation time tau[i][k]
 * of the node k streams from, and target[i] = (rho, u, v) of the
 * high-fidelity data. For m = m_lo..m_hi, dtau = m * step: collide those
 * populations with tau + dtau, sum rho, sum e_x f, sum e_y f, and score
 * |rho - rho*| + |jx - u*| + |jy - v*|. The first m with the smallest score
 * wins (best_err = 1e10, best_dtau = 1e10 when no score is below 1e10). */
void or_tau_search(int n, const double* fnew, const double* feq, const double* tau, const double* target, int m_lo,
                   int m_hi, double step, double* best_dtau, double* best_err);

#ifdef __cplusplus
}
#endif
#endif
