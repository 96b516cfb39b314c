/*
 * IPPI-LBM D2Q9 time step -- CPU oracle, FP64 (TEST INFRASTRUCTURE ONLY).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this file's library, and only as the checker. The HIP library never
 * links it.
 *
 * Each function cites the reference lines it restates. Floating-point
 * expressions keep the reference's evaluation order (the file is built
 * with -ffp-contract=off) so the FP64 HIP path can be compared bit for bit.
 *
 * Pins (tests/test_oracle_*.py):
 *   or_feq, or_feq_init, or_macro, or_collide_bgk/mrt/rr, or_guo
 *       bitwise against the reference's own Lattice methods, compiled from
 *       IPPI-LBM/NodeAndLattice.cpp by `make -C oracle ref` (golden fixtures
 *       in tests/golden/), plus moment identities.
 *   or_les_tau   the quadratic it solves, tau (tau - tau0) = 4.5 Cs^2 |S|/rho,
 *                with |S| from Pi_neq computed independently; Cs = 0 and
 *                Pi_neq = 0 give tau0.
 *   or_delta     Peskin identities (sum = 1, first moment 0, sum of
 *                squares 3/8).
 *   streaming / BC / IB composition: mass conservation, permutation check
 *                by brute force and numpy.roll, decay of a shear wave between
 *                halfway bounce-back walls, interpolation of linear
 *                fields, spread = transpose of interpolation.
 *   inlet / outlet NEE and free-slip rows: per-node closed forms (rho = 1
 *                and u = u_in at the inlet, u copied at the outlet, the
 *                specular mirror at the walls).
 *   stream_flatplate  each rule checked node by node on a noisy flow (NEE
 *                inlet/outlet/top, 'm' mirror upstream of the plate, 'd'
 *                no-slip NEE on it, interior pull = shift, frozen bottom
 *                corners); uniform inflow with no plate is a fixed point to
 *                1e-14; a boundary layer grows behind the leading edge.
 *   or_noise     the published splitmix64 outputs for seed 0.
 *   MLP          torch autograd (FP64) and central finite differences.
 *   or_adjoint_drag  central finite differences (h = 1e-4) of the forward
 *                step for single nodes, random directions and the body
 *                centre, on horizons of 1 to 50 steps; exact zeros outside
 *                the domain of dependence; translation equivariance.
 *   or_bnn_forward, or_set_bnn  the boundary-row / column networks
 *                (mesh.cpp:2767-2898, 3128-3260): the forward pass against
 *                torch.nn in FP64 and FP32; which node each type writes and
 *                which five neighbours it reads, in the reference's input
 *                order (pass-through networks); the >= 0.501 guard and the
 *                row-over-column order where they meet.
 * The step as a whole (mesh.cpp) is not compiled (it needs MPI and
 * libtorch), so the composition is pinned by the checks above rather than
 * bit for bit. Parity unpinned: the outputs of the reference's TorchScript
 * archives (AutoT0x.pt), which are not loaded; or_bnn_forward follows the
 * layer shapes recorded from the reference, with ReLU.
 */
#include "ippi_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* NodeAndLattice.cpp:3-14 */
static const int EX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
static const int EY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
static const int OPP[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6}; /* mesh.cpp:236 */
#define W0 (4. / 9.)
#define W1 (1. / 9.)
#define W2 (1. / 36.)
static const double WEIGHT[9] = {W0, W1, W1, W1, W1, W2, W2, W2, W2};

/* NodeAndLattice.cpp:18-29 (MRT transform and its inverse) */
static const double TM[9][9] = {
    {1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0},
    {-4.0, -1.0, -1.0, -1.0, -1.0, 2.0, 2.0, 2.0, 2.0},
    {4.0, -2.0, -2.0, -2.0, -2.0, 1.0, 1.0, 1.0, 1.0},
    {0.0, 1.0, 0.0, -1.0, 0.0, 1.0, -1.0, -1.0, 1.0},
    {0.0, -2.0, 0.0, 2.0, 0.0, 1.0, -1.0, -1.0, 1.0},
    {0.0, 0.0, 1.0, 0.0, -1.0, 1.0, 1.0, -1.0, -1.0},
    {0.0, 0.0, -2.0, 0.0, 2.0, 1.0, 1.0, -1.0, -1.0},
    {0.0, 1.0, -1.0, 1.0, -1.0, 0.0, 0.0, 0.0, 0.0},
    {0.0, 0.0, 0.0, 0.0, 0.0, 1.0, -1.0, 1.0, -1.0}};
/* entries are integer multiples of a1 = 1/36, each product rounded once */
static const int TMINV_K[9][9] = {
    {4, -4, 4, 0, 0, 0, 0, 0, 0},    {4, -1, -2, 6, -6, 0, 0, 9, 0},
    {4, -1, -2, 0, 0, 6, -6, -9, 0}, {4, -1, -2, -6, 6, 0, 0, 9, 0},
    {4, -1, -2, 0, 0, -6, 6, -9, 0}, {4, 2, 1, 6, 3, 6, 3, 0, 9},
    {4, 2, 1, -6, -3, 6, 3, 0, -9},  {4, 2, 1, -6, -3, -6, -3, 0, 9},
    {4, 2, 1, 6, 3, -6, -3, 0, -9}};

static double tminv(int i, int j) {
    const double a1 = 1.0 / 36.0;
    return (double)TMINV_K[i][j] * a1;
}

/* Compute_feq, NodeAndLattice.cpp:124-149 */
void or_feq(double rho, double ux, double uy, double feq[9]) {
    double w0 = rho * (4.0 / 9.0);
    double w1 = rho / 9.0;
    double w2 = rho / 36.0;
    double uu = 1 - 1.5 * (ux * ux + uy * uy);
    feq[0] = w0 * (uu);
    feq[1] = w1 * (uu + 3.0 * (ux) + 4.5 * (ux) * (ux));
    feq[2] = w1 * (uu + 3.0 * (uy) + 4.5 * (uy) * (uy));
    feq[3] = w1 * (uu - 3.0 * (ux) + 4.5 * (ux) * (ux));
    feq[4] = w1 * (uu - 3.0 * (uy) + 4.5 * (uy) * (uy));
    feq[5] = w2 * (uu + 3.0 * (ux + uy) + 4.5 * (ux + uy) * (ux + uy));
    feq[6] = w2 * (uu - 3.0 * (ux - uy) + 4.5 * (ux - uy) * (ux - uy));
    feq[7] = w2 * (uu - 3.0 * (ux + uy) + 4.5 * (ux + uy) * (ux + uy));
    feq[8] = w2 * (uu + 3.0 * (ux - uy) + 4.5 * (ux - uy) * (ux - uy));
}

/* InitialLatticeFeq, NodeAndLattice.cpp:104-119 (note rho*4.0/9.0) */
void or_feq_init(double rho, double ux, double uy, double f[9]) {
    double w0 = rho * 4.0 / 9.0;
    double w1 = rho / 9.0;
    double w2 = rho / 36.0;
    double uu = 1 - 1.5 * (ux * ux + uy * uy);
    f[0] = w0 * (uu);
    f[1] = w1 * (uu + 3.0 * (ux) + 4.5 * (ux) * (ux));
    f[2] = w1 * (uu + 3.0 * (uy) + 4.5 * (uy) * (uy));
    f[3] = w1 * (uu - 3.0 * (ux) + 4.5 * (ux) * (ux));
    f[4] = w1 * (uu - 3.0 * (uy) + 4.5 * (uy) * (uy));
    f[5] = w2 * (uu + 3.0 * (ux + uy) + 4.5 * (ux + uy) * (ux + uy));
    f[6] = w2 * (uu - 3.0 * (ux - uy) + 4.5 * (ux - uy) * (ux - uy));
    f[7] = w2 * (uu - 3.0 * (ux + uy) + 4.5 * (ux + uy) * (ux + uy));
    f[8] = w2 * (uu + 3.0 * (ux - uy) + 4.5 * (ux - uy) * (ux - uy));
}

/* CalMacro, NodeAndLattice.cpp:196-226 (half-force shift commented out there) */
void or_macro(const double f[9], double out[3]) {
    double trho = 0.0, tx = 0.0, ty = 0.0;
    for (int k = 0; k < 9; k++) {
        trho += f[k];
        tx += f[k] * EX[k];
        ty += f[k] * EY[k];
    }
    if (trho < 1.e-6) {
        tx = 0.0;
        ty = 0.0;
    } else {
        tx /= trho;
        ty /= trho;
    }
    out[0] = trho;
    out[1] = tx;
    out[2] = ty;
}

/* ComputeBGKCollision, NodeAndLattice.cpp:150-161 */
void or_collide_bgk(const double f[9], double rho, double ux, double uy, double tau, double out[9]) {
    double feq[9];
    or_feq(rho, ux, uy, feq);
    for (int k = 0; k < 9; k++) out[k] = (1.0 - 1.0 / tau) * f[k] + feq[k] / tau;
}

/* ComputeMRTCollision, NodeAndLattice.cpp:290-360 (relaxation rates :306) */
void or_collide_mrt(const double f[9], double rho, double ux, double uy, double tau, double out[9]) {
    double sm[9] = {0, 1.2, 1.1, 0, 1.2, 0, 1.2, 1. / tau, 1. / tau};
    double fmom[9], fmeq[9];
    double t1 = ux * ux + uy * uy;
    double t2 = ux * ux - uy * uy;
    fmeq[0] = rho;
    fmeq[1] = rho * (-2.0 + 3.0 * t1);
    fmeq[2] = rho * (1.0 - 3.0 * t1);
    fmeq[3] = rho * ux;
    fmeq[4] = -rho * ux;
    fmeq[5] = rho * uy;
    fmeq[6] = -rho * uy;
    fmeq[7] = rho * t2;
    fmeq[8] = rho * ux * uy;
    for (int k = 0; k < 9; k++) {
        double s = 0.0;
        for (int l = 0; l < 9; l++) s += TM[k][l] * f[l];
        fmom[k] = s;
    }
    for (int k = 0; k < 9; k++) {
        double s = 0.;
        for (int l = 0; l < 9; l++) s += (tminv(k, l) * sm[l]) * (fmom[l] - fmeq[l]);
        out[k] = f[k] - s;
    }
}

/* ComputeRRCollision, NodeAndLattice.cpp:228-286 */
void or_collide_rr(const double f[9], double rho, double ux, double uy, double tau, double out[9]) {
    double feq[9];
    or_feq(rho, ux, uy, feq);
    double xx = 0, txx = 0, xy = 0, txy = 0, yx = 0, tyx = 0, yy = 0, tyy = 0;
    for (int k = 0; k < 9; k++) {
        xx += f[k] * EX[k] * EX[k];
        txx += feq[k] * EX[k] * EX[k];
        xy += f[k] * EX[k] * EY[k];
        txy += feq[k] * EX[k] * EY[k];
        yx += f[k] * EY[k] * EX[k];
        tyx += feq[k] * EY[k] * EX[k];
        yy += f[k] * EY[k] * EY[k];
        tyy += feq[k] * EY[k] * EY[k];
    }
    double nxx = xx - txx, nxy = xy - txy, nyx = yx - tyx, nyy = yy - tyy;
    for (int k = 0; k < 9; k++) {
        double q0 = EX[k] * EX[k] - 1. / 3;
        double q1 = EX[k] * EY[k];
        double q2 = EY[k] * EX[k];
        double q3 = EY[k] * EY[k] - 1. / 3;
        double fone = 4.5 * WEIGHT[k] * (q0 * nxx + q1 * nxy + q2 * nyx + q3 * nyy);
        out[k] = feq[k] + (1.0 - 1.0 / tau) * fone;
    }
}

/* Compute_Disforce (Guo forcing term), NodeAndLattice.cpp:162-195 */
void or_guo(double ux, double uy, double fx, double fy, double tau_begin, double out[9]) {
    double omega = 1.0 - 0.5 * (1.0 / tau_begin);
    for (int k = 0; k < 9; k++) {
        double eu = (EX[k] * ux + EY[k] * uy);
        double gx = WEIGHT[k] * omega * (3.0 * (EX[k] - ux) + 9.0 * eu * EX[k]) * fx;
        double gy = WEIGHT[k] * omega * (3.0 * (EY[k] - uy) + 9.0 * eu * EY[k]) * fy;
        out[k] = gx + gy;
    }
}

/* CaculateLESTau live body (hybrid Smagorinsky), mesh.cpp:877-910 */
double or_les_tau(const double f[9], double rho, double ux, double uy, double tau0, double cs) {
    double rhoTemp = 1.0;
    rhoTemp *= rho;
    double feq[9];
    or_feq(rho, ux, uy, feq);
    double Sxx = 0.0, Sxy = 0.0, Syx, Syy = 0.0;
    for (int k = 0; k < 9; k++) {
        Sxx += EX[k] * EX[k] * (f[k] - feq[k]);
        Sxy += EX[k] * EY[k] * (f[k] - feq[k]);
        Syy += EY[k] * EY[k] * (f[k] - feq[k]);
    }
    Syx = Sxy;
    Sxx *= -3.0 / (2 * rhoTemp * tau0);
    Sxy *= -3.0 / (2 * rhoTemp * tau0);
    Syx *= -3.0 / (2 * rhoTemp * tau0);
    Syy *= -3.0 / (2 * rhoTemp * tau0);
    double S2 = Sxx * Sxx + Sxy * Sxy + Syx * Syx + Syy * Syy;
    double Q = sqrt(2.0 * S2);
    double g1 = tau0 * tau0 + 18.0 * cs * cs * Q / (rhoTemp);
    return (sqrt(g1) + tau0) / 2.0;
}

/* DiracDelta (Peskin 4-point), mesh.cpp:1354-1385 */
double or_delta(double x) {
    if (fabs(x) < 1) {
        double R = 1.0 + 4.0 * fabs(x) - 4 * (x * x);
        return (1.0 / 8.0) * (3.0 - 2.0 * fabs(x) + sqrt(R));
    } else if (fabs(x) < 2) {
        double R = -7.0 + 12.0 * fabs(x) - 4 * (x * x);
        return (1.0 / 8.0) * (5.0 - 2.0 * fabs(x) - sqrt(R));
    }
    return 0;
}

/* Seeded uniform noise in [-1, 1): splitmix64 finaliser keyed by
 * (seed, node, component). Identical formula in the HIP library. */
double or_noise(uint64_t seed, int64_t node, int comp) {
    uint64_t z = seed + (uint64_t)(2 * node + comp + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    double r = (double)(z >> 11) * (1.0 / 9007199254740992.0);
    return 2.0 * r - 1.0;
}

/* ------------------------------------------------------------------ */
/* MLP relaxation-time model (6 -> 32 -> 32 -> 1, SiLU, softplus + 0.5)  */
/* Parameter blob: W1[32][6] b1[32] W2[32][32] b2[32] W3[32] b3          */
/*                 then feature normalisation mu[6] isig[6].             */
/* ------------------------------------------------------------------ */
#define NIN OR_MLP_NIN
#define NH OR_MLP_NH
#define OFF_W1 0
#define OFF_B1 (OFF_W1 + NH * NIN)
#define OFF_W2 (OFF_B1 + NH)
#define OFF_B2 (OFF_W2 + NH * NH)
#define OFF_W3 (OFF_B2 + NH)
#define OFF_B3 (OFF_W3 + NH)
#define OFF_MU (OFF_B3 + 1)
#define OFF_IS (OFF_MU + NIN)

/* Features: rho, ux, uy, |S| (the LES strain magnitude Q of mesh.cpp:886-906,
 * built on tau0 = 1 so it is the raw |Pi_neq| scale), Pi_xx - Pi_yy, Pi_xy. */
void or_mlp_features(const double f[9], double rho, double ux, double uy, double feat[6]) {
    double feq[9];
    or_feq(rho, ux, uy, feq);
    double pxx = 0, pxy = 0, pyy = 0;
    for (int k = 0; k < 9; k++) {
        double d = f[k] - feq[k];
        pxx += EX[k] * EX[k] * d;
        pxy += EX[k] * EY[k] * d;
        pyy += EY[k] * EY[k] * d;
    }
    double c = -3.0 / (2 * rho);
    double sxx = pxx * c, sxy = pxy * c, syy = pyy * c;
    feat[0] = rho;
    feat[1] = ux;
    feat[2] = uy;
    feat[3] = sqrt(2.0 * (sxx * sxx + 2.0 * sxy * sxy + syy * syy));
    feat[4] = pxx - pyy;
    feat[5] = pxy;
}

static double silu(double z) { return z / (1.0 + exp(-z)); }
static double dsilu(double z) {
    double s = 1.0 / (1.0 + exp(-z));
    return s * (1.0 + z * (1.0 - s));
}
static double softplus(double z) { return z > 30.0 ? z : log1p(exp(z)); }
static double sigm(double z) { return 1.0 / (1.0 + exp(-z)); }

static double mlp_fwd(const double* p, const double in[6], double* z1, double* h1, double* z2,
                      double* h2, double* xo, double* o_pre) {
    double x[NIN];
    for (int i = 0; i < NIN; i++) x[i] = (in[i] - p[OFF_MU + i]) * p[OFF_IS + i];
    for (int j = 0; j < NH; j++) {
        double s = p[OFF_B1 + j];
        for (int i = 0; i < NIN; i++) s += p[OFF_W1 + j * NIN + i] * x[i];
        z1[j] = s;
        h1[j] = silu(s);
    }
    for (int j = 0; j < NH; j++) {
        double s = p[OFF_B2 + j];
        for (int i = 0; i < NH; i++) s += p[OFF_W2 + j * NH + i] * h1[i];
        z2[j] = s;
        h2[j] = silu(s);
    }
    double o = p[OFF_B3];
    for (int i = 0; i < NH; i++) o += p[OFF_W3 + i] * h2[i];
    if (xo)
        for (int i = 0; i < NIN; i++) xo[i] = x[i];
    *o_pre = o;
    return softplus(o) + 0.5;
}

double or_mlp_forward(const double* params, const double in[6]) {
    double z1[NH], h1[NH], z2[NH], h2[NH], o;
    return mlp_fwd(params, in, z1, h1, z2, h2, NULL, &o);
}

double or_mlp_loss_grad(const double* p, const double* xs, const double* ys, int n, double* g) {
    memset(g, 0, sizeof(double) * OR_MLP_NPARAM);
    double loss = 0.0;
    for (int s = 0; s < n; s++) {
        double z1[NH], h1[NH], z2[NH], h2[NH], x[NIN], o;
        double t = mlp_fwd(p, xs + (size_t)s * NIN, z1, h1, z2, h2, x, &o);
        double r = 1.0 / t - ys[s];
        loss += 0.5 * r * r;
        double dt = -r / (t * t) / n;
        double dout = dt * sigm(o);
        g[OFF_B3] += dout;
        double dz2[NH];
        for (int i = 0; i < NH; i++) {
            g[OFF_W3 + i] += dout * h2[i];
            dz2[i] = dout * p[OFF_W3 + i] * dsilu(z2[i]);
        }
        double dh1[NH];
        for (int i = 0; i < NH; i++) dh1[i] = 0.0;
        for (int j = 0; j < NH; j++) {
            g[OFF_B2 + j] += dz2[j];
            for (int i = 0; i < NH; i++) {
                g[OFF_W2 + j * NH + i] += dz2[j] * h1[i];
                dh1[i] += dz2[j] * p[OFF_W2 + j * NH + i];
            }
        }
        for (int j = 0; j < NH; j++) {
            double dz1 = dh1[j] * dsilu(z1[j]);
            g[OFF_B1 + j] += dz1;
            for (int i = 0; i < NIN; i++) g[OFF_W1 + j * NIN + i] += dz1 * x[i];
        }
    }
    return loss / n;
}

/* ------------------------------------------------------------------ */
/* Whole-domain solver                                                 */
/* ------------------------------------------------------------------ */
typedef struct {
    double x, y;      /* position (IBParticle.position) */
    double udx, udy;  /* desired velocity (IBParticle.vel) */
    double uix, uiy;  /* interpolated velocity (IBParticle.vel_ref) */
    double fx, fy;    /* marker force */
} or_marker;

struct or_sim {
    or_config c;
    int N;
    double *fnew, *fcol, *rho, *ux, *uy, *tau, *tau_base, *Fx, *Fy;
    int nbody;
    int body_first[64], body_count[64];
    or_marker* mk;
    int nmk, cap_mk;
    double* forces; /* [rec][nbody][2] */
    int nrec, cap_rec;
    double mlp[OR_MLP_NPARAM + 2 * OR_MLP_NIN];
    int has_mlp;
    float* bnn[4];     /* boundary networks T01, T03, T05, T07 (FP32 parameters), NULL = not set */
    double* bnn_tau;   /* per node: accepted boundary-network tau of this step, 0 = none */
};

static int g_threads = 0;
void or_set_threads(int n) { g_threads = n; }

or_sim* or_create(const or_config* cfg) {
    or_sim* s = (or_sim*)calloc(1, sizeof(or_sim));
    s->c = *cfg;
    s->N = cfg->nx * cfg->ny;
    size_t N = (size_t)s->N;
    s->fnew = (double*)calloc(9 * N, sizeof(double));
    s->fcol = (double*)calloc(9 * N, sizeof(double));
    s->rho = (double*)calloc(N, sizeof(double));
    s->ux = (double*)calloc(N, sizeof(double));
    s->uy = (double*)calloc(N, sizeof(double));
    s->tau = (double*)calloc(N, sizeof(double));
    s->tau_base = (double*)calloc(N, sizeof(double));
    s->Fx = (double*)calloc(N, sizeof(double));
    s->Fy = (double*)calloc(N, sizeof(double));
    for (size_t i = 0; i < N; i++) s->tau[i] = s->tau_base[i] = cfg->tau0;
    return s;
}

void or_destroy(or_sim* s) {
    if (!s) return;
    free(s->fnew); free(s->fcol); free(s->rho); free(s->ux); free(s->uy);
    free(s->tau); free(s->tau_base); free(s->Fx); free(s->Fy);
    free(s->mk); free(s->forces);
    for (int t = 0; t < 4; t++) free(s->bnn[t]);
    free(s->bnn_tau);
    free(s);
}

/* InitialLatMesh + InitialLatticeFeq (mesh.cpp:88-167): rho, u, f = feq.
 * The initial velocity is U0 plus seeded uniform noise of amplitude `noise`. */
void or_init_fields(or_sim* s, uint64_t seed, double rho0, double ux0, double uy0, double noise) {
    int N = s->N;
    for (int n = 0; n < N; n++) {
        double ux = ux0 + noise * or_noise(seed, n, 0);
        double uy = uy0 + noise * or_noise(seed, n, 1);
        double f[9];
        or_feq_init(rho0, ux, uy, f);
        for (int k = 0; k < 9; k++) s->fnew[(size_t)k * N + n] = s->fcol[(size_t)k * N + n] = f[k];
        s->rho[n] = rho0;
        s->ux[n] = ux;
        s->uy[n] = uy;
    }
    s->nrec = 0;
}

void or_set_state(or_sim* s, const double* fnew, const double* rho, const double* ux, const double* uy) {
    size_t N = (size_t)s->N;
    memcpy(s->fnew, fnew, 9 * N * sizeof(double));
    memcpy(s->rho, rho, N * sizeof(double));
    memcpy(s->ux, ux, N * sizeof(double));
    memcpy(s->uy, uy, N * sizeof(double));
    s->nrec = 0;
}

/* InitialIBParticle, mesh.cpp:173-191 */
int or_add_body(or_sim* s, double cx, double cy, double diameter, int M, double udx, double udy) {
    if (s->nbody >= 64 || M <= 0) return -1;
    double R = 0.5 * diameter;
    if (s->nmk + M > s->cap_mk) {
        s->cap_mk = (s->nmk + M) * 2;
        s->mk = (or_marker*)realloc(s->mk, sizeof(or_marker) * s->cap_mk);
    }
    s->body_first[s->nbody] = s->nmk;
    s->body_count[s->nbody] = M;
    for (int i = 0; i < M; i++) {
        or_marker* m = &s->mk[s->nmk + i];
        m->x = cx + R * sin(2. * M_PI * (double)i / M);
        m->y = cy + R * cos(2. * M_PI * (double)i / M);
        m->udx = udx;
        m->udy = udy;
        m->uix = m->uiy = m->fx = m->fy = 0.0;
    }
    s->nmk += M;
    s->nbody++;
    return s->nbody - 1;
}

void or_set_tau_field(or_sim* s, const double* tau) {
    memcpy(s->tau_base, tau, sizeof(double) * s->N);
    memcpy(s->tau, tau, sizeof(double) * s->N);
}

void or_set_mlp(or_sim* s, const double* params) {
    memcpy(s->mlp, params, sizeof(s->mlp));
    s->has_mlp = 1;
}

/* ------------------------------------------------------------------ */
/* Boundary-row and -column networks: Mesh::NNmodel_Type01/03/05/07     */
/* (mesh.cpp:3128-3194, 2767-2832, 3195-3260, 2833-2898). Each reads     */
/* (u.x, u.y, rho) of five neighbours as float (mesh.cpp:2806-2808),     */
/* runs Linear(15,128) -> 6 x Linear(128,128) -> Linear(128,1) in FP32     */
/* and sets the node's tau where the output is >= 0.501 (mesh.cpp:       */
/* 2821-2827), the LES value otherwise. The activation between layers is  */
/* ReLU (the source does not state it). Products are fused multiply-adds  */
/* in increasing input order, so the GPU evaluation matches bit for bit.  */
/* ------------------------------------------------------------------ */
float or_bnn_forward(const float* p, const float in[OR_BNN_NIN]) {
    float a[OR_BNN_NH], b[OR_BNN_NH];
    const float* W = p;
    const float* B = p + OR_BNN_NH * OR_BNN_NIN;
    for (int o = 0; o < OR_BNN_NH; o++) {
        float v = B[o];
        for (int k = 0; k < OR_BNN_NIN; k++) v = fmaf(W[o * OR_BNN_NIN + k], in[k], v);
        a[o] = v > 0.0f ? v : 0.0f;
    }
    size_t off = OR_BNN_NH * OR_BNN_NIN + OR_BNN_NH;
    for (int l = 0; l < OR_BNN_NHID; l++) {
        W = p + off;
        B = W + OR_BNN_NH * OR_BNN_NH;
        for (int o = 0; o < OR_BNN_NH; o++) {
            float v = B[o];
            for (int k = 0; k < OR_BNN_NH; k++) v = fmaf(W[o * OR_BNN_NH + k], a[k], v);
            b[o] = v > 0.0f ? v : 0.0f;
        }
        memcpy(a, b, sizeof(a));
        off += OR_BNN_NH * OR_BNN_NH + OR_BNN_NH;
    }
    W = p + off;
    float v = W[OR_BNN_NH];
    for (int k = 0; k < OR_BNN_NH; k++) v = fmaf(W[k], a[k], v);
    return v;
}

/* node (x, y) of type index t (0: T01, 1: T03, 2: T05, 3: T07) and input j's
 * neighbour offset; the neighbours' order is the reference's input order */
static const int BNN_OFF[4][5][2] = {
    {{0, 1}, {1, 1}, {1, 0}, {1, -1}, {0, -1}},      /* T01, column 0, mesh.cpp:3149-3165 */
    {{1, 0}, {1, -1}, {0, -1}, {-1, -1}, {-1, 0}},   /* T03, row ny-2, mesh.cpp:2788-2804 */
    {{0, 1}, {-1, 1}, {-1, 0}, {-1, -1}, {0, -1}},   /* T05, column nx-1, mesh.cpp:3216-3232 */
    {{-1, 0}, {-1, 1}, {0, 1}, {1, 1}, {1, 0}}};     /* T07, row 1, mesh.cpp:2854-2870 */

int or_set_bnn(or_sim* s, int type, const double* params) {
    const int t = type == 1 ? 0 : type == 3 ? 1 : type == 5 ? 2 : type == 7 ? 3 : -1;
    if (t < 0) return -1;
    if (!s->bnn[t]) s->bnn[t] = (float*)malloc(sizeof(float) * OR_BNN_NPARAM);
    for (int i = 0; i < OR_BNN_NPARAM; i++) s->bnn[t][i] = (float)params[i];
    if (!s->bnn_tau) s->bnn_tau = (double*)calloc((size_t)s->N, sizeof(double));
    return 0;
}

/* The reference's call order in one step (mesh.cpp:363-378, one rank):
 * T01, T05, then T03, T07, each writing the tau or the rejection of its
 * nodes, so the row networks decide at the four nodes a column shares
 * with them. Neighbours beyond x = 0 or nx-1 are the periodic ring's
 * (the reference's ghost columns, mesh.cpp:440-501). */
static void bnn_phase(or_sim* s) {
    const int nx = s->c.nx, ny = s->c.ny;
    static const int order[4] = {0, 2, 1, 3};
    for (int q = 0; q < 4; q++) {
        const int t = order[q];
        if (!s->bnn[t]) continue;
        const int cnt = (t == 1 || t == 3) ? nx : ny - 2;
        for (int i = 0; i < cnt; i++) {
            int x, y;
            if (t == 0) { x = 0; y = 1 + i; }
            else if (t == 2) { x = nx - 1; y = 1 + i; }
            else if (t == 1) { x = i; y = ny - 2; }
            else { x = i; y = 1; }
            float in[OR_BNN_NIN];
            for (int j = 0; j < 5; j++) {
                int xn = x + BNN_OFF[t][j][0], yn = y + BNN_OFF[t][j][1];
                if (xn < 0) xn += nx;
                if (xn >= nx) xn -= nx;
                const int n = yn * nx + xn;
                in[3 * j] = (float)s->ux[n];
                in[3 * j + 1] = (float)s->uy[n];
                in[3 * j + 2] = (float)s->rho[n];
            }
            const double tau = (double)or_bnn_forward(s->bnn[t], in);
            s->bnn_tau[y * nx + x] = tau >= 0.501 ? tau : 0.0;
        }
    }
}

static void record_forces(or_sim* s) {
    if (s->nrec + 1 > s->cap_rec) {
        s->cap_rec = s->cap_rec ? s->cap_rec * 2 : 1024;
        s->forces = (double*)realloc(s->forces, sizeof(double) * 2 * (size_t)s->cap_rec * (s->nbody ? s->nbody : 1));
    }
    for (int b = 0; b < s->nbody; b++) {
        double fx = 0.0, fy = 0.0;
        for (int m = s->body_first[b]; m < s->body_first[b] + s->body_count[b]; m++) {
            fx += s->mk[m].fx;
            fy += s->mk[m].fy;
        }
        s->forces[((size_t)s->nrec * s->nbody + b) * 2 + 0] = -fx;
        s->forces[((size_t)s->nrec * s->nbody + b) * 2 + 1] = -fy;
    }
    s->nrec++;
}

/* Immersed boundary, wired before collision:
 *   InterpolateVel (4-point delta)  mesh.cpp:555-605
 *   CalParticleForce (DirectForce)  mesh.cpp:514-532
 *   SpreadForce (4-point delta)     mesh.cpp:659-741
 *   UpdateVelocity                  mesh.cpp:1244-1281 */
static void ib_phase(or_sim* s) {
    const int nx = s->c.nx;
    for (int i = 0; i < s->nmk; i++) {
        or_marker* m = &s->mk[i];
        m->uix = 0;
        m->uiy = 0;
        int XI = (int)(m->x), YI = (int)(m->y);
        for (int X = XI - 1; X <= XI + 2; X++)
            for (int Y = YI - 1; Y <= YI + 2; Y++) {
                double dx = m->x - X, dy = m->y - Y;
                double D = or_delta(dx) * or_delta(dy);
                int n = Y * nx + X;
                m->uix += (s->ux[n] * D);
                m->uiy += (s->uy[n] * D);
            }
    }
    for (int i = 0; i < s->nmk; i++) {
        or_marker* m = &s->mk[i];
        m->fx = 2 * (m->udx - m->uix);
        m->fy = 2 * (m->udy - m->uiy);
    }
    memset(s->Fx, 0, sizeof(double) * s->N);
    memset(s->Fy, 0, sizeof(double) * s->N);
    for (int i = 0; i < s->nmk; i++) {
        or_marker* m = &s->mk[i];
        int XI = (int)(m->x), YI = (int)(m->y);
        for (int X = XI - 1; X <= XI + 2; X++)
            for (int Y = YI - 1; Y <= YI + 2; Y++) {
                double dx = m->x - X, dy = m->y - Y;
                double D = or_delta(dx) * or_delta(dy);
                int n = Y * nx + X;
                s->Fx[n] += (m->fx * D);
                s->Fy[n] += (m->fy * D);
            }
    }
    for (int n = 0; n < s->N; n++) {
        s->ux[n] = s->ux[n] + 0.5 * 1 * s->Fx[n] / s->rho[n];
        s->uy[n] = s->uy[n] + 0.5 * 1 * s->Fy[n] / s->rho[n];
    }
    record_forces(s);
}

/* Collision at every node (mesh.cpp:386-438): tau from the NN guard
 * (mesh.cpp:2821-2827, :393), the LES model or the base value, then the
 * collision operator, plus the Guo term where IB force acts. */
static void collide_phase(or_sim* s) {
    const int N = s->N;
    const or_config* c = &s->c;
    const int ib = s->nbody > 0;
#pragma omp parallel for schedule(static) if (N > 4096) num_threads(g_threads > 0 ? g_threads : 1)
    for (int n = 0; n < N; n++) {
        double f[9], out[9];
        for (int k = 0; k < 9; k++) f[k] = s->fnew[(size_t)k * N + n];
        double rho = s->rho[n], ux = s->ux[n], uy = s->uy[n];
        double tb = s->tau_base[n];
        int nn_ok = 0;
        double tau = tb;
        if (s->bnn_tau && s->bnn_tau[n] > 0.0) { /* a boundary network's accepted value */
            tau = s->bnn_tau[n];
            nn_ok = 1;
        } else if (c->tau_mode == OR_TAU_NN && s->has_mlp) {
            double feat[6];
            or_mlp_features(f, rho, ux, uy, feat);
            double t = or_mlp_forward(s->mlp, feat);
            if (t >= 0.501 && (c->tau_nn_max <= 0 || t <= c->tau_nn_max)) {
                tau = t;
                nn_ok = 1;
            }
        }
        if (!nn_ok && c->les) tau = or_les_tau(f, rho, ux, uy, tb, c->cs);
        s->tau[n] = tau;
        if (c->collision == OR_COLL_MRT)
            or_collide_mrt(f, rho, ux, uy, tau, out);
        else if (c->collision == OR_COLL_RR)
            or_collide_rr(f, rho, ux, uy, tau, out);
        else
            or_collide_bgk(f, rho, ux, uy, tau, out);
        if (ib) {
            double g[9];
            or_guo(ux, uy, s->Fx[n], s->Fy[n], tb, g);
            for (int k = 0; k < 9; k++) out[k] = out[k] + g[k];
        }
        for (int k = 0; k < 9; k++) s->fcol[(size_t)k * N + n] = out[k];
    }
}

/* Pull streaming (mesh.cpp:192-226) with the boundary kinds:
 *  OR_BC_PERIODIC_BB: periodic x (the reference's ring halo, mesh.cpp:467-484,
 *     with inlet/outlet off) and halfway bounce-back walls below row 0 and
 *     above row ny-1 (the reference declares inv[] at mesh.cpp:236 but has no
 *     bounce-back; this kind is the channel of the plain-method workload).
 *  OR_BC_INFLOW: columns 0 / nx-1 are non-equilibrium-extrapolation inlet /
 *     outlet (mesh.cpp:239-286: rho = 1, inlet u = u_in, outlet u copied from
 *     column nx-2); rows 0 / ny-1 (1 <= x <= nx-2) are specular free-slip
 *     walls (the 'm' rule, mesh.cpp:325-337, mirrored for the top row). */
static inline double fc(const or_sim* s, int k, int x, int y) {
    return s->fcol[(size_t)k * s->N + (size_t)y * s->c.nx + x];
}

static void pull_node(const or_sim* s, int x, int y, double f[9]) {
    const int nx = s->c.nx, ny = s->c.ny;
    if (s->c.bc == OR_BC_PERIODIC_BB) {
        for (int k = 0; k < 9; k++) {
            int sx = x - EX[k];
            int sy = y - EY[k];
            if (sx < 0) sx += nx;
            if (sx >= nx) sx -= nx;
            if (sy < 0 || sy >= ny)
                f[k] = fc(s, OPP[k], x, y);
            else
                f[k] = fc(s, k, sx, sy);
        }
        return;
    }
    /* OR_BC_INFLOW, non-NEE nodes */
    if (y == 0) { /* mesh.cpp:327-335 */
        f[0] = fc(s, 0, x, y);
        f[1] = fc(s, 1, x - 1, y);
        f[2] = fc(s, 4, x, y + 1);
        f[3] = fc(s, 3, x + 1, y);
        f[4] = fc(s, 4, x, y + 1);
        f[5] = fc(s, 8, x - 1, y + 1);
        f[6] = fc(s, 7, x + 1, y + 1);
        f[7] = fc(s, 7, x + 1, y + 1);
        f[8] = fc(s, 8, x - 1, y + 1);
    } else if (y == ny - 1) { /* mirror image of the 'm' rule */
        f[0] = fc(s, 0, x, y);
        f[1] = fc(s, 1, x - 1, y);
        f[2] = fc(s, 2, x, y - 1);
        f[3] = fc(s, 3, x + 1, y);
        f[4] = fc(s, 2, x, y - 1);
        f[5] = fc(s, 5, x - 1, y - 1);
        f[6] = fc(s, 6, x + 1, y - 1);
        f[7] = fc(s, 6, x + 1, y - 1);
        f[8] = fc(s, 5, x - 1, y - 1);
    } else {
        for (int k = 0; k < 9; k++) f[k] = fc(s, k, x - EX[k], y - EY[k]);
    }
}

/* fnew and macros of node nb by non-equilibrium extrapolation from node nn
 * with boundary density rb and velocity (uxb, uyb): f = feq_b + f_n - feq_n
 * (the form of mesh.cpp:245-250, 279-284, 311-318, 343-349). */
static void nee_node(or_sim* s, int nb, int nn, double rb, double uxb, double uyb) {
    const size_t N = (size_t)s->N;
    double fb[9], fq[9];
    or_feq(rb, uxb, uyb, fb);
    or_feq(s->rho[nn], s->ux[nn], s->uy[nn], fq);
    for (int k = 0; k < 9; k++) s->fnew[k * N + nb] = fb[k] + s->fnew[k * N + nn] - fq[k];
    s->rho[nb] = rb;
    s->ux[nb] = uxb;
    s->uy[nb] = uyb;
}

/* Flat plate: the reference's own boundary set for one domain
 * (Mesh::BoundaryCondition, mesh.cpp:228-356, with the styles of
 * Mesh::JudgeStyle, mesh.cpp:1391-1454); x is the reference's global
 * column pos.x (mesh.cpp:105). Pull streaming on the 'f' nodes
 * 1 <= x <= nx-2, 1 <= y <= ny-2 (mesh.cpp:192-226), then in the
 * reference's order:
 *   inlet  x = 0,    1 <= y <= ny-2: rho = 1, u = u_in, NEE from (1, y)
 *          (mesh.cpp:239-253; the reference writes u = (0.1, 0))
 *   outlet x = nx-1, 1 <= y <= ny-2: rho = 1, u = u(nx-2, y), NEE from (nx-2, y)
 *          (mesh.cpp:273-287)
 *   top    y = ny-1, 0 <= x <= nx-1: rho = 1, u = u(x, ny-2), NEE from (x, ny-2)
 *          (mesh.cpp:305-319; at x = 0 and nx-1 the source is the inlet /
 *          outlet node just set)
 *   bottom y = 0,    1 <= x <= nx-2: x < plate_x: specular 'm' rule
 *          (mesh.cpp:325-337); else 'd': rho = rho(x, 1), u = 0, NEE from
 *          (x, 1) (mesh.cpp:338-350)
 * The corners (0, 0) and (nx-1, 0) carry the inlet / outlet style, which
 * no rule of the bottom loop handles: their fnew and macros never change
 * (they are still collided every step). */
static void stream_flatplate(or_sim* s) {
    const int nx = s->c.nx, ny = s->c.ny;
    const size_t N = (size_t)s->N;
#pragma omp parallel for schedule(static) if (N > 4096) num_threads(g_threads > 0 ? g_threads : 1)
    for (int y = 1; y < ny - 1; y++)
        for (int x = 1; x < nx - 1; x++) {
            double f[9], m[3];
            for (int k = 0; k < 9; k++) f[k] = fc(s, k, x - EX[k], y - EY[k]);
            or_macro(f, m);
            const int n = y * nx + x;
            for (int k = 0; k < 9; k++) s->fnew[k * N + n] = f[k];
            s->rho[n] = m[0];
            s->ux[n] = m[1];
            s->uy[n] = m[2];
        }
    for (int y = 1; y < ny - 1; y++) nee_node(s, y * nx, y * nx + 1, 1.0, s->c.u_in_x, s->c.u_in_y);
    for (int y = 1; y < ny - 1; y++) {
        const int nn = y * nx + nx - 2;
        nee_node(s, y * nx + nx - 1, nn, 1.0, s->ux[nn], s->uy[nn]);
    }
    for (int x = 0; x < nx; x++) {
        const int nn = (ny - 2) * nx + x;
        nee_node(s, (ny - 1) * nx + x, nn, 1.0, s->ux[nn], s->uy[nn]);
    }
    for (int x = 1; x < nx - 1; x++) {
        if (x < s->c.plate_x) {
            double f[9], m[3];
            f[0] = fc(s, 0, x, 0);
            f[1] = fc(s, 1, x - 1, 0);
            f[2] = fc(s, 4, x, 1);
            f[3] = fc(s, 3, x + 1, 0);
            f[4] = fc(s, 4, x, 1);
            f[5] = fc(s, 8, x - 1, 1);
            f[6] = fc(s, 7, x + 1, 1);
            f[7] = fc(s, 7, x + 1, 1);
            f[8] = fc(s, 8, x - 1, 1);
            or_macro(f, m);
            for (int k = 0; k < 9; k++) s->fnew[k * N + x] = f[k];
            s->rho[x] = m[0];
            s->ux[x] = m[1];
            s->uy[x] = m[2];
        } else {
            nee_node(s, x, nx + x, s->rho[nx + x], 0.0, 0.0);
        }
    }
}

static void stream_phase(or_sim* s) {
    const int nx = s->c.nx, ny = s->c.ny, N = s->N;
    if (s->c.bc == OR_BC_FLATPLATE) {
        stream_flatplate(s);
        return;
    }
    const int inflow = s->c.bc == OR_BC_INFLOW;
#pragma omp parallel for schedule(static) if (N > 4096) num_threads(g_threads > 0 ? g_threads : 1)
    for (int y = 0; y < ny; y++) {
        for (int x = 0; x < nx; x++) {
            if (inflow && (x == 0 || x == nx - 1)) continue;
            double f[9], m[3];
            pull_node(s, x, y, f);
            or_macro(f, m);
            int n = y * nx + x;
            for (int k = 0; k < 9; k++) s->fnew[(size_t)k * N + n] = f[k];
            s->rho[n] = m[0];
            s->ux[n] = m[1];
            s->uy[n] = m[2];
        }
    }
    if (!inflow) return;
    /* inlet / outlet NEE, mesh.cpp:239-286 */
    for (int y = 0; y < ny; y++) {
        for (int side = 0; side < 2; side++) {
            int xb = side == 0 ? 0 : nx - 1;
            int xn = side == 0 ? 1 : nx - 2;
            int nb = y * nx + xb, nn = y * nx + xn;
            double rb = 1.0, uxb, uyb;
            if (side == 0) {
                uxb = s->c.u_in_x;
                uyb = s->c.u_in_y;
            } else {
                uxb = s->ux[nn];
                uyb = s->uy[nn];
            }
            double fb[9], fq[9];
            or_feq(rb, uxb, uyb, fb);
            or_feq(s->rho[nn], s->ux[nn], s->uy[nn], fq);
            for (int k = 0; k < 9; k++)
                s->fnew[(size_t)k * N + nb] = fb[k] + s->fnew[(size_t)k * N + nn] - fq[k];
            s->rho[nb] = rb;
            s->ux[nb] = uxb;
            s->uy[nb] = uyb;
        }
    }
}

/* One time step in the reference's OneLBMStep order (mesh.cpp:357-513),
 * with the immersed boundary wired before the collision. */
int or_step(or_sim* s, int nsteps) {
    for (int t = 0; t < nsteps; t++) {
        if (s->bnn_tau) bnn_phase(s);
        if (s->nbody > 0) ib_phase(s);
        collide_phase(s);
        stream_phase(s);
        /* ErrorCatch, mesh.cpp:2101-2111 (NaN also caught here) */
        for (int n = 0; n < s->N; n++) {
            double r = s->rho[n];
            if (!(r >= 0.0 && r <= 10.0)) return 1;
        }
    }
    return 0;
}

void or_get_fields(const or_sim* s, double* rho, double* ux, double* uy) {
    size_t N = (size_t)s->N;
    if (rho) memcpy(rho, s->rho, N * sizeof(double));
    if (ux) memcpy(ux, s->ux, N * sizeof(double));
    if (uy) memcpy(uy, s->uy, N * sizeof(double));
}
void or_get_fnew(const or_sim* s, double* f) { memcpy(f, s->fnew, 9 * (size_t)s->N * sizeof(double)); }
void or_get_fcol(const or_sim* s, double* f) { memcpy(f, s->fcol, 9 * (size_t)s->N * sizeof(double)); }
void or_get_tau(const or_sim* s, double* t) { memcpy(t, s->tau, (size_t)s->N * sizeof(double)); }
int or_num_force_records(const or_sim* s) { return s->nrec; }
void or_get_forces(const or_sim* s, double* out) {
    memcpy(out, s->forces, sizeof(double) * 2 * (size_t)s->nrec * s->nbody);
}
int or_num_markers(const or_sim* s) { return s->nmk; }
void or_get_markers(const or_sim* s, double* out) {
    for (int i = 0; i < s->nmk; i++) {
        out[6 * i + 0] = s->mk[i].x;
        out[6 * i + 1] = s->mk[i].y;
        out[6 * i + 2] = s->mk[i].uix;
        out[6 * i + 3] = s->mk[i].uiy;
        out[6 * i + 4] = s->mk[i].fx;
        out[6 * i + 5] = s->mk[i].fy;
    }
}
double or_mass(const or_sim* s) {
    double m = 0.0;
    for (int n = 0; n < s->N; n++) m += s->rho[n];
    return m;
}
void or_get_force_field(const or_sim* s, double* fx, double* fy) {
    memcpy(fx, s->Fx, sizeof(double) * s->N);
    memcpy(fy, s->Fy, sizeof(double) * s->N);
}

/* ------------------------------------------------------------------ */
/* Discrete adjoint of the BGK + IB step (periodic x, bounce-back walls),  */
/* objective J = mean over steps [avg_from, nsteps] of the body's drag     */
/* record (minus the sum of the marker forces, x component).               */
/* Gradients: dJ/dtau (per node base relaxation time) and dJ/d(cx, cy).    */
/* Forward model = or_step with tau_mode FIELD, les off, one body.         */
/* Every stage is transposed by hand; states are kept for all steps.       */
/* ------------------------------------------------------------------ */
static double dphi(double x) { /* derivative of or_delta */
    double ax = fabs(x), sg = x < 0 ? -1.0 : 1.0;
    if (ax < 1) {
        double R = 1.0 + 4.0 * ax - 4 * (x * x);
        return sg * (1.0 / 8.0) * (-2.0 + (4.0 - 8.0 * ax) / (2.0 * sqrt(R)));
    } else if (ax < 2) {
        double R = -7.0 + 12.0 * ax - 4 * (x * x);
        return sg * (1.0 / 8.0) * (-2.0 - (12.0 - 8.0 * ax) / (2.0 * sqrt(R)));
    }
    return 0.0;
}

int or_adjoint_drag(int nx, int ny, uint64_t seed, double ux0, double uy0, double noise, const double* tau,
                    double cx, double cy, double diam, int M, int nsteps, int avg_from, double* J, double* g_tau,
                    double* g_c) {
    or_config c = {nx, ny, OR_COLL_BGK, OR_BC_PERIODIC_BB, 0, OR_TAU_FIELD, 0.6, 0.0, 0.0, 0.0, 0.0, 0};
    or_sim* s = or_create(&c);
    or_init_fields(s, seed, 1.0, ux0, uy0, noise);
    or_set_tau_field(s, tau);
    or_add_body(s, cx, cy, diam, M, 0.0, 0.0);
    const int N = s->N;
    const size_t SZ = 9 * (size_t)N;
    /* states fnew^t and the fields u^t, rho^t for t = 0..nsteps-1 */
    double* F = (double*)malloc(sizeof(double) * SZ * nsteps);
    double* R = (double*)malloc(sizeof(double) * 3 * (size_t)N * nsteps);
    const int nwin = nsteps - avg_from + 1;
    double Jv = 0.0;
    for (int t = 0; t < nsteps; t++) {
        memcpy(F + SZ * t, s->fnew, sizeof(double) * SZ);
        memcpy(R + 3 * (size_t)N * t, s->rho, sizeof(double) * N);
        memcpy(R + 3 * (size_t)N * t + N, s->ux, sizeof(double) * N);
        memcpy(R + 3 * (size_t)N * t + 2 * N, s->uy, sizeof(double) * N);
        or_step(s, 1);
        if (t + 1 >= avg_from) Jv += s->forces[(size_t)t * 2] / nwin;
    }
    if (J) *J = Jv;
    double* lam = (double*)calloc(SZ, sizeof(double));   /* adjoint of fnew^{t+1} */
    double* lamn = (double*)calloc(SZ, sizeof(double));  /* adjoint of fnew^t */
    double* lc = (double*)calloc(SZ, sizeof(double));    /* adjoint of fcol^t */
    double* lFx = (double*)calloc(N, sizeof(double));    /* adjoint of Eulerian force */
    double* lFy = (double*)calloc(N, sizeof(double));
    double* luIx = (double*)calloc(N, sizeof(double));   /* velocity adjoint from interpolation */
    double* luIy = (double*)calloc(N, sizeof(double));
    double* Ex = (double*)calloc(N, sizeof(double));     /* Eulerian force of the step */
    double* Ey = (double*)calloc(N, sizeof(double));
    double* mUx = (double*)malloc(sizeof(double) * M);
    double* mUy = (double*)malloc(sizeof(double) * M);
    double* mFx = (double*)malloc(sizeof(double) * M);
    double* mFy = (double*)malloc(sizeof(double) * M);
    if (g_tau) memset(g_tau, 0, sizeof(double) * N);
    double gcx = 0.0, gcy = 0.0;
    or_marker* mk = s->mk;
    for (int t = nsteps - 1; t >= 0; t--) {
        const double* f = F + SZ * t;
        const double* rho = R + 3 * (size_t)N * t;
        const double* ux = rho + N;
        const double* uy = rho + 2 * N;
        const double cJ = (t + 1 >= avg_from) ? 1.0 / nwin : 0.0; /* weight of this step's drag */
        /* recompute the IB forward of this step */
        memset(Ex, 0, sizeof(double) * N);
        memset(Ey, 0, sizeof(double) * N);
        for (int m = 0; m < M; m++) {
            double Ux = 0, Uy = 0;
            int XI = (int)mk[m].x, YI = (int)mk[m].y;
            for (int X = XI - 1; X <= XI + 2; X++)
                for (int Y = YI - 1; Y <= YI + 2; Y++) {
                    double D = or_delta(mk[m].x - X) * or_delta(mk[m].y - Y);
                    Ux += ux[Y * nx + X] * D;
                    Uy += uy[Y * nx + X] * D;
                }
            mUx[m] = Ux;
            mUy[m] = Uy;
            mFx[m] = 2 * (0.0 - Ux);
            mFy[m] = 2 * (0.0 - Uy);
            for (int X = XI - 1; X <= XI + 2; X++)
                for (int Y = YI - 1; Y <= YI + 2; Y++) {
                    double D = or_delta(mk[m].x - X) * or_delta(mk[m].y - Y);
                    Ex[Y * nx + X] += mFx[m] * D;
                    Ey[Y * nx + X] += mFy[m] * D;
                }
        }
        /* streaming transpose: lc_j(m) = lam_j(m + e_j), or lam_opp(j)(m) when m + e_j leaves the walls */
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                int n = y * nx + x;
                for (int j = 0; j < 9; j++) {
                    int yd = y + EY[j], xd = (x + EX[j] + nx) % nx;
                    lc[(size_t)j * N + n] = (yd < 0 || yd >= ny) ? lam[(size_t)OPP[j] * N + n]
                                                                 : lam[(size_t)j * N + yd * nx + xd];
                }
            }
        /* collision transpose, first for the force adjoint (needed by the markers) */
        for (int n = 0; n < N; n++) {
            double r = rho[n], tb = tau[n], om = 1.0 / tb, ct = 1.0 - 0.5 * om;
            double us_x = ux[n] + 0.5 * Ex[n] / r, us_y = uy[n] + 0.5 * Ey[n] / r;
            double lux = 0, luy = 0, lfx = 0, lfy = 0;
            for (int k = 0; k < 9; k++) {
                double l = lc[(size_t)k * N + n];
                double eu = EX[k] * us_x + EY[k] * us_y;
                lux += l * (om * WEIGHT[k] * r * (3.0 * EX[k] + 9.0 * eu * EX[k] - 3.0 * us_x) +
                            WEIGHT[k] * ct * (-3.0 * Ex[n] + 9.0 * EX[k] * (EX[k] * Ex[n] + EY[k] * Ey[n])));
                luy += l * (om * WEIGHT[k] * r * (3.0 * EY[k] + 9.0 * eu * EY[k] - 3.0 * us_y) +
                            WEIGHT[k] * ct * (-3.0 * Ey[n] + 9.0 * EY[k] * (EX[k] * Ex[n] + EY[k] * Ey[n])));
                lfx += l * WEIGHT[k] * ct * (3.0 * (EX[k] - us_x) + 9.0 * eu * EX[k]);
                lfy += l * WEIGHT[k] * ct * (3.0 * (EY[k] - us_y) + 9.0 * eu * EY[k]);
            }
            lFx[n] = lfx + lux / (2.0 * r);
            lFy[n] = lfy + luy / (2.0 * r);
        }
        /* markers: adjoint of F_m (spread + drag), of U_m, position gradient */
        memset(luIx, 0, sizeof(double) * N);
        memset(luIy, 0, sizeof(double) * N);
        for (int m = 0; m < M; m++) {
            int XI = (int)mk[m].x, YI = (int)mk[m].y;
            double lFmx = -cJ, lFmy = 0.0; /* d(drag)/dF_m = -1 on x */
            for (int X = XI - 1; X <= XI + 2; X++)
                for (int Y = YI - 1; Y <= YI + 2; Y++) {
                    double D = or_delta(mk[m].x - X) * or_delta(mk[m].y - Y);
                    lFmx += D * lFx[Y * nx + X];
                    lFmy += D * lFy[Y * nx + X];
                }
            double lUx = -2.0 * lFmx, lUy = -2.0 * lFmy;
            for (int X = XI - 1; X <= XI + 2; X++)
                for (int Y = YI - 1; Y <= YI + 2; Y++) {
                    int n = Y * nx + X;
                    double px = or_delta(mk[m].x - X), py = or_delta(mk[m].y - Y);
                    double D = px * py, Dx = dphi(mk[m].x - X) * py, Dy = px * dphi(mk[m].y - Y);
                    luIx[n] += D * lUx;
                    luIy[n] += D * lUy;
                    /* d/dx_m of U_m (interpolation) and of F(n) (spreading) */
                    gcx += lUx * Dx * ux[n] + lUy * Dx * uy[n] + Dx * (lFx[n] * mFx[m] + lFy[n] * mFy[m]);
                    gcy += lUx * Dy * ux[n] + lUy * Dy * uy[n] + Dy * (lFx[n] * mFx[m] + lFy[n] * mFy[m]);
                }
        }
        /* collision and macro transpose -> adjoint of fnew^t; tau gradient */
        for (int n = 0; n < N; n++) {
            double r = rho[n], tb = tau[n], om = 1.0 / tb, ct = 1.0 - 0.5 * om;
            double us_x = ux[n] + 0.5 * Ex[n] / r, us_y = uy[n] + 0.5 * Ey[n] / r;
            double feq[9], fc[9];
            or_feq(r, us_x, us_y, feq);
            for (int k = 0; k < 9; k++) fc[k] = f[(size_t)k * N + n];
            double lux = 0, luy = 0, lr = 0, lt = 0;
            for (int k = 0; k < 9; k++) {
                double l = lc[(size_t)k * N + n];
                double eu = EX[k] * us_x + EY[k] * us_y;
                double S = WEIGHT[k] * ct * (3.0 * ((EX[k] - us_x) * Ex[n] + (EY[k] - us_y) * Ey[n]) +
                                             9.0 * eu * (EX[k] * Ex[n] + EY[k] * Ey[n]));
                lux += l * (om * WEIGHT[k] * r * (3.0 * EX[k] + 9.0 * eu * EX[k] - 3.0 * us_x) +
                            WEIGHT[k] * ct * (-3.0 * Ex[n] + 9.0 * EX[k] * (EX[k] * Ex[n] + EY[k] * Ey[n])));
                luy += l * (om * WEIGHT[k] * r * (3.0 * EY[k] + 9.0 * eu * EY[k] - 3.0 * us_y) +
                            WEIGHT[k] * ct * (-3.0 * Ey[n] + 9.0 * EY[k] * (EX[k] * Ex[n] + EY[k] * Ey[n])));
                lr += l * om * feq[k] / r;
                lt += l * ((fc[k] - feq[k]) * om * om + S * 0.5 * om * om / ct);
            }
            /* u* = u + F/(2 rho) */
            lr += -(lux * Ex[n] + luy * Ey[n]) / (2.0 * r * r);
            lux += luIx[n];
            luy += luIy[n];
            for (int k = 0; k < 9; k++)
                lamn[(size_t)k * N + n] = (1.0 - om) * lc[(size_t)k * N + n] + lr +
                                          (lux * (EX[k] - ux[n]) + luy * (EY[k] - uy[n])) / r;
            if (g_tau) g_tau[n] += lt;
        }
        double* tmp = lam;
        lam = lamn;
        lamn = tmp;
    }
    if (g_c) {
        g_c[0] = gcx;
        g_c[1] = gcy;
    }
    free(F); free(R); free(lam); free(lamn); free(lc); free(lFx); free(lFy); free(luIx); free(luIy);
    free(Ex); free(Ey); free(mUx); free(mUy); free(mFx); free(mFy);
    or_destroy(s);
    return 0;
}

double or_drag_objective(int nx, int ny, uint64_t seed, double ux0, double uy0, double noise, const double* tau,
                         double cx, double cy, double diam, int M, int nsteps, int avg_from) {
    double J = 0.0;
    or_config c = {nx, ny, OR_COLL_BGK, OR_BC_PERIODIC_BB, 0, OR_TAU_FIELD, 0.6, 0.0, 0.0, 0.0, 0.0, 0};
    or_sim* s = or_create(&c);
    or_init_fields(s, seed, 1.0, ux0, uy0, noise);
    or_set_tau_field(s, tau);
    or_add_body(s, cx, cy, diam, M, 0.0, 0.0);
    const int nwin = nsteps - avg_from + 1;
    for (int t = 0; t < nsteps; t++) {
        or_step(s, 1);
        if (t + 1 >= avg_from) J += s->forces[(size_t)t * 2] / nwin;
    }
    or_destroy(s);
    return J;
}

/* FindAndWriteTau's search loop, IPPI-LBM/mesh.cpp:2667-2704, for each of
 * n nodes: candidate fcol_k = (1 - 1/(tau_k + dtau)) fnew_k + feq_k /
 * (tau_k + dtau) in the reference's operation order, macros summed over
 * k = 0..8, the score |rho - rho*| + |jx - u*| + |jy - v*| summed left to
 * right, kept when strictly below the best so far. */
void or_tau_search(int n, const double* fnew, const double* feq, const double* tau, const double* target, int m_lo,
                   int m_hi, double step, double* best_dtau, double* best_err) {
    for (int i = 0; i < n; i++) {
        const double* fn = fnew + 9 * (size_t)i;
        const double* fq = feq + 9 * (size_t)i;
        const double* tk = tau + 9 * (size_t)i;
        const double rt = target[3 * i], ut = target[3 * i + 1], vt = target[3 * i + 2];
        double bestd = 1e10, beste = 1e10;
        for (int m = m_lo; m <= m_hi; m++) {
            const double dt = (double)m * step;
            double fc[9];
            for (int k = 0; k < 9; k++) fc[k] = (1.0 - 1.0 / (tk[k] + dt)) * fn[k] + fq[k] / (tk[k] + dt);
            double r = 0, u = 0, v = 0;
            for (int k = 0; k < 9; k++) {
                r += fc[k];
                u += fc[k] * EX[k];
                v += fc[k] * EY[k];
            }
            const double err = fabs(r - rt) + fabs(u - ut) + fabs(v - vt);
            if (err < beste) {
                beste = err;
                bestd = dt;
            }
        }
        best_dtau[i] = bestd;
        best_err[i] = beste;
    }
}
