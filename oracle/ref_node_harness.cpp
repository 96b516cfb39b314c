// Thin C entry points around the reference's own Lattice methods, compiled
// together with the unmodified IPPI-LBM/NodeAndLattice.cpp by
// oracle/Makefile (target `ref`) into oracle/_ref/libippi_refnode.so.
// Used only by tests/golden/make_golden.py to write golden vectors that pin
// the oracle's per-node functions. Nothing here is product code.
#include "NodeAndLattice.h"
#include <cstring>

static void load(Lattice& L, const double* f, double rho, double ux, double uy, double tau) {
    L.fnew.assign(f, f + 9);
    L.fcol.assign(9, 0.0);
    L.feq.assign(9, 0.0);
    L.rho = rho;
    L.u.x = ux;
    L.u.y = uy;
    L.tau = tau;
    L.TauAtBeginning = tau;
}

extern "C" {
void ref_feq(double rho, double ux, double uy, double* out) {
    Lattice L;
    L.rho = rho; L.u.x = ux; L.u.y = uy;
    L.Compute_feq();                     // NodeAndLattice.cpp:124-149
    std::memcpy(out, L.feq.data(), 9 * sizeof(double));
}
void ref_feq_init(double rho, double ux, double uy, double* out) {
    Lattice L;
    L.rho = rho; L.u.x = ux; L.u.y = uy;
    L.InitialLatticeFeq();               // NodeAndLattice.cpp:69-122
    std::memcpy(out, L.fcol.data(), 9 * sizeof(double));
}
void ref_macro(const double* f, double* out) {
    Lattice L;
    L.fnew.assign(f, f + 9);
    L.CalMacro();                        // NodeAndLattice.cpp:196-226
    out[0] = L.rho; out[1] = L.u.x; out[2] = L.u.y;
}
void ref_bgk(const double* f, double rho, double ux, double uy, double tau, double* out) {
    Lattice L; load(L, f, rho, ux, uy, tau);
    L.ComputeBGKCollision();             // NodeAndLattice.cpp:150-161
    std::memcpy(out, L.fcol.data(), 9 * sizeof(double));
}
void ref_mrt(const double* f, double rho, double ux, double uy, double tau, double* out) {
    Lattice L; load(L, f, rho, ux, uy, tau);
    L.ComputeMRTCollision();             // NodeAndLattice.cpp:290-360
    std::memcpy(out, L.fcol.data(), 9 * sizeof(double));
}
void ref_rr(const double* f, double rho, double ux, double uy, double tau, double* out) {
    Lattice L; load(L, f, rho, ux, uy, tau);
    L.ComputeRRCollision();              // NodeAndLattice.cpp:228-286
    std::memcpy(out, L.fcol.data(), 9 * sizeof(double));
}
void ref_guo(double ux, double uy, double fx, double fy, double tau_begin, double* out) {
    Lattice L;
    L.u.x = ux; L.u.y = uy; L.force.x = fx; L.force.y = fy;
    L.TauAtBeginning = tau_begin;
    L.Compute_Disforce();                // NodeAndLattice.cpp:162-195
    std::memcpy(out, L.lat_force.data(), 9 * sizeof(double));
}
}
