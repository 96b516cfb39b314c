// Discrete adjoint of one BGK + immersed-boundary time step, node-local
// parts (the reference has no gradient code; this differentiates the step
// that the forward kernels compute: NodeAndLattice.cpp:150-195 collision and
// Guo source, mesh.cpp:192-226 streaming, mesh.cpp:514-741, 1244-1281 IB).
//
// Forward map of one step at a node, with the IB force E of the node (zero
// outside the marker supports), base relaxation time tb, omega = 1/tb and
// c = 1 - omega/2:
//     rho = sum f,  u = sum e f / rho,  us = u + E / (2 rho)
//     fcol_k = f_k - omega (f_k - feq_k(rho, us)) + S_k(us, E)
//     S_k = w_k c [3 (e_k - us).E + 9 (e_k.us)(e_k.E)]
// Given the adjoint l of fcol at the node, everything the transpose needs is
// linear in six weighted moments of l:
//     W0 = sum w l, Wa = sum w l e_a, Wab = sum w l e_a e_b.
#pragma once
#include "lbm_node.hpp"

namespace lbm {
namespace adj {

template <typename T>
struct WMom {
    T w0, wx, wy, wxx, wxy, wyy;
};

template <typename T>
__device__ __forceinline__ WMom<T> wmoments(const T* l) {
    const T a = l[1] + l[3], b = l[2] + l[4], c = (l[5] + l[7]) + (l[6] + l[8]);
    const T w1 = T(1) / T(9), w5 = T(1) / T(36);
    WMom<T> m;
    m.w0 = T(4) / T(9) * l[0] + w1 * (a + b) + w5 * c;
    m.wx = w1 * (l[1] - l[3]) + w5 * ((l[5] - l[7]) + (l[8] - l[6]));
    m.wy = w1 * (l[2] - l[4]) + w5 * ((l[5] - l[7]) + (l[6] - l[8]));
    m.wxx = w1 * a + w5 * c;
    m.wyy = w1 * b + w5 * c;
    m.wxy = w5 * ((l[5] + l[7]) - (l[6] + l[8]));
    return m;
}

// Forward quantities of the node the transpose is evaluated at.
template <typename T>
struct NodeFwd {
    T rho, ux, uy;  // macroscopic (before the IB velocity correction)
    T fx, fy;       // IB force of the node
    T om, ct, dl;   // 1/tb, 1 - 1/(2 tb) = dl/2, dl = 2 - 1/tb
};

// Relaxation quantities from tx = tb - 1/2 (dl keeps its digits near
// tb = 1/2, where omega and 1 - omega/2 computed from tb would not).
template <typename T>
__device__ __forceinline__ NodeFwd<T> node_fwd(T rho, T ux, T uy, T fx, T fy, T tx) {
    const T dl = (T(4) * tx) / (T(1) + T(2) * tx);
    return NodeFwd<T>{rho, ux, uy, fx, fy, T(2) - dl, T(0.5) * dl, dl};
}

// Adjoint of the corrected velocity us (through feq and S) and of the force
// through S alone.
template <typename T>
struct UsAdj {
    T usx, usy, aux, auy, afx, afy;
};

template <typename T>
__device__ __forceinline__ UsAdj<T> us_adjoint(const WMom<T>& m, const NodeFwd<T>& s, bool forced) {
    UsAdj<T> r;
    const T h = T(0.5) / s.rho;
    r.usx = forced ? s.ux + s.fx * h : s.ux;
    r.usy = forced ? s.uy + s.fy * h : s.uy;
    const T qx = m.wx + T(3) * (r.usx * m.wxx + r.usy * m.wxy) - r.usx * m.w0;  // (sum w l (e_x + 3(e.us) e_x - us_x))/3
    const T qy = m.wy + T(3) * (r.usx * m.wxy + r.usy * m.wyy) - r.usy * m.w0;
    r.aux = T(3) * s.om * s.rho * qx;
    r.auy = T(3) * s.om * s.rho * qy;
    r.afx = T(0);
    r.afy = T(0);
    if (forced) {
        r.aux += s.ct * (T(-3) * s.fx * m.w0 + T(9) * (s.fx * m.wxx + s.fy * m.wxy));
        r.auy += s.ct * (T(-3) * s.fy * m.w0 + T(9) * (s.fx * m.wxy + s.fy * m.wyy));
        r.afx = T(3) * s.ct * qx;
        r.afy = T(3) * s.ct * qy;
    }
    return r;
}

// Adjoint of the node's IB force E (through S and through us).
template <typename T>
__device__ __forceinline__ void force_adjoint(const UsAdj<T>& a, const NodeFwd<T>& s, T& lfx, T& lfy) {
    const T h = T(0.5) / s.rho;
    lfx = a.afx + a.aux * h;
    lfy = a.afy + a.auy * h;
}

// Adjoint of the node's pre-collision populations (nu) and the derivative
// of sum l fcol with respect to tb (dt). d[k] = f_k - feq_k(rho, us) in the
// representation of the stored populations; (lix, liy) = adjoint of u from
// the IB interpolation (zero off the supports).
template <typename T>
__device__ __forceinline__ void node_adjoint(const T* l, const T* d, const WMom<T>& m, const UsAdj<T>& a,
                                             const NodeFwd<T>& s, bool forced, T lix, T liy, T* nu, T& dt) {
    const T us2 = a.usx * a.usx + a.usy * a.usy;
    // sum l feq / rho
    const T p = m.w0 * (T(1) - T(1.5) * us2) + T(3) * (a.usx * m.wx + a.usy * m.wy) +
                T(4.5) * (a.usx * a.usx * m.wxx + T(2) * a.usx * a.usy * m.wxy + a.usy * a.usy * m.wyy);
    T lr = s.om * p;
    T sd = T(0);
#pragma unroll
    for (int k = 0; k < 9; k++) sd += l[k] * d[k];
    dt = s.om * s.om * sd;
    if (forced) {
        const T ir = T(1) / s.rho;
        lr -= T(0.5) * (a.aux * s.fx + a.auy * s.fy) * ir * ir;
        // sum l S / c
        const T se = T(3) * (m.wx * s.fx + m.wy * s.fy) - T(3) * m.w0 * (a.usx * s.fx + a.usy * s.fy) +
                     T(9) * (a.usx * s.fx * m.wxx + (a.usx * s.fy + a.usy * s.fx) * m.wxy + a.usy * s.fy * m.wyy);
        dt += T(0.5) * s.om * s.om * se;
    }
    const T lux = a.aux + lix, luy = a.auy + liy;
    const T ir = T(1) / s.rho;
    const T vx = lux * ir, vy = luy * ir;
    const T base = lr - (vx * s.ux + vy * s.uy);
    // (1 - omega) l = dl l - l
#pragma unroll
    for (int k = 0; k < 9; k++) nu[k] = (s.dl * l[k] - l[k]) + base + (T)EX[k] * vx + (T)EY[k] * vy;
}

// Derivative of the 4-point delta (delta4).
__host__ __device__ inline double ddelta4(double x) {
    const double ax = x < 0 ? -x : x, sg = x < 0 ? -1.0 : 1.0;
    if (ax < 1) {
        const double R = 1.0 + 4.0 * ax - 4 * (x * x);
        return sg * 0.125 * (-2.0 + (4.0 - 8.0 * ax) / (2.0 * sqrt(R)));
    } else if (ax < 2) {
        const double R = -7.0 + 12.0 * ax - 4 * (x * x);
        return sg * 0.125 * (-2.0 - (12.0 - 8.0 * ax) / (2.0 * sqrt(R)));
    }
    return 0.0;
}

}  // namespace adj
}  // namespace lbm
