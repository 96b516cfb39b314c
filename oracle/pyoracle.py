"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
always as the checker or the CPU baseline, never as the product path.
The library is oracle/_build/libippi_oracle.so, built by `make -C oracle`.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libippi_oracle.so")
REF_LIB_PATH = os.path.join(_HERE, "_ref", "libippi_refnode.so")

COLL = {"bgk": 0, "mrt": 1, "rr": 2}
BC = {"periodic_bb": 0, "inflow": 1, "flatplate": 2}
TAU = {"const": 0, "field": 1, "nn": 2}
MLP_NPARAM = 6 * 32 + 32 + 32 * 32 + 32 + 32 + 1
MLP_BLOB = MLP_NPARAM + 12


class OrConfig(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("collision", ctypes.c_int),
        ("bc", ctypes.c_int), ("les", ctypes.c_int), ("tau_mode", ctypes.c_int),
        ("tau0", ctypes.c_double), ("cs", ctypes.c_double),
        ("u_in_x", ctypes.c_double), ("u_in_y", ctypes.c_double),
        ("tau_nn_max", ctypes.c_double), ("plate_x", ctypes.c_int),
    ]


_lib = None
_D = ctypes.POINTER(ctypes.c_double)


def _p(a):
    return a.ctypes.data_as(_D)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
        L = ctypes.CDLL(LIB_PATH)
        d, i, vp = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
        L.or_feq.argtypes = [d, d, d, _D]
        L.or_feq_init.argtypes = [d, d, d, _D]
        L.or_macro.argtypes = [_D, _D]
        for n in ("or_collide_bgk", "or_collide_mrt", "or_collide_rr"):
            getattr(L, n).argtypes = [_D, d, d, d, d, _D]
        L.or_guo.argtypes = [d, d, d, d, d, _D]
        L.or_les_tau.argtypes = [_D, d, d, d, d, d]
        L.or_les_tau.restype = d
        L.or_delta.argtypes = [d]
        L.or_delta.restype = d
        L.or_noise.argtypes = [ctypes.c_uint64, ctypes.c_int64, i]
        L.or_noise.restype = d
        L.or_mlp_features.argtypes = [_D, d, d, d, _D]
        L.or_bnn_forward.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
        L.or_bnn_forward.restype = ctypes.c_float
        L.or_set_bnn.argtypes = [vp, i, _D]
        L.or_set_bnn.restype = i
        L.or_mlp_forward.argtypes = [_D, _D]
        L.or_mlp_forward.restype = d
        L.or_mlp_loss_grad.argtypes = [_D, _D, _D, i, _D]
        L.or_mlp_loss_grad.restype = d
        L.or_create.argtypes = [ctypes.POINTER(OrConfig)]
        L.or_create.restype = vp
        L.or_destroy.argtypes = [vp]
        L.or_set_threads.argtypes = [i]
        L.or_init_fields.argtypes = [vp, ctypes.c_uint64, d, d, d, d]
        L.or_set_state.argtypes = [vp, _D, _D, _D, _D]
        L.or_add_body.argtypes = [vp, d, d, d, i, d, d]
        L.or_add_body.restype = i
        L.or_set_tau_field.argtypes = [vp, _D]
        L.or_set_mlp.argtypes = [vp, _D]
        L.or_step.argtypes = [vp, i]
        L.or_step.restype = i
        L.or_get_fields.argtypes = [vp, _D, _D, _D]
        L.or_get_fnew.argtypes = [vp, _D]
        L.or_get_fcol.argtypes = [vp, _D]
        L.or_get_tau.argtypes = [vp, _D]
        L.or_num_force_records.argtypes = [vp]
        L.or_num_force_records.restype = i
        L.or_get_forces.argtypes = [vp, _D]
        L.or_num_markers.argtypes = [vp]
        L.or_num_markers.restype = i
        L.or_get_markers.argtypes = [vp, _D]
        L.or_mass.argtypes = [vp]
        L.or_mass.restype = d
        L.or_get_force_field.argtypes = [vp, _D, _D]
        L.or_adjoint_drag.argtypes = [i, i, ctypes.c_uint64, d, d, d, _D, d, d, d, i, i, i, _D, _D, _D]
        L.or_adjoint_drag.restype = i
        L.or_drag_objective.argtypes = [i, i, ctypes.c_uint64, d, d, d, _D, d, d, d, i, i, i]
        L.or_drag_objective.restype = d
        L.or_tau_search.argtypes = [i, _D, _D, _D, _D, i, i, d, _D, _D]
        _lib = L
    return _lib


# ---- per-node functions -------------------------------------------------
def feq(rho, ux, uy):
    out = np.zeros(9)
    lib().or_feq(rho, ux, uy, _p(out))
    return out


def feq_init(rho, ux, uy):
    out = np.zeros(9)
    lib().or_feq_init(rho, ux, uy, _p(out))
    return out


def macro(f):
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.zeros(3)
    lib().or_macro(_p(f), _p(out))
    return out


def collide(kind, f, rho, ux, uy, tau):
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.zeros(9)
    getattr(lib(), "or_collide_" + kind)(_p(f), rho, ux, uy, tau, _p(out))
    return out


def guo(ux, uy, fx, fy, tau_begin):
    out = np.zeros(9)
    lib().or_guo(ux, uy, fx, fy, tau_begin, _p(out))
    return out


def les_tau(f, rho, ux, uy, tau0, cs):
    f = np.ascontiguousarray(f, dtype=np.float64)
    return lib().or_les_tau(_p(f), rho, ux, uy, tau0, cs)


def delta(r):
    return lib().or_delta(r)


def noise(seed, node, comp):
    return lib().or_noise(seed, node, comp)


def mlp_features(f, rho, ux, uy):
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.zeros(6)
    lib().or_mlp_features(_p(f), rho, ux, uy, _p(out))
    return out


def mlp_forward(params, x):
    params = np.ascontiguousarray(params, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    return lib().or_mlp_forward(_p(params), _p(x))


def mlp_loss_grad(params, xs, ys):
    params = np.ascontiguousarray(params, dtype=np.float64)
    xs = np.ascontiguousarray(xs, dtype=np.float64)
    ys = np.ascontiguousarray(ys, dtype=np.float64)
    g = np.zeros(MLP_NPARAM)
    loss = lib().or_mlp_loss_grad(_p(params), _p(xs), _p(ys), len(ys), _p(g))
    return loss, g


BNN_NPARAM = 15 * 128 + 128 + 6 * (128 * 128 + 128) + 128 + 1


def bnn_forward(params, x):
    """The boundary network of Mesh::NNmodel_Type0x in FP32: params
    W0[128][15] b0[128], 6 x (W[128][128] b[128]), W7[128] b7; x [15]."""
    p = np.ascontiguousarray(params, dtype=np.float32)
    v = np.ascontiguousarray(x, dtype=np.float32)
    fp = ctypes.POINTER(ctypes.c_float)
    return float(lib().or_bnn_forward(p.ctypes.data_as(fp), v.ctypes.data_as(fp)))


# ---- whole-domain solver ------------------------------------------------
class Sim:
    """FP64 CPU solver; same configuration vocabulary as the HIP library."""

    def __init__(self, nx, ny, collision="bgk", bc="periodic_bb", les=False, tau_mode="const",
                 tau0=0.6, cs=0.0, u_in=(0.0, 0.0), threads=1, tau_nn_max=0.0, plate_x=0):
        self.nx, self.ny = nx, ny
        self.N = nx * ny
        cfg = OrConfig(nx, ny, COLL[collision], BC[bc], int(bool(les)), TAU[tau_mode],
                       tau0, cs, u_in[0], u_in[1], tau_nn_max, int(plate_x))
        lib().or_set_threads(int(threads))
        self.h = lib().or_create(ctypes.byref(cfg))

    def close(self):
        if self.h:
            lib().or_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init_fields(self, seed, rho0=1.0, u0=(0.0, 0.0), noise_amp=0.0):
        lib().or_init_fields(self.h, seed, rho0, u0[0], u0[1], noise_amp)

    def set_state(self, fnew, rho, ux, uy):
        a = [np.ascontiguousarray(x, dtype=np.float64).ravel() for x in (fnew, rho, ux, uy)]
        lib().or_set_state(self.h, *[_p(x) for x in a])

    def add_body(self, cx, cy, diameter, nmarkers, ud=(0.0, 0.0)):
        b = lib().or_add_body(self.h, cx, cy, diameter, nmarkers, ud[0], ud[1])
        if b < 0:
            raise ValueError("oracle rejected body")
        self._nbodies = b + 1
        return b

    def set_tau_field(self, tau):
        t = np.ascontiguousarray(tau, dtype=np.float64).ravel()
        lib().or_set_tau_field(self.h, _p(t))

    def set_mlp(self, blob):
        b = np.ascontiguousarray(blob, dtype=np.float64)
        assert b.size == MLP_BLOB
        lib().or_set_mlp(self.h, _p(b))

    def set_bnn(self, net_type, params):
        """Boundary-row / column network of type 1, 3, 5 or 7 (FP64 values,
        evaluated in FP32), layout of BNN_NPARAM parameters as bnn_forward."""
        p = np.ascontiguousarray(params, dtype=np.float64)
        assert p.size == BNN_NPARAM
        if lib().or_set_bnn(self.h, int(net_type), _p(p)) != 0:
            raise ValueError("boundary network type must be 1, 3, 5 or 7")

    def step(self, n):
        return lib().or_step(self.h, n)

    def fields(self):
        r, ux, uy = (np.zeros(self.N) for _ in range(3))
        lib().or_get_fields(self.h, _p(r), _p(ux), _p(uy))
        s = (self.ny, self.nx)
        return r.reshape(s), ux.reshape(s), uy.reshape(s)

    def fnew(self):
        f = np.zeros(9 * self.N)
        lib().or_get_fnew(self.h, _p(f))
        return f.reshape(9, self.ny, self.nx)

    def fcol(self):
        f = np.zeros(9 * self.N)
        lib().or_get_fcol(self.h, _p(f))
        return f.reshape(9, self.ny, self.nx)

    def tau(self):
        t = np.zeros(self.N)
        lib().or_get_tau(self.h, _p(t))
        return t.reshape(self.ny, self.nx)

    def forces(self):
        n = lib().or_num_force_records(self.h)
        nm = self.num_bodies()
        out = np.zeros(max(1, 2 * n * nm))
        if n and nm:
            lib().or_get_forces(self.h, _p(out))
        return out[: 2 * n * nm].reshape(n, nm, 2)

    def num_bodies(self):
        return getattr(self, "_nbodies", 0)

    def markers(self):
        n = lib().or_num_markers(self.h)
        out = np.zeros(max(1, 6 * n))
        if n:
            lib().or_get_markers(self.h, _p(out))
        return out[: 6 * n].reshape(n, 6)

    def force_field(self):
        fx, fy = np.zeros(self.N), np.zeros(self.N)
        lib().or_get_force_field(self.h, _p(fx), _p(fy))
        s = (self.ny, self.nx)
        return fx.reshape(s), fy.reshape(s)

    def mass(self):
        return lib().or_mass(self.h)


def drag_objective(nx, ny, seed, u0, noise, tau, center, diam, M, nsteps, avg_from):
    t = np.ascontiguousarray(tau, dtype=np.float64).ravel()
    return lib().or_drag_objective(nx, ny, seed, u0[0], u0[1], noise, _p(t), center[0], center[1], diam, M,
                                   nsteps, avg_from)


def adjoint_drag(nx, ny, seed, u0, noise, tau, center, diam, M, nsteps, avg_from):
    """(J, dJ/dtau [ny, nx], dJ/d(cx, cy)) of the FP64 BGK + IB step."""
    t = np.ascontiguousarray(tau, dtype=np.float64).ravel()
    J = np.zeros(1)
    gt = np.zeros(nx * ny)
    gc = np.zeros(2)
    lib().or_adjoint_drag(nx, ny, seed, u0[0], u0[1], noise, _p(t), center[0], center[1], diam, M, nsteps,
                          avg_from, _p(J), _p(gt), _p(gc))
    return J[0], gt.reshape(ny, nx), gc


def tau_search(fnew, feq, tau, target, m_lo=-10000, m_hi=10000, step=1e-5):
    """FindAndWriteTau's candidate search (mesh.cpp:2667-2704) for n nodes:
    fnew, feq, tau [n, 9] (of the node each population streams from),
    target [n, 3] = (rho, u, v). Returns (best dtau [n], best score [n])."""
    fnew, feq, tau, target = (np.ascontiguousarray(a, dtype=np.float64) for a in (fnew, feq, tau, target))
    n = fnew.shape[0]
    bd, be = np.zeros(n), np.zeros(n)
    lib().or_tau_search(n, _p(fnew), _p(feq), _p(tau), _p(target), int(m_lo), int(m_hi), float(step), _p(bd), _p(be))
    return bd, be


# ---- effective-relaxation-time pairs (stress matching) -------------------
_EX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1], dtype=np.float64)
_EY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1], dtype=np.float64)
_W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)


def _feq_grid(rho, ux, uy):
    """feq of NodeAndLattice.cpp:124-149 on whole fields, [9, ...]."""
    eu = _EX.reshape((9,) + (1,) * rho.ndim) * ux + _EY.reshape((9,) + (1,) * rho.ndim) * uy
    uu = ux * ux + uy * uy
    return _W.reshape((9,) + (1,) * rho.ndim) * rho * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * uu)


def _stress(d):
    """Pi_xx, Pi_xy, Pi_yy of populations d [9, ...]."""
    sh = (9,) + (1,) * (d.ndim - 1)
    ex, ey = _EX.reshape(sh), _EY.reshape(sh)
    return (d * ex * ex).sum(0), (d * ex * ey).sum(0), (d * ey * ey).sum(0)


def _blk(a):
    return 0.25 * (a[..., 0::2, 0::2] + a[..., 0::2, 1::2] + a[..., 1::2, 0::2] + a[..., 1::2, 1::2])


def stress_pairs(fp, tau_l, tau_coarse, tau0, excl=()):
    """Effective-relaxation-time training pairs at half resolution (the
    restatement ippi_coarse_stress_pairs is checked against).

    fp [9, ny, nx]: post-streaming fine populations; tau_l [ny, nx]: the
    relaxation time of each fine node's next collision (hybrid Smagorinsky
    included, mesh.cpp:877-910); tau0: the fine base relaxation time;
    excl: coarse boxes (x0, y0, x1, y1) inclusive that get weight 0.

    Per fine node: sigma = -(1 - 1/(2 tau_l)) Pi^neq (the viscous stress of
    its collision) and S = -Pi^neq / (2 rho c_s^2 tau_l) (its strain). Per
    2x2 block: T = <sigma> - (<rho u u> - rho_c u_c u_c) (resolved minus
    subgrid stress), C = 2 <S> (strain in coarse time units), label tau =
    1/2 + 3 (dev T : dev C) / (2 rho_c dev C : dev C), returned as omega =
    1/clamp(tau, 0.5005, 2), weight dev C : dev C. Features: rho, ux, uy,
    |S|, Pi_xx - Pi_yy, Pi_xy of the restricted state feq(rho_c, u_c) +
    s <f - feq> with s = 2 tau_coarse / tau0 (the network's inputs, as the
    oracle's or_mlp_features computes them from a node's populations).
    No counterpart in the reference (its labels come from FindAndWriteTau's
    one-step search, mesh.cpp:2625-2715)."""
    fp = np.asarray(fp, dtype=np.float64)
    rho = fp.sum(0)
    ux = (fp * _EX[:, None, None]).sum(0) / rho
    uy = (fp * _EY[:, None, None]).sum(0) / rho
    neq = fp - _feq_grid(rho, ux, uy)
    pxx, pxy, pyy = _stress(neq)
    g = -(1.0 - 0.5 / tau_l)
    c = -3.0 / (2.0 * rho * tau_l)
    rc = _blk(rho)
    ucx, ucy = _blk(rho * ux) / rc, _blk(rho * uy) / rc
    txx = _blk(g * pxx) - (_blk(rho * ux * ux) - rc * ucx * ucx)
    txy = _blk(g * pxy) - (_blk(rho * ux * uy) - rc * ucx * ucy)
    tyy = _blk(g * pyy) - (_blk(rho * uy * uy) - rc * ucy * ucy)
    cxx, cxy, cyy = 2 * _blk(c * pxx), 2 * _blk(c * pxy), 2 * _blk(c * pyy)
    a_t, a_c = 0.5 * (txx - tyy), 0.5 * (cxx - cyy)
    num = 2 * a_t * a_c + 2 * txy * cxy
    den = 2 * a_c * a_c + 2 * cxy * cxy
    ok = den > 0
    for (x0, y0, x1, y1) in excl:
        ok[max(y0, 0):y1 + 1, max(x0, 0):x1 + 1] = False
    tau = np.where(ok, 0.5 + 3 * num / (2 * rc * np.where(ok, den, 1.0)), 1.0)
    label = np.where(ok, 1.0 / np.clip(tau, 0.5005, 2.0), 0.0)
    weight = np.where(ok, den, 0.0)
    # features of the restricted state
    fa, na = _blk(fp), _blk(neq)
    r2 = fa.sum(0)
    vx, vy = (fa * _EX[:, None, None]).sum(0) / r2, (fa * _EY[:, None, None]).sum(0) / r2
    fc = _feq_grid(r2, vx, vy) + (2.0 * tau_coarse / tau0) * na
    qx, qxy, qy = _stress(fc - _feq_grid(r2, vx, vy))
    k = -3.0 / (2.0 * r2)
    sxx, sxy, syy = qx * k, qxy * k, qy * k
    feat = np.stack([r2, vx, vy, np.sqrt(2.0 * (sxx * sxx + 2.0 * sxy * sxy + syy * syy)), qx - qy, qxy], axis=-1)
    return feat.reshape(-1, 6), label.ravel(), weight.ravel(), tau.ravel()
