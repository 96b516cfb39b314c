"""MI355X-native D2Q9 time step of IPPI-LBM -- Python binding of libippi_lbm.so.

The compute path is the HIP library built from csrc/ (gfx950); this module
only marshals host arrays across the C ABI declared in include/ippi_lbm.h.
There is no CPU fallback: if the library is missing, or no GPU is visible,
constructing a Solver raises.

Reference counterpart: the `Mesh` object of IPPI-LBM (mesh.h:33-115), driven
by main.cpp:52-120; see facade/ for the C++ `Mesh` mirror over the same ABI.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libippi_lbm.so")

IPPI_OK, IPPI_ERR_ARG, IPPI_ERR_HIP, IPPI_ERR_DIVERGED, IPPI_ERR_STATE, IPPI_ERR_NOMEM, IPPI_ERR_CAPACITY = range(7)
COLLISION = {"bgk": 0, "mrt": 1, "rr": 2}
PRECISION = {"fp64": 0, "fp32": 1}
BC = {"periodic_bb": 0, "inflow": 1, "flatplate": 2}
TAU_MODE = {"const": 0, "field": 1, "nn": 2, "nn_bf16": 3}
MLP_BLOB = 6 * 32 + 32 + 32 * 32 + 32 + 32 + 1 + 12
IPPI_ADJ_FP64 = 1


class IppiError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ippi error {code}: {msg}")
        self.code = code


class Config(ctypes.Structure):
    """Mirror of struct ippi_config (include/ippi_lbm.h)."""
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int),
        ("collision", ctypes.c_int), ("precision", ctypes.c_int), ("bc", ctypes.c_int),
        ("les", ctypes.c_int), ("tau_mode", ctypes.c_int),
        ("tau0", ctypes.c_double), ("cs", ctypes.c_double),
        ("u_in_x", ctypes.c_double), ("u_in_y", ctypes.c_double),
        ("device", ctypes.c_int), ("force_history", ctypes.c_int),
        ("tau_nn_max", ctypes.c_double), ("plate_x", ctypes.c_int),
    ]


_D = ctypes.POINTER(ctypes.c_double)
_F = ctypes.POINTER(ctypes.c_float)
_lib = None

# every entry point of include/ippi_lbm.h with its ctypes signature
_SIGS = {
    "ippi_version": ([], ctypes.c_int),
    "ippi_create": ([ctypes.POINTER(Config), ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
    "ippi_destroy": ([ctypes.c_void_p], ctypes.c_int),
    "ippi_last_error": ([ctypes.c_void_p], ctypes.c_char_p),
    "ippi_init_fields": ([ctypes.c_void_p, ctypes.c_uint64, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                          ctypes.c_double], ctypes.c_int),
    "ippi_set_fields": ([ctypes.c_void_p, _D, _D, _D], ctypes.c_int),
    "ippi_add_body": ([ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                       ctypes.c_double, ctypes.c_double, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "ippi_num_bodies": ([ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "ippi_set_tau_field": ([ctypes.c_void_p, _D], ctypes.c_int),
    "ippi_set_mlp": ([ctypes.c_void_p, _D, ctypes.c_int], ctypes.c_int),
    "ippi_step": ([ctypes.c_void_p, ctypes.c_int], ctypes.c_int),
    "ippi_synchronize": ([ctypes.c_void_p], ctypes.c_int),
    "ippi_steps_done": ([ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
    "ippi_get_fields": ([ctypes.c_void_p, _D, _D, _D], ctypes.c_int),
    "ippi_get_populations": ([ctypes.c_void_p, _D], ctypes.c_int),
    "ippi_get_tau": ([ctypes.c_void_p, _D], ctypes.c_int),
    "ippi_get_body_forces": ([ctypes.c_void_p, _D, ctypes.c_int, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "ippi_get_markers": ([ctypes.c_void_p, _D, ctypes.c_int, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "ippi_mass": ([ctypes.c_void_p, _D], ctypes.c_int),
    "ippi_set_timing": ([ctypes.c_void_p, ctypes.c_int], ctypes.c_int),
    "ippi_get_timing": ([ctypes.c_void_p, _D, ctypes.POINTER(ctypes.c_int64), _D, ctypes.POINTER(ctypes.c_int64)],
                        ctypes.c_int),
    "ippi_set_populations": ([ctypes.c_void_p, _D], ctypes.c_int),
    "ippi_coarse_pairs": ([ctypes.c_void_p, ctypes.c_double, ctypes.c_int, _F, _F, _F], ctypes.c_int),
    "ippi_coarse_stress_pairs": ([ctypes.c_void_p, ctypes.c_double, _F, _F, _F], ctypes.c_int),
    "ippi_restrict": ([ctypes.c_void_p, ctypes.c_double, _D, _D], ctypes.c_int),
    "ippi_trainer_create": ([ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
    "ippi_trainer_destroy": ([ctypes.c_void_p], ctypes.c_int),
    "ippi_trainer_last_error": ([ctypes.c_void_p], ctypes.c_char_p),
    "ippi_trainer_set_data": ([ctypes.c_void_p, _F, _F, _F, ctypes.c_int], ctypes.c_int),
    "ippi_trainer_set_params": ([ctypes.c_void_p, _D, ctypes.c_int], ctypes.c_int),
    "ippi_trainer_get_params": ([ctypes.c_void_p, _D, ctypes.c_int], ctypes.c_int),
    "ippi_trainer_loss_grad": ([ctypes.c_void_p, _D, _D], ctypes.c_int),
    "ippi_trainer_adam": ([ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_uint64, _D],
                          ctypes.c_int),
    "ippi_adjoint_drag": ([ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _D, _D, _D, _D],
                          ctypes.c_int),
    "ippi_get_adjoint_timing": ([ctypes.c_void_p, _D, ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
    "ippi_pattern_probe": ([ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _D],
                           ctypes.c_int),
    "ippi_get_stencil": ([ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int), _D, _D, _D], ctypes.c_int),
    "ippi_set_boundary_mlp": ([ctypes.c_void_p, ctypes.c_int, _D, ctypes.c_int], ctypes.c_int),
    "ippi_tau_search": ([ctypes.c_int, ctypes.c_int, _D, _D, _D, _D, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                         _D, _D], ctypes.c_int),
    "ippi_tau_search_noalloc": ([ctypes.c_int, ctypes.c_int, _D, _D, _D, _D, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_double, _D, _D], ctypes.c_int),
    "ippi_accumulate_mean": ([ctypes.c_void_p, ctypes.c_int64], ctypes.c_int),
    "ippi_get_mean": ([ctypes.c_void_p, _D, _D, _D], ctypes.c_int),
    "ippi_save": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t)], ctypes.c_int),
    "ippi_load": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t], ctypes.c_int),
}


def library():
    """Load libippi_lbm.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library not built: {LIB_PATH} (run build_lib.py)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_D)


def _pf(a):
    return a.ctypes.data_as(_F) if a is not None else None


def _f64(a, n):
    a = np.ascontiguousarray(a, dtype=np.float64).ravel()
    if a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


class Solver:
    """One whole-domain D2Q9 solver on one GPU (a handle of the C ABI)."""

    def __init__(self, nx, ny, collision="bgk", precision="fp64", bc="periodic_bb", les=False,
                 tau_mode="const", tau0=0.6, cs=0.0, u_in=(0.0, 0.0), device=0, force_history=0,
                 tau_nn_max=0.0, plate_x=0):
        L = library()
        self.nx, self.ny, self.N = int(nx), int(ny), int(nx) * int(ny)
        self.precision = precision
        cfg = Config(self.nx, self.ny, COLLISION[collision], PRECISION[precision], BC[bc], int(bool(les)),
                     TAU_MODE[tau_mode], float(tau0), float(cs), float(u_in[0]), float(u_in[1]), int(device),
                     int(force_history), float(tau_nn_max), int(plate_x))
        h = ctypes.c_void_p()
        rc = L.ippi_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc != IPPI_OK:
            raise IppiError(rc, "ippi_create failed (bad configuration or no usable GPU)")
        self.h = h
        self.nbodies = 0

    # -- helpers --------------------------------------------------------------
    def _check(self, rc):
        if rc != IPPI_OK:
            msg = library().ippi_last_error(self.h)
            raise IppiError(rc, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "h", None):
            library().ippi_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- state ---------------------------------------------------------------
    def init_fields(self, seed, rho0=1.0, u0=(0.0, 0.0), noise_amp=0.0):
        self._check(library().ippi_init_fields(self.h, int(seed), float(rho0), float(u0[0]), float(u0[1]),
                                               float(noise_amp)))

    def set_fields(self, rho, ux, uy):
        a = [_f64(v, self.N) for v in (rho, ux, uy)]
        self._check(library().ippi_set_fields(self.h, *[_p(v) for v in a]))

    def add_body(self, cx, cy, diameter, nmarkers, ud=(0.0, 0.0)):
        b = ctypes.c_int()
        self._check(library().ippi_add_body(self.h, float(cx), float(cy), float(diameter), int(nmarkers),
                                            float(ud[0]), float(ud[1]), ctypes.byref(b)))
        self.nbodies = b.value + 1
        return b.value

    def set_populations(self, f):
        a = _f64(f, 9 * self.N)
        self._check(library().ippi_set_populations(self.h, _p(a)))

    def coarse_pairs(self, tau_coarse, coarse_collision="rr", criterion="step"):
        """Training pairs at half resolution: (features [n,6], labels omega
        [n], weights [n]) over the coarse grid. criterion "step": the
        reference's one-step search (advances this solver two steps);
        "stress": effective relaxation times by stress matching (no advance,
        coarse_collision unused)."""
        nc = (self.nx // 2) * (self.ny // 2)
        feat = np.empty(6 * nc, dtype=np.float32)
        lab = np.empty(nc, dtype=np.float32)
        w = np.empty(nc, dtype=np.float32)
        if criterion == "step":
            self._check(library().ippi_coarse_pairs(self.h, float(tau_coarse), COLLISION[coarse_collision],
                                                    _pf(feat), _pf(lab), _pf(w)))
        elif criterion == "stress":
            self._check(library().ippi_coarse_stress_pairs(self.h, float(tau_coarse), _pf(feat), _pf(lab), _pf(w)))
        else:
            raise ValueError(f"unknown criterion {criterion!r}")
        return feat.reshape(nc, 6), lab, w

    def restrict(self, tau_coarse, populations=True):
        """This state at half resolution: (coarse populations [9, ny/2, nx/2]
        or None, (rho, ux, uy) each [ny/2, nx/2])."""
        nxc, nyc = self.nx // 2, self.ny // 2
        mc = np.empty(3 * nxc * nyc)
        fc = np.empty(9 * nxc * nyc) if populations else None
        self._check(library().ippi_restrict(self.h, float(tau_coarse), None if fc is None else _p(fc), _p(mc)))
        m = mc.reshape(3, nyc, nxc)
        return (None if fc is None else fc.reshape(9, nyc, nxc)), (m[0], m[1], m[2])

    def set_tau_field(self, tau):
        t = _f64(tau, self.N)
        self._check(library().ippi_set_tau_field(self.h, _p(t)))

    def set_mlp(self, blob):
        b = _f64(blob, MLP_BLOB)
        self._check(library().ippi_set_mlp(self.h, _p(b), b.size))

    def set_boundary_mlp(self, net_type, blob):
        """The reference's boundary-row / column network of type 1, 3, 5 or
        7 (NNmodel_Type01/03/05/07): 101,249 parameters, W0[128][15] b0[128],
        6 x (W[128][128] b[128]), W7[128] b7 (nn.Linear layout)."""
        b = _f64(blob, BNN_NPARAM)
        self._check(library().ippi_set_boundary_mlp(self.h, int(net_type), _p(b), b.size))

    def step(self, n=1):
        self._check(library().ippi_step(self.h, int(n)))

    def synchronize(self):
        self._check(library().ippi_synchronize(self.h))

    def steps_done(self):
        s = ctypes.c_int64()
        self._check(library().ippi_steps_done(self.h, ctypes.byref(s)))
        return s.value

    # -- readback -------------------------------------------------------------
    def fields(self):
        r, ux, uy = (np.empty(self.N) for _ in range(3))
        self._check(library().ippi_get_fields(self.h, _p(r), _p(ux), _p(uy)))
        s = (self.ny, self.nx)
        return r.reshape(s), ux.reshape(s), uy.reshape(s)

    def accumulate_mean(self, k):
        """mean = (mean * k + current) / (k + 1) on the device (ComputeAvgVelocity)."""
        self._check(library().ippi_accumulate_mean(self.h, int(k)))

    def mean(self):
        r, ux, uy = (np.empty(self.N) for _ in range(3))
        self._check(library().ippi_get_mean(self.h, _p(r), _p(ux), _p(uy)))
        s = (self.ny, self.nx)
        return r.reshape(s), ux.reshape(s), uy.reshape(s)

    def tau(self):
        t = np.empty(self.N)
        self._check(library().ippi_get_tau(self.h, _p(t)))
        return t.reshape(self.ny, self.nx)

    def populations(self):
        f = np.empty(9 * self.N)
        self._check(library().ippi_get_populations(self.h, _p(f)))
        return f.reshape(9, self.ny, self.nx)

    def body_forces(self, max_records=1 << 30):
        n = ctypes.c_int()
        self._check(library().ippi_get_body_forces(self.h, None, int(max_records), ctypes.byref(n)))
        out = np.empty(max(1, 2 * n.value * self.nbodies))
        self._check(library().ippi_get_body_forces(self.h, _p(out), n.value, ctypes.byref(n)))
        return out[: 2 * n.value * self.nbodies].reshape(n.value, self.nbodies, 2)

    def markers(self):
        n = ctypes.c_int()
        self._check(library().ippi_get_markers(self.h, None, 1 << 30, ctypes.byref(n)))
        out = np.empty(max(1, 6 * n.value))
        self._check(library().ippi_get_markers(self.h, _p(out), n.value, ctypes.byref(n)))
        return out[: 6 * n.value].reshape(n.value, 6)

    def mass(self):
        m = ctypes.c_double()
        self._check(library().ippi_mass(self.h, ctypes.byref(m)))
        return m.value

    def set_timing(self, enable):
        """False/0: off; True/1: HIP events around every kernel launch;
        "calls"/2: one event pair around the launches of each step() call."""
        mode = 2 if enable == "calls" else int(enable)
        self._check(library().ippi_set_timing(self.h, mode))

    def timing(self):
        """(fused-kernel ms, launches, IB-kernels ms, IB launches) since the last call."""
        sm, im = ctypes.c_double(), ctypes.c_double()
        sc, ic = ctypes.c_int64(), ctypes.c_int64()
        self._check(library().ippi_get_timing(self.h, ctypes.byref(sm), ctypes.byref(sc), ctypes.byref(im),
                                              ctypes.byref(ic)))
        return sm.value, sc.value, im.value, ic.value

    def stencil(self, nodes):
        """(fnew, feq, tau) [n, 9] of the nodes each population of the given
        nodes ([n, 2] of (x, y)) streams in from, as the last pull step's
        collision used them: the inputs of tau_search."""
        xy = np.ascontiguousarray(np.asarray(nodes, dtype=np.int32).reshape(-1, 2))
        n = xy.shape[0]
        fn, fq, tk = np.zeros((n, 9)), np.zeros((n, 9)), np.zeros((n, 9))
        self._check(library().ippi_get_stencil(self.h, n, xy.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _p(fn),
                                               _p(fq), _p(tk)))
        return fn, fq, tk

    def adjoint_drag(self, nsteps, avg_from, ckpt=100, body_w=None, fp64_adjoint=False):
        """Run nsteps steps and the discrete adjoint of the averaged body force.

        Returns (J, dJ/dtau [ny, nx], dJ/d(cx, cy) [nbody, 2]); J is the mean
        over steps avg_from..nsteps of sum_b w_b . record_b (drag of body 0
        when body_w is None). The solver ends advanced by nsteps.
        fp64_adjoint: FP64 adjoint over an FP32 forward state.
        """
        nb = ctypes.c_int()
        self._check(library().ippi_num_bodies(self.h, ctypes.byref(nb)))
        w = None if body_w is None else _f64(body_w, 2 * nb.value)
        J = ctypes.c_double()
        gt = np.zeros(self.N)
        gc = np.zeros(2 * max(nb.value, 1))
        self._check(library().ippi_adjoint_drag(self.h, int(nsteps), int(avg_from), int(ckpt),
                                                IPPI_ADJ_FP64 if fp64_adjoint else 0,
                                                None if w is None else _p(w), ctypes.byref(J), _p(gt), _p(gc)))
        return J.value, gt.reshape(self.ny, self.nx), gc[: 2 * nb.value].reshape(nb.value, 2)

    def adjoint_timing(self):
        """(ms, launches) of the adjoint's all-node kernel since the last call."""
        ms, c = ctypes.c_double(), ctypes.c_int64()
        self._check(library().ippi_get_adjoint_timing(self.h, ctypes.byref(ms), ctypes.byref(c)))
        return ms.value, c.value

    def save(self):
        n = ctypes.c_size_t()
        self._check(library().ippi_save(self.h, None, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        self._check(library().ippi_save(self.h, buf, ctypes.byref(n)))
        return buf.raw[: n.value]

    def load(self, blob):
        buf = ctypes.create_string_buffer(blob, len(blob))
        self._check(library().ippi_load(self.h, buf, len(blob)))


class Trainer:
    """GPU trainer of the relaxation-time network (include/ippi_lbm.h)."""

    def __init__(self, capacity):
        h = ctypes.c_void_p()
        rc = library().ippi_trainer_create(int(capacity), ctypes.byref(h))
        if rc != IPPI_OK:
            raise IppiError(rc, "ippi_trainer_create failed")
        self.h = h

    def _check(self, rc):
        if rc != IPPI_OK:
            msg = library().ippi_trainer_last_error(self.h)
            raise IppiError(rc, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "h", None):
            library().ippi_trainer_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_data(self, x, y, w=None):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 6)
        y = np.ascontiguousarray(y, dtype=np.float32).ravel()
        w = None if w is None else np.ascontiguousarray(w, dtype=np.float32).ravel()
        self._keep = (x, y, w)
        self._check(library().ippi_trainer_set_data(self.h, _pf(x), _pf(y), _pf(w), len(y)))

    def set_params(self, blob):
        b = _f64(blob, MLP_BLOB)
        self._check(library().ippi_trainer_set_params(self.h, _p(b), MLP_BLOB))

    def params(self):
        b = np.empty(MLP_BLOB)
        self._check(library().ippi_trainer_get_params(self.h, _p(b), MLP_BLOB))
        return b

    def loss_grad(self):
        loss = ctypes.c_double()
        g = np.empty(MLP_BLOB - 12)
        self._check(library().ippi_trainer_loss_grad(self.h, ctypes.byref(loss), _p(g)))
        return loss.value, g

    def adam(self, iters, batch, lr, seed=0):
        loss = ctypes.c_double()
        self._check(library().ippi_trainer_adam(self.h, int(iters), int(batch), float(lr), int(seed),
                                                ctypes.byref(loss)))
        return loss.value


def drag_coefficient(force_x, u_ref, radius):
    """C = F / (0.5 U^2 2R), the normalisation of OutputCd (mesh.cpp:1501-1502)."""
    return force_x / (0.5 * u_ref * u_ref * 2 * radius)


def pattern_probe(nx, ny, precision="fp32", mode="pull", reps=200, device=0):
    """Microseconds per launch of a kernel with the fused step's memory
    pattern and no arithmetic (mode "pull": the nine shifted loads of an
    inflow step; "pull_xcd": the same with the rows grouped by XCD;
    "copy": a nine-array copy; "copy16": a flat copy of the same elements
    with 16 bytes per lane) on an nx x ny lattice."""
    us = ctypes.c_double()
    rc = library().ippi_pattern_probe(int(device), int(nx), int(ny), PRECISION[precision],
                                      {"copy": 0, "pull": 1, "pull_xcd": 2, "copy16": 3}[mode], int(reps),
                                      ctypes.byref(us))
    if rc != 0:
        raise IppiError(rc, "ippi_pattern_probe failed")
    return us.value


TAU_SEARCH_NOALLOC_MAX = 32768
BNN_NPARAM = 15 * 128 + 128 + 6 * (128 * 128 + 128) + 128 + 1  # 101249


def tau_search(fnew, feq, tau, target, m_lo=-10000, m_hi=10000, step=1e-5, device=0, noalloc=False):
    """The candidate loop of FindAndWriteTau (mesh.cpp:2667-2704) for n nodes
    on the GPU: fnew, feq, tau [n, 9] of the node each population streams in
    from, target [n, 3] = (rho, u, v). Returns (best dtau [n], score [n]).
    noalloc: ippi_tau_search_noalloc (fixed device staging, n at most
    TAU_SEARCH_NOALLOC_MAX)."""
    fnew, feq, tau, target = (np.ascontiguousarray(a, dtype=np.float64) for a in (fnew, feq, tau, target))
    n = fnew.shape[0]
    if fnew.shape != (n, 9) or feq.shape != (n, 9) or tau.shape != (n, 9) or target.shape != (n, 3):
        raise ValueError("fnew, feq, tau must be [n, 9] and target [n, 3]")
    bd, be = np.zeros(n), np.zeros(n)
    fn = library().ippi_tau_search_noalloc if noalloc else library().ippi_tau_search
    rc = fn(int(device), n, _p(fnew), _p(feq), _p(tau), _p(target), int(m_lo), int(m_hi), float(step), _p(bd), _p(be))
    if rc != 0:
        raise IppiError(rc, "ippi_tau_search_noalloc failed" if noalloc else "ippi_tau_search failed")
    return bd, be
