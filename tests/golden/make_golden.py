"""Write golden vectors for the oracle's per-node functions from the
reference's own Lattice methods (IPPI-LBM/NodeAndLattice.cpp), compiled in
this container by `make -C oracle ref` into oracle/_ref/libippi_refnode.so.

Run:  python tests/golden/make_golden.py
Output: tests/golden/node_kernels.npz (inputs and the reference's outputs).
The reference library is only needed to regenerate the file; the tests read
the committed .npz.
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "libippi_refnode.so")

W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)


def main():
    if not os.path.exists(REF):
        sys.exit(f"missing {REF}: run `make -C oracle ref` first")
    L = ctypes.CDLL(REF)
    D = ctypes.POINTER(ctypes.c_double)
    d = ctypes.c_double
    L.ref_feq.argtypes = [d, d, d, D]
    L.ref_feq_init.argtypes = [d, d, d, D]
    L.ref_macro.argtypes = [D, D]
    for n in ("ref_bgk", "ref_mrt", "ref_rr"):
        getattr(L, n).argtypes = [D, d, d, d, d, D]
    L.ref_guo.argtypes = [d, d, d, d, d, D]
    p = lambda a: a.ctypes.data_as(D)

    rng = np.random.default_rng(20240531)
    n = 512
    rho = rng.uniform(0.9, 1.1, n)
    ux = rng.uniform(-0.15, 0.15, n)
    uy = rng.uniform(-0.15, 0.15, n)
    tau = rng.uniform(0.5005, 1.6, n)
    fxy = rng.uniform(-2e-3, 2e-3, (n, 2))
    # populations: equilibrium of a nearby state plus a non-equilibrium part
    f = np.empty((n, 9))
    for i in range(n):
        base = np.zeros(9)
        L.ref_feq(rho[i], ux[i], uy[i], p(base))
        f[i] = base + W * rng.uniform(-2e-2, 2e-2, 9)
    # edge cases: zero velocity, tau = 1, rho below the CalMacro threshold
    ux[:4] = 0.0
    uy[:4] = 0.0
    tau[4:8] = 1.0
    f[8] = 1e-8 * W
    f[9] = np.array([0.0, 1e-7, 0, 0, 0, 0, 0, 0, 0])

    out = {k: np.empty((n, 9)) for k in ("feq", "feq_init", "bgk", "mrt", "rr", "guo")}
    out["macro"] = np.empty((n, 3))
    for i in range(n):
        fi = np.ascontiguousarray(f[i])
        o = np.zeros(9)
        L.ref_feq(rho[i], ux[i], uy[i], p(o)); out["feq"][i] = o
        o = np.zeros(9)
        L.ref_feq_init(rho[i], ux[i], uy[i], p(o)); out["feq_init"][i] = o
        m = np.zeros(3)
        L.ref_macro(p(fi), p(m)); out["macro"][i] = m
        for k in ("bgk", "mrt", "rr"):
            o = np.zeros(9)
            getattr(L, "ref_" + k)(p(fi), rho[i], ux[i], uy[i], tau[i], p(o))
            out[k][i] = o
        o = np.zeros(9)
        L.ref_guo(ux[i], uy[i], fxy[i, 0], fxy[i, 1], tau[i], p(o)); out["guo"][i] = o

    np.savez_compressed(os.path.join(HERE, "node_kernels.npz"), f=f, rho=rho, ux=ux, uy=uy,
                        tau=tau, fxy=fxy, **{"out_" + k: v for k, v in out.items()})
    print("wrote", os.path.join(HERE, "node_kernels.npz"))


if __name__ == "__main__":
    main()
