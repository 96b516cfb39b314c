"""Several immersed bodies: HIP library (C ABI) against the FP64 CPU oracle.

The reference's marker list holds any number of markers
(InitialIBParticle, IPPI-LBM/mesh.cpp:173-191) and its interpolation and
spread loop over all of them (mesh.cpp:514-741); bodies here are placed by
workloads.place_cylinders, whose supports do not overlap. Both ways the
library runs the immersed boundary are covered:

* in-step: one block per body inside the fused step kernel, the rows of
  the bodies' patches (at most half the lattice) waiting for its forces;
* separate kernel (k_ib): chosen when the patches span more than half the
  rows, and always for the FP64 network.

Which one ran is read from the library's timing counters (IB-kernel
launches). Bars: FP64 fields, per-body force records and marker states
within 1e-12 relative L2 of the oracle; FP32 velocity within 1e-5 and
per-body force records within 1e-3 of the largest record.

At the 16-body size (8192 x 4096, MRT + LES + network tau, FP32) the
oracle is out of reach, so size-independent properties are checked:
chunked calls equal one call bit for bit, and the FP32 per-body drag and
the velocity field agree with the FP64 path.
"""
import numpy as np
import pytest

from tests import _pkg

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_U_TOL = 1e-5


@pytest.fixture(scope="module")
def ippi():
    m = _pkg.load()
    m.library()
    return m


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def vel_err(a, b):
    return float(np.sqrt(np.sum((a[1] - b[1]) ** 2 + (a[2] - b[2]) ** 2) / np.sum(b[1] ** 2 + b[2] ** 2)))


def layouts(ippi):
    from ippi_amd import workloads as wl
    # four bodies in a band of 60 rows of 128: in-step (one block per body needs nx >= 4 * 256)
    band = [(cx, cy + 34.0, d, m) for (cx, cy, d, m) in wl.place_cylinders(1024, 60, 4, 12.0, 20.0, seed=3)]
    # four bodies whose patches span more than half of 128 rows: the separate IB kernel
    spread = wl.place_cylinders(512, 128, 4, 14.0, 24.0, seed=5)
    return {"band": (1024, 128, band), "spread": (512, 128, spread)}


CASES = {
    # name: (layout, collision, network, steps)
    "mrt_les_band": ("band", "mrt", False, 250),
    "mrt_les_nn_band": ("band", "mrt", True, 200),
    "mrt_les_spread": ("spread", "mrt", False, 250),
    "mrt_les_nn_spread": ("spread", "mrt", True, 200),
    "rr_les_nn_spread": ("spread", "rr", True, 200),
}
U0, TAU0, CS = 0.06, 0.505, 0.1


def run_pair(oracle, ippi, case, precision):
    lay, coll, nn, steps = CASES[case]
    nx, ny, bodies = layouts(ippi)[lay]
    tm = "nn" if nn else "const"
    from ippi_amd import workloads as wl
    blob = wl.xavier_blob(7)
    o = oracle.Sim(nx, ny, coll, "inflow", les=True, tau_mode=tm, tau0=TAU0, cs=CS, u_in=(U0, 0.0), threads=16)
    g = ippi.Solver(nx, ny, coll, precision, "inflow", les=True, tau_mode=tm, tau0=TAU0, cs=CS, u_in=(U0, 0.0))
    for s in (o, g):
        s.init_fields(9, 1.0, (U0, 0.0), 1e-3)
        for (cx, cy, d, m) in bodies:
            s.add_body(cx, cy, d, m)
        if nn:
            s.set_mlp(blob)
    print(f"case {case} precision {precision} grid {nx}x{ny} bodies {len(bodies)} steps {steps}", flush=True)
    assert o.step(steps) == 0
    g.step(1)  # first collision from the initial state (no pull, no IB in the step kernel)
    g.set_timing(True)
    g.step(steps - 1)
    _, nstep, _, nib = g.timing()
    g.set_timing(False)
    assert nstep == steps - 1
    return o, g, steps, len(bodies), nib


def expected_ib_kernels(case, precision, steps):
    lay, _, nn, _ = CASES[case]
    in_step = lay == "band" and not (precision == "fp64" and nn)
    return 0 if in_step else steps - 1


@pytest.mark.parametrize("case", sorted(CASES))
def test_multibody_fp64_matches_oracle(oracle, ippi, case):
    o, g, steps, nb, nib = run_pair(oracle, ippi, case, "fp64")
    assert nib == expected_ib_kernels(case, "fp64", steps)
    r0, r1 = o.fields(), g.fields()
    assert rel_l2(r1[0], r0[0]) <= FP64_TOL
    assert vel_err(r1, r0) <= FP64_TOL
    d0, d1 = o.forces(), g.body_forces()
    assert d1.shape == d0.shape == (steps, nb, 2)
    for b in range(nb):
        assert rel_l2(d1[:, b], d0[:, b]) <= FP64_TOL, b
    m0, m1 = o.markers(), g.markers()
    assert np.array_equal(m0[:, :2], m1[:, :2])
    assert rel_l2(m1[:, 2:], m0[:, 2:]) <= FP64_TOL
    print(f"{case}: FP64 bitwise-equal populations: {np.array_equal(o.fnew(), g.populations())}")
    g.close()
    o.close()


@pytest.mark.parametrize("case", sorted(CASES))
def test_multibody_fp32_within_budget(oracle, ippi, case):
    o, g, steps, nb, nib = run_pair(oracle, ippi, case, "fp32")
    assert nib == expected_ib_kernels(case, "fp32", steps)
    r0, r1 = o.fields(), g.fields()
    eu = vel_err(r1, r0)
    d0, d1 = o.forces(), g.body_forces()
    fd = [float(np.abs(d1[:, b] - d0[:, b]).max() / np.abs(d0[:, b]).max()) for b in range(nb)]
    print(f"{case}: FP32 vs oracle rel L2 u {eu:.3e}; per-body force max rel err {max(fd):.3e}")
    assert eu <= FP32_U_TOL
    assert max(fd) <= 1e-3
    g.close()
    o.close()


def test_multibody_16_bodies_full_size(ippi):
    """8192 x 4096, 16 cylinders D ~ U[40, 120] (the large multi-body
    workload), MRT + LES + network tau: FP32 chunked calls equal one call
    bit for bit; FP32 per-body mean drag within 1e-4 of the FP64 path's
    over 200 steps; every body's record is finite and shows a drag."""
    from ippi_amd import workloads as wl
    nx, ny, U = 8192, 4096, 0.06
    bodies = wl.place_cylinders(nx, ny, 16, 40.0, 120.0, seed=11)
    assert sum(b[3] for b in bodies) > 3500
    tau0 = 0.5 + 3 * U * float(np.mean([b[2] for b in bodies])) / 1e4
    blob = wl.xavier_blob(7)

    def make(prec):
        s = ippi.Solver(nx, ny, "mrt", prec, "inflow", les=True, tau_mode="nn", tau0=tau0, cs=CS, u_in=(U, 0.0))
        s.init_fields(20240601, 1.0, (U, 0.0), 1e-3)
        for (cx, cy, d, m) in bodies:
            s.add_body(cx, cy, d, m)
        s.set_mlp(blob)
        return s
    steps = 200
    print(f"16 bodies 8192x4096: fp32 one call of {steps} steps", flush=True)
    a = make("fp32")
    a.step(steps)
    fa, da = a.fields(), a.body_forces()
    a.close()
    print("16 bodies 8192x4096: fp32 in chunks", flush=True)
    b = make("fp32")
    for n in (1, 37, 62, 100):
        b.step(n)
    fb, db = b.fields(), b.body_forces()
    b.close()
    for x, y in zip(fa, fb):
        assert np.array_equal(x, y)
    assert np.array_equal(da, db)
    print("16 bodies 8192x4096: fp64", flush=True)
    c = make("fp64")
    c.step(steps)
    dc = c.body_forces()
    ev = vel_err(fa, c.fields())
    c.close()
    assert da.shape == dc.shape == (steps, 16, 2)
    assert np.all(np.isfinite(da))
    m32, m64 = da[steps // 2:, :, 0].mean(0), dc[steps // 2:, :, 0].mean(0)
    err = np.abs(m32 - m64) / np.abs(m64)
    print(f"16 bodies: velocity rel L2 FP32 vs FP64 {ev:.3e}; per-body mean drag rel err max {err.max():.3e}")
    assert err.max() <= 1e-4
    assert ev <= FP32_U_TOL
    assert np.all(m64 > 0)  # every cylinder in the inflow is dragged downstream


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_large_body_separate_kernel_matches_oracle(oracle, ippi, precision):
    """One cylinder of diameter 320 (1005 markers, about 4,100 support
    nodes: more than 3,072, within the separate IB kernel's 6,144) on
    768 x 448, RR + LES: the body of the 2x finer training run."""
    nx, ny, D = 768, 448, 320.0
    M = int(round(np.pi * D))
    kw = dict(les=True, tau0=0.5077, cs=0.1, u_in=(0.08, 0.0))
    o = oracle.Sim(nx, ny, "rr", "inflow", threads=16, **kw)
    g = ippi.Solver(nx, ny, "rr", precision, "inflow", **kw)
    for s in (o, g):
        s.init_fields(5, 1.0, (0.08, 0.0), 1e-3)
        s.add_body(384.0, 224.0, D, M)
    steps = 40
    print(f"large body D={D} markers {M} precision {precision} steps {steps}", flush=True)
    assert o.step(steps) == 0
    g.step(1)
    g.set_timing(True)
    g.step(steps - 1)
    _, _, _, nib = g.timing()
    g.set_timing(False)
    assert nib == steps - 1  # the separate kernel ran
    r0, r1 = o.fields(), g.fields()
    d0, d1 = o.forces(), g.body_forces()
    if precision == "fp64":
        assert vel_err(r1, r0) <= FP64_TOL
        assert rel_l2(d1[:, 0], d0[:, 0]) <= FP64_TOL
    else:
        assert vel_err(r1, r0) <= FP32_U_TOL
        assert float(np.abs(d1[:, 0] - d0[:, 0]).max() / np.abs(d0[:, 0]).max()) <= 1e-3
    print(f"FP{precision[2:]}: rel L2 u {vel_err(r1, r0):.3e}, drag record rel L2 {rel_l2(d1[:, 0], d0[:, 0]):.3e}")
    g.close()
    o.close()
