"""FP64 pull steps without the network take one row per thread below 2^20
nodes and four from there on. Both kernels must give the same bits: each
case runs with IPPI_F64_ROWS=1 and =4 (read at ippi_create) and compares
fields, τ, body records and marker state, on row counts that are and are
not multiples of four, with the IB inside the step, as a separate kernel
and without a body; and the FP64 adjoint, whose recomputed segments use the
same kernels. The one-row kernels are pinned to the oracle by the parity
tests at these sizes; the four-row ones by this equality.
"""
import os

import numpy as np
import pytest

from tests import _pkg

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ippi():
    m = _pkg.load()
    m.library()
    return m


def make(ippi, rows, nx, ny, coll, bc, les, body, ib_in_step="1"):
    os.environ["IPPI_F64_ROWS"] = rows
    os.environ["IPPI_IB_IN_STEP"] = ib_in_step
    try:
        s = ippi.Solver(nx, ny, coll, "fp64", bc, les=les, tau0=0.58, cs=0.12, u_in=(0.06, 0.0))
    finally:
        del os.environ["IPPI_F64_ROWS"]
        del os.environ["IPPI_IB_IN_STEP"]
    s.init_fields(17, 1.0, (0.06, 0.0), 1e-3)
    if body:
        s.add_body(*body)
    return s


CASES = {
    "bgk_periodic_body_256x128": (256, 128, "bgk", "periodic_bb", False, (64.3, 64.1, 20.0, 63), "1"),
    "mrt_les_inflow_body_300x91": (300, 91, "mrt", "inflow", True, (80.2, 45.6, 16.0, 50), "1"),
    "mrt_les_inflow_body_separate_ib_300x91": (300, 91, "mrt", "inflow", True, (80.2, 45.6, 16.0, 50), "0"),
    "rr_inflow_no_body_130x67": (130, 67, "rr", "inflow", False, None, "1"),
    "bgk_periodic_no_body_97x30": (97, 30, "bgk", "periodic_bb", False, None, "1"),
}


@pytest.mark.parametrize("name", list(CASES))
def test_f64_one_and_four_rows_bitwise(ippi, name):
    nx, ny, coll, bc, les, body, ink = CASES[name]
    print(f"[f64 rows] {name}: 1 vs 4 rows per thread, 150 steps", flush=True)
    out = []
    for rows in ("1", "4"):
        with make(ippi, rows, nx, ny, coll, bc, les, body, ink) as s:
            s.step(100)
            s.step(50)
            r = [*s.fields(), s.tau()]
            if body:
                r += [s.body_forces(), s.markers()]
            out.append(r)
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_f64_adjoint_one_and_four_rows_bitwise(ippi):
    from ippi_amd import adjoint
    print("[f64 rows] FP64 adjoint 160x96, 120 steps, checkpoints every 10: 1 vs 4 rows", flush=True)
    tau = 0.6 + np.random.default_rng(2).uniform(-0.02, 0.02, (96, 160))
    res = []
    for rows in ("1", "4"):
        case = adjoint.DragCase(nx=160, ny=96, diameter=16.0, markers=50, nsteps=120, avg_from=61, ckpt=10,
                                precision="fp64", seed=11, u0=(0.05, 0.0))
        os.environ["IPPI_F64_ROWS"] = rows
        try:
            s = case.solver(tau, (48.3, 47.8))
        finally:
            del os.environ["IPPI_F64_ROWS"]
        with s:
            res.append(s.adjoint_drag(case.nsteps, case.avg_from, case.ckpt))
    assert res[0][0] == res[1][0]
    assert np.array_equal(res[0][1], res[1][1])
    assert np.array_equal(res[0][2], res[1][2])
