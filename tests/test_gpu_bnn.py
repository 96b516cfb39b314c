"""Boundary-row / column networks (the reference's Mesh::NNmodel_Type01 /
03 / 05 / 07) on the GPU (C ABI) against the CPU oracle.

The networks run in FP32 with fused multiply-adds in a fixed order on both
sides, so the FP64 lattice path matches the oracle bit for bit; the bar is
relative L2 <= 1e-12 on density, velocity and the relaxation time the next
collision uses. The FP32 lattice path keeps the FP32 path's budget
(velocity 1e-5). Parity with the reference's TorchScript files
(AutoT0x.pt) is unpinned: the build does not load them.
"""
import numpy as np
import pytest

from tests import _pkg
from tests.test_oracle_bnn import random_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ippi():
    m = _pkg.load()
    m.library()
    return m


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def vel_err(a, b):
    return float(np.sqrt(np.sum((a[1] - b[1]) ** 2 + (a[2] - b[2]) ** 2) / np.sum(b[1] ** 2 + b[2] ** 2)))


CASES = {
    # name: (nx, ny, collision, bc, les, tau0, cs, u_in, types, body, steps)
    "flatplate_bgk_les_all": (160, 48, "bgk", "flatplate", True, 0.52, 0.173, (0.1, 0.0), (1, 3, 5, 7), None, 200),
    "inflow_mrt_les_rows_ib": (128, 64, "mrt", "inflow", True, 0.53, 0.1, (0.08, 0.0), (3, 7), (40, 32, 16, 51), 150),
    "channel_bgk_columns": (96, 40, "bgk", "periodic_bb", False, 0.6, 0.0, (0.05, 0.0), (1, 5), None, 150),
    # every type's node list shorter than one 16-node tile (mostly padding)
    "tiny_flatplate_all": (9, 6, "bgk", "flatplate", True, 0.52, 0.173, (0.1, 0.0), (1, 3, 5, 7), None, 60),
}


def run(oracle, ippi, case, precision, extra_oracle_steps=0):
    nx, ny, coll, bc, les, tau0, cs, u0, types, body, steps = CASES[case]
    kw = dict(les=les, tau0=tau0, cs=cs, u_in=u0, plate_x=40)
    o = oracle.Sim(nx, ny, coll, bc, threads=4, **kw)
    g = ippi.Solver(nx, ny, coll, precision, bc, **kw)
    for s in (o, g):
        s.init_fields(3, 1.0, u0, 1e-3)
        if body:
            s.add_body(*body)
    for t in types:
        p = random_params(100 + t)
        o.set_bnn(t, p)
        g.set_boundary_mlp(t, p)
    print(f"case {case} precision {precision} types {types} steps {steps}", flush=True)
    assert o.step(steps + extra_oracle_steps) == 0
    g.step(steps)
    return o, g


@pytest.mark.parametrize("case", sorted(CASES))
def test_bnn_fp64_matches_oracle(oracle, ippi, case):
    o, g = run(oracle, ippi, case, "fp64")
    a, b = g.fields(), o.fields()
    assert rel_l2(a[0], b[0]) <= 1e-12
    assert vel_err(a, b) <= 1e-12
    print(f"{case}: fields bitwise equal: {all(np.array_equal(x, y) for x, y in zip(a, b))}")
    g.close()
    o.close()


@pytest.mark.parametrize("case", sorted(CASES))
def test_bnn_tau_readback(oracle, ippi, case):
    """The tau the next collision uses (ippi_get_tau) against the oracle's
    tau of that collision: networks accepted where >= 0.501, LES (or base)
    elsewhere (outside the body's marker supports, where the read-back's LES
    omits the IB velocity correction)."""
    o, g = run(oracle, ippi, case, "fp64", extra_oracle_steps=1)
    t_gpu, t_ref = g.tau(), o.tau()
    # the read-back evaluates LES without the IB velocity correction: compare away from the body's force
    fx, fy = o.force_field()
    free = (fx == 0) & (fy == 0)
    assert rel_l2(t_gpu[free], t_ref[free]) <= 1e-12
    nx, ny = CASES[case][0], CASES[case][1]
    edge = np.zeros((ny, nx), bool)
    edge[1, :] = edge[ny - 2, :] = True
    edge[1:ny - 1, 0] = edge[1:ny - 1, nx - 1] = True
    acc = np.mean(t_ref[edge] >= 0.501) if CASES[case][4] else None
    print(f"{case}: tau rel L2 {rel_l2(t_gpu[free], t_ref[free]):.2e}; edge nodes with tau >= 0.501: {acc}")
    g.close()
    o.close()


@pytest.mark.parametrize("case", sorted(CASES))
def test_bnn_fp32_within_budget(oracle, ippi, case):
    o, g = run(oracle, ippi, case, "fp32")
    e = vel_err(g.fields(), o.fields())
    print(f"{case}: FP32 vs oracle velocity rel L2 {e:.3e}")
    assert e <= 1e-5
    g.close()
    o.close()


def test_bnn_rejects_bad_input(ippi):
    with ippi.Solver(32, 16, "bgk", "fp64", "inflow", u_in=(0.05, 0.0)) as s:
        with pytest.raises(ippi.IppiError):
            s.set_boundary_mlp(2, np.zeros(ippi.BNN_NPARAM))
        p = np.zeros(ippi.BNN_NPARAM)
        p[5] = np.nan
        with pytest.raises(ippi.IppiError):
            s.set_boundary_mlp(3, p)
        with pytest.raises((ValueError, ippi.IppiError)):
            s.set_boundary_mlp(3, np.zeros(10))


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_bnn_mfma_matches_fma_chains(ippi, monkeypatch, case, precision):
    """The networks on FP32 MFMA tiles (k_bnn_mfma, the default) and as
    sequential FMA chains (IPPI_BNN_MFMA=0, bit for bit the oracle's): an
    MFMA adds its four products to the accumulator in order with a rounding
    each, so the two give the same relaxation times and fields bit for bit
    over the case's steps."""
    nx, ny, coll, bc, les, tau0, cs, u0, types, body, steps = CASES[case]
    kw = dict(les=les, tau0=tau0, cs=cs, u_in=u0, plate_x=40)

    def make():
        g = ippi.Solver(nx, ny, coll, precision, bc, **kw)
        g.init_fields(3, 1.0, u0, 1e-3)
        if body:
            g.add_body(*body)
        for t in types:
            g.set_boundary_mlp(t, random_params(100 + t))
        return g
    mm = make()
    monkeypatch.setenv("IPPI_BNN_MFMA", "0")
    ch = make()
    monkeypatch.delenv("IPPI_BNN_MFMA")
    for g in (mm, ch):
        g.step(steps)
    edge = np.zeros((ny, nx), bool)
    edge[1, :] = edge[ny - 2, :] = True
    edge[1:ny - 1, 0] = edge[1:ny - 1, nx - 1] = True
    t_m, t_c = mm.tau(), ch.tau()
    print(f"{case} {precision}: edge nodes accepting the network: {np.mean(t_c[edge] >= 0.501):.2f}")
    assert np.array_equal(t_m, t_c)
    assert all(np.array_equal(a, b) for a, b in zip(mm.fields(), ch.fields()))
    mm.close()
    ch.close()


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_bnn_mfma_full_size_flat_plate(ippi, monkeypatch, precision):
    """The reference's own run at full size (2401 x 501 flat plate, BGK +
    LES Cs = 0.173, Re = 5e5) with all four boundary networks: the MFMA
    networks and the FMA chains give the same relaxation times and fields
    bit for bit after 50 steps (366 tiles of 16 nodes, each type's list
    padded to whole tiles)."""
    from ippi_amd import workloads
    nx, ny, u = 2401, 501, 0.1
    tau = 0.5 + 6.0 * (ny * u) / 5e5

    def make():
        g = ippi.Solver(nx, ny, "bgk", precision, "flatplate", les=True, tau0=tau, cs=0.173, u_in=(u, 0.0),
                        plate_x=400)
        g.init_fields(1, 1.0, (u, 0.0), 1e-3)
        for t in (1, 3, 5, 7):
            g.set_boundary_mlp(t, workloads.boundary_blob(t))
        return g
    mm = make()
    monkeypatch.setenv("IPPI_BNN_MFMA", "0")
    ch = make()
    monkeypatch.delenv("IPPI_BNN_MFMA")
    print(f"flat plate 2401x501 {precision}: 50 steps, MFMA networks vs FMA chains", flush=True)
    for g in (mm, ch):
        g.step(50)
    t_m, t_c = mm.tau(), ch.tau()
    edge = np.zeros((ny, nx), bool)
    edge[1, :] = edge[ny - 2, :] = True
    edge[1:ny - 1, 0] = edge[1:ny - 1, nx - 1] = True
    print(f"  edge nodes accepting the network: {np.mean(t_c[edge] >= 0.501):.3f}")
    assert np.array_equal(t_m, t_c)
    assert all(np.array_equal(a, b) for a, b in zip(mm.fields(), ch.fields()))
    mm.close()
    ch.close()
