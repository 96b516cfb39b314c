"""HIP discrete adjoint (ippi_adjoint_drag) against the FP64 CPU oracle's
adjoint, against central finite differences of the HIP forward run, and
FP32 against FP64.

Bars: FP64 adjoint vs the oracle adjoint 1e-10 relative (different
summation orders); FP64 adjoint vs central differences (h = 1e-4) 1e-4
relative; FP32 forward with FP64 adjoint vs the FP64 adjoint and the
differences 1e-4 (norm-wise for the tau field, per component for the
centre); FP32 adjoint storage (FP64 arithmetic) 1e-4 for both.
"""
import dataclasses

import numpy as np
import pytest
import torch

from tests import _pkg

pytestmark = pytest.mark.gpu

H = 1e-4


@pytest.fixture(scope="module")
def ippi():
    m = _pkg.load()
    m.library()
    return m


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


SMALL = dict(nx=40, ny=28, seed=5, u0=(0.06, 0.004), noise=1e-3, center=(13.3, 13.6), diam=9.0, M=29)


def small_tau():
    return 0.6 + np.random.default_rng(0).uniform(-0.02, 0.02, (SMALL["ny"], SMALL["nx"]))


def small_gpu(ippi, precision, tau, nsteps, avg_from, ckpt):
    c = SMALL
    with ippi.Solver(c["nx"], c["ny"], "bgk", precision, "periodic_bb", tau_mode="field", tau0=0.6,
                     force_history=nsteps) as s:
        s.set_tau_field(tau)
        s.init_fields(c["seed"], 1.0, c["u0"], c["noise"])
        s.add_body(c["center"][0], c["center"][1], c["diam"], c["M"])
        return s.adjoint_drag(nsteps, avg_from, ckpt)


@pytest.mark.parametrize("nsteps,avg_from,ckpt", [(2, 2, 1), (60, 31, 7), (60, 1, 60)])
def test_fp64_adjoint_matches_oracle(ippi, oracle, nsteps, avg_from, ckpt):
    print(f"[adjoint] fp64 vs oracle nsteps={nsteps} avg_from={avg_from} ckpt={ckpt}", flush=True)
    tau = small_tau()
    J, gt, gc = small_gpu(ippi, "fp64", tau, nsteps, avg_from, ckpt)
    J0, gt0, gc0 = oracle.adjoint_drag(tau=tau, nsteps=nsteps, avg_from=avg_from, **SMALL)
    assert J == pytest.approx(J0, rel=1e-13)
    assert rel_l2(gt, gt0) <= 1e-10
    np.testing.assert_allclose(gc[0], gc0, rtol=1e-10)


def test_checkpoint_interval_does_not_change_result(ippi):
    print("[adjoint] checkpoint invariance", flush=True)
    tau = small_tau()
    ref = small_gpu(ippi, "fp64", tau, 45, 20, 45)
    for ck in (1, 4, 16):
        J, gt, gc = small_gpu(ippi, "fp64", tau, 45, 20, ck)
        assert J == ref[0]
        assert rel_l2(gt, ref[1]) <= 1e-14
        np.testing.assert_allclose(gc, ref[2], rtol=1e-14)


@pytest.mark.parametrize("precision,ckpt", [("fp64", 7), ("fp32", 4), ("fp32", 1)])
def test_recompute_overlap_is_bitwise_neutral(ippi, monkeypatch, precision, ckpt):
    """The next segment's recompute on a second stream beside the reverse
    sweep (the default) and between the sweeps (IPPI_ADJ_OVERLAP=0, read
    when the handle is created) give bit-identical objective and gradients,
    down to one-step segments."""
    print(f"[adjoint] recompute overlap {precision} ckpt={ckpt}", flush=True)
    tau = small_tau()
    on = small_gpu(ippi, precision, tau, 60, 31, ckpt)
    monkeypatch.setenv("IPPI_ADJ_OVERLAP", "0")
    off = small_gpu(ippi, precision, tau, 60, 31, ckpt)
    monkeypatch.delenv("IPPI_ADJ_OVERLAP")
    assert on[0] == off[0]
    assert np.array_equal(on[1], off[1]) and np.array_equal(on[2], off[2])


@pytest.mark.parametrize("precision,ckpt", [("fp64", 7), ("fp32", 4), ("fp32", 1)])
def test_fused_reverse_step_is_bitwise_neutral(ippi, monkeypatch, precision, ckpt):
    """One launch per reverse step (k_adj_fused, the default: the body's
    block computes the force adjoint and the marker transpose and publishes
    the velocity adjoint to the lattice rows of the same launch) against the
    three kernels in sequence (IPPI_ADJ_FUSED=0): bit-identical objective
    and gradients."""
    print(f"[adjoint] fused reverse step {precision} ckpt={ckpt}", flush=True)
    tau = small_tau()
    fused = small_gpu(ippi, precision, tau, 60, 31, ckpt)
    monkeypatch.setenv("IPPI_ADJ_FUSED", "0")
    split = small_gpu(ippi, precision, tau, 60, 31, ckpt)
    monkeypatch.delenv("IPPI_ADJ_FUSED")
    assert fused[0] == split[0]
    assert np.array_equal(fused[1], split[1]) and np.array_equal(fused[2], split[2])


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_fused_reverse_step_two_bodies(ippi, monkeypatch, precision):
    """Two bodies (two body blocks, patch rows of each waiting on their own
    body's flag; the objective weights both records): fused and three-kernel
    sweeps bit-identical."""
    print(f"[adjoint] fused reverse step, two bodies, {precision}", flush=True)
    nx, ny, nsteps, avg_from = 96, 48, 50, 21
    tau = 0.6 + np.random.default_rng(3).uniform(-0.02, 0.02, (ny, nx))
    bw = np.array([1.0, 0.0, 0.5, 0.25])

    def run():
        with ippi.Solver(nx, ny, "bgk", precision, "periodic_bb", tau_mode="field", tau0=0.6,
                         force_history=nsteps) as s:
            s.set_tau_field(tau)
            s.init_fields(8, 1.0, (0.05, 0.003), 1e-3)
            s.add_body(24.3, 15.6, 10.0, 31)
            s.add_body(66.7, 31.2, 12.0, 38)
            return s.adjoint_drag(nsteps, avg_from, 10, body_w=bw)
    fused = run()
    monkeypatch.setenv("IPPI_ADJ_FUSED", "0")
    split = run()
    monkeypatch.delenv("IPPI_ADJ_FUSED")
    assert fused[0] == split[0]
    assert np.array_equal(fused[1], split[1]) and np.array_equal(fused[2], split[2])
    assert np.all(np.abs(fused[2]) > 0)  # both centres carry a gradient


def test_fp32_adjoint_small(ippi, oracle):
    print("[adjoint] fp32 vs oracle, small", flush=True)
    tau = small_tau()
    J, gt, gc = small_gpu(ippi, "fp32", tau, 60, 31, 10)
    J0, gt0, gc0 = oracle.adjoint_drag(tau=tau, nsteps=60, avg_from=31, **SMALL)
    assert J == pytest.approx(J0, rel=1e-5)
    assert rel_l2(gt, gt0) <= 1e-4
    np.testing.assert_allclose(gc[0], gc0, rtol=1e-4)


# The differentiation workload: BGK + IB, 1024 x 512, cylinder D = 40
# (126 markers) near (256, 256), tau = 0.6 +- U(-0.02, 0.02) per node,
# Re = 200 on the mean tau (U = Re nu / D), 4000 steps, checkpoints every
# 100, mean drag over steps 2001-4000. The centre is offset so that no
# marker coordinate lies within 0.028 of an integer: the 4-point delta's
# second derivative jumps at |r| = 1 and 2, and a central difference whose
# stencil straddles such a jump has an O(h) error instead of O(h^2).
C4_CENTER = (256.47, 256.27)
def c4_case(ippi, precision):
    from importlib import import_module
    adj = import_module("ippi_amd.adjoint")
    nu = (0.6 - 0.5) / 3.0
    return adj.DragCase(nx=1024, ny=512, diameter=40.0, markers=126, nsteps=4000, avg_from=2001, ckpt=100,
                        precision=precision, seed=20240607, u0=(200 * nu / 40.0, 0.0), noise=1e-3)


def c4_tau():
    return 0.6 + np.random.default_rng(4).uniform(-0.02, 0.02, (512, 1024))


@pytest.fixture(scope="module")
def c4(ippi):
    tau, center = c4_tau(), np.array(C4_CENTER)
    i = np.arange(126)
    mk = np.concatenate([center[0] + 20 * np.sin(2 * np.pi * i / 126), center[1] + 20 * np.cos(2 * np.pi * i / 126)])
    assert np.min(np.abs(mk - np.round(mk))) > 0.02
    case64, case32 = c4_case(ippi, "fp64"), c4_case(ippi, "fp32")
    out = dict(tau=tau, center=center, case64=case64)
    print("[adjoint] C4 fp64 adjoint", flush=True)
    out["r64"] = case64.gradient(tau, center)
    print("[adjoint] C4 fp32 adjoint", flush=True)
    out["r32"] = case32.gradient(tau, center)
    print("[adjoint] C4 fp32 forward + fp64 adjoint", flush=True)
    out["r3264"] = dataclasses.replace(case32, fp64_adjoint=True).gradient(tau, center)
    # central differences of the FP64 forward run
    print("[adjoint] C4 finite differences", flush=True)
    out["J"] = case64.objective(tau, center)
    d = np.random.default_rng(9).choice([-1.0, 1.0], tau.shape)
    out["dir"] = d
    out["fd_dir"] = (case64.objective(tau + H * d, center) - case64.objective(tau - H * d, center)) / (2 * H)
    fdc = []
    for comp in (0, 1):
        e = np.zeros(2)
        e[comp] = H
        fdc.append((case64.objective(tau, center + e) - case64.objective(tau, center - e)) / (2 * H))
    out["fd_c"] = np.array(fdc)
    return out


def test_c4_fp64_adjoint_vs_finite_differences(c4):
    case, tau, c = c4["case64"], c4["tau"], c4["center"]
    J, gt, gc = c4["r64"]
    print(f"[adjoint] C4 J = {J:.10e}", flush=True)
    assert J == pytest.approx(c4["J"], rel=1e-13)
    ad = float(np.sum(gt * c4["dir"]))
    print(f"[adjoint] C4 directional tau: adjoint {ad:.10e} fd {c4['fd_dir']:.10e}", flush=True)
    assert ad == pytest.approx(c4["fd_dir"], rel=1e-4)
    for comp in (0, 1):
        print(f"[adjoint] C4 centre[{comp}]: adjoint {gc[comp]:.10e} fd {c4['fd_c'][comp]:.10e}", flush=True)
        assert gc[comp] == pytest.approx(c4["fd_c"][comp], rel=1e-4)
    # the two nodes of largest sensitivity
    for flat in np.argsort(-np.abs(gt).ravel())[:2]:
        j, i = divmod(int(flat), tau.shape[1])
        e = np.zeros_like(tau)
        e[j, i] = H
        fdn = (case.objective(tau + e, c) - case.objective(tau - e, c)) / (2 * H)
        print(f"[adjoint] C4 tau[{j},{i}]: adjoint {gt[j, i]:.10e} fd {fdn:.10e}", flush=True)
        assert gt[j, i] == pytest.approx(fdn, rel=1e-4)


def errors(c4, key):
    J64, gt64, gc64 = c4["r64"]
    J, gt, gc = c4[key]
    return dict(J=abs(J - J64) / abs(J64), tau=rel_l2(gt, gt64),
                dir=abs(float(np.sum(gt * c4["dir"])) - c4["fd_dir"]) / abs(c4["fd_dir"]),
                centre=np.abs(gc - gc64) / np.abs(gc64), centre_fd=np.abs(gc - c4["fd_c"]) / np.abs(c4["fd_c"]))


def test_c4_fp32_forward_fp64_adjoint(c4):
    e = errors(c4, "r3264")
    print(f"[adjoint] C4 fp32 forward + fp64 adjoint: {e}", flush=True)
    assert e["J"] <= 1e-5
    assert e["tau"] <= 1e-4 and e["dir"] <= 1e-4
    assert np.all(e["centre"] <= 1e-4) and np.all(e["centre_fd"] <= 1e-4)


def test_c4_fp32_adjoint(c4):
    # FP32 forward and FP32 adjoint populations: the tau field and the
    # centre gradient both meet 1e-4 against the FP64 adjoint and central
    # differences. The centre gradient is a sum over markers and steps that
    # cancels by ~1e4 (front and back of the cylinder); the reverse step's
    # arithmetic is FP64 over the FP32 storage, without which it was 2e-4 /
    # 4e-4 off.
    e = errors(c4, "r32")
    print(f"[adjoint] C4 fp32 adjoint: {e}", flush=True)
    assert e["J"] <= 1e-5
    assert e["tau"] <= 1e-4 and e["dir"] <= 1e-4
    assert np.all(e["centre"] <= 1e-4) and np.all(e["centre_fd"] <= 1e-4)


def test_c4_recompute_overlap_is_bitwise_neutral(c4, monkeypatch):
    """At the full C4 size (1024 x 512, 40 segments of 100 steps), where the
    recompute of each next segment on the second stream and the current
    segment's reverse sweep really run at the same time: the one-stream
    order (IPPI_ADJ_OVERLAP=0) gives the same objective and gradients bit
    for bit, in FP32."""
    monkeypatch.setenv("IPPI_ADJ_OVERLAP", "0")
    print("[adjoint] C4 fp32 adjoint, recompute between the sweeps", flush=True)
    off = c4_case(None, "fp32").gradient(c4["tau"], c4["center"])
    monkeypatch.delenv("IPPI_ADJ_OVERLAP")
    on = c4["r32"]
    assert on[0] == off[0]
    assert np.array_equal(on[1], off[1]) and np.array_equal(on[2], off[2])


def test_c4_fused_reverse_step_is_bitwise_neutral(c4, monkeypatch):
    """At the full C4 size in FP32: the fused reverse step against the three
    kernels in sequence (IPPI_ADJ_FUSED=0), bit for bit."""
    monkeypatch.setenv("IPPI_ADJ_FUSED", "0")
    print("[adjoint] C4 fp32 adjoint, three kernels per reverse step", flush=True)
    split = c4_case(None, "fp32").gradient(c4["tau"], c4["center"])
    monkeypatch.delenv("IPPI_ADJ_FUSED")
    fused = c4["r32"]
    assert fused[0] == split[0]
    assert np.array_equal(fused[1], split[1]) and np.array_equal(fused[2], split[2])


def test_autograd_function(ippi):
    from importlib import import_module
    adj = import_module("ippi_amd.adjoint")
    case = adj.DragCase(nx=40, ny=28, diameter=9.0, markers=29, nsteps=60, avg_from=31, ckpt=10, seed=5,
                        u0=(0.06, 0.004), noise=1e-3)
    tau = torch.tensor(small_tau(), requires_grad=True)
    c = torch.tensor([13.3, 13.6], dtype=torch.float64, requires_grad=True)
    J = adj.drag(tau, c, case)
    (3.0 * J).backward()
    Jr, gt, gc = case.gradient(small_tau(), np.array([13.3, 13.6]))
    assert J.item() == Jr
    np.testing.assert_array_equal(tau.grad.numpy(), 3.0 * gt)
    np.testing.assert_array_equal(c.grad.numpy(), 3.0 * gc)


def test_adjoint_rejects_unsupported(ippi):
    with ippi.Solver(64, 32, "mrt", "fp64", "periodic_bb") as s:
        s.init_fields(1, 1.0, (0.05, 0.0), 0.0)
        s.add_body(20, 16, 8, 25)
        with pytest.raises(ippi.IppiError):
            s.adjoint_drag(10, 5)
    with ippi.Solver(64, 32, "bgk", "fp64", "periodic_bb") as s:
        s.init_fields(1, 1.0, (0.05, 0.0), 0.0)
        with pytest.raises(ippi.IppiError):  # no body
            s.adjoint_drag(10, 5)
