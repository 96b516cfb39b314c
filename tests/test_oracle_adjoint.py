"""Discrete adjoint of the FP64 oracle (BGK + IB, periodic x, bounce-back
walls) against central finite differences of the oracle's own forward run.

The adjoint transposes every stage of the step by hand (streaming with
bounce-back, macros, IB interpolation / direct forcing / spreading,
velocity correction, BGK collision, Guo source); a dropped term, a wrong
sign or a transposed index in any stage shows up as a mismatch with the
finite differences below, which only run the forward step.
"""
import numpy as np
import pytest

H = 1e-4


def case(nsteps, avg_from, **kw):
    c = dict(nx=40, ny=28, seed=5, u0=(0.06, 0.004), noise=1e-3, center=(13.3, 13.6), diam=9.0, M=29,
             nsteps=nsteps, avg_from=avg_from)
    c.update(kw)
    return c


def tau_field(c, seed=0):
    rng = np.random.default_rng(seed)
    return 0.6 + rng.uniform(-0.02, 0.02, (c["ny"], c["nx"]))


def fd_tau(oracle, c, tau, d):
    return (oracle.drag_objective(tau=tau + H * d, **c) - oracle.drag_objective(tau=tau - H * d, **c)) / (2 * H)


def fd_center(oracle, c, tau, comp):
    lo, hi = list(c["center"]), list(c["center"])
    lo[comp] -= H
    hi[comp] += H
    return (oracle.drag_objective(tau=tau, **{**c, "center": tuple(hi)}) -
            oracle.drag_objective(tau=tau, **{**c, "center": tuple(lo)})) / (2 * H)


@pytest.mark.parametrize("nsteps,avg_from", [(1, 1), (2, 2), (3, 1), (50, 26)])
def test_adjoint_matches_fd(oracle, nsteps, avg_from):
    c = case(nsteps, avg_from)
    tau = tau_field(c)
    J, gt, gc = oracle.adjoint_drag(tau=tau, **c)
    assert J == pytest.approx(oracle.drag_objective(tau=tau, **c), rel=1e-15, abs=0)
    # single nodes: on the marker support, next to it, and in the open flow
    for (j, i) in [(13, 13), (10, 9), (17, 18), (13, 8), (5, 30)]:
        d = np.zeros_like(tau)
        d[j, i] = 1.0
        fd = fd_tau(oracle, c, tau, d)
        if abs(gt[j, i]) < 1e-12 and abs(fd) < 1e-12:
            continue
        assert gt[j, i] == pytest.approx(fd, rel=1e-5, abs=1e-12), (j, i)
    # a random direction over the whole field
    d = np.random.default_rng(1).standard_normal(tau.shape)
    assert float(np.sum(gt * d)) == pytest.approx(fd_tau(oracle, c, tau, d), rel=1e-6)
    for comp in (0, 1):
        assert gc[comp] == pytest.approx(fd_center(oracle, c, tau, comp), rel=1e-5, abs=1e-12)


def test_adjoint_zero_outside_domain_of_dependence(oracle):
    # the record of step 1 is taken before any collision, so it does not
    # depend on tau at all; the record of step 2 sees the collisions of step
    # 1 at the 4x4 supports and at their pull neighbours only
    c = case(1, 1)
    tau = tau_field(c)
    _, gt, _ = oracle.adjoint_drag(tau=tau, **c)
    assert np.all(gt == 0.0)
    c = case(2, 2)
    _, gt, _ = oracle.adjoint_drag(tau=tau, **c)
    # markers lie on the circle of radius 4.5, supports reach 2 further per axis
    y, x = np.mgrid[0:c["ny"], 0:c["nx"]]
    r = np.hypot(x - c["center"][0], y - c["center"][1])
    assert np.all(gt[r > 4.5 + 3 * np.sqrt(2) + 1e-9] == 0.0)
    assert np.all(gt[r < 4.5 - 3 * np.sqrt(2) - 1e-9] == 0.0)
    assert np.count_nonzero(gt) > 50


def test_adjoint_translation_of_body_and_field(oracle):
    # shifting the body and the tau field by whole lattice units along the
    # periodic axis shifts the gradient (no noise: the initial state is
    # uniform, so the whole problem is translation invariant in x)
    c = case(20, 11, noise=0.0)
    tau = tau_field(c)
    J0, g0, gc0 = oracle.adjoint_drag(tau=tau, **c)
    c2 = dict(c, center=(c["center"][0] + 7, c["center"][1]))
    J1, g1, gc1 = oracle.adjoint_drag(tau=np.roll(tau, 7, axis=1), **c2)
    assert J1 == pytest.approx(J0, rel=1e-12)
    np.testing.assert_allclose(g1, np.roll(g0, 7, axis=1), rtol=1e-9, atol=1e-15)
    np.testing.assert_allclose(gc1, gc0, rtol=1e-9)
