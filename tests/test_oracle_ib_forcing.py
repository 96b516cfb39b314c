"""Pins of the oracle's immersed-boundary velocity correction
(UpdateVelocity, IPPI-LBM/mesh.cpp:1244-1281: u += F / (2 rho)) and of the
Guo source added to the collision (Lattice::Compute_Disforce,
IPPI-LBM/NodeAndLattice.cpp:162-195, with TauAtBeginning), against moment
identities of the published scheme rather than the oracle's own formulas:

* Guo et al. (2002) forcing: sum_k S_k = 0, sum_k S_k e = (1 - 1/(2 tau0)) F,
  sum_k S_k e e = (1 - 1/(2 tau0)) (u* F + F u*), with u* the corrected
  velocity and tau0 the base relaxation time the reference passes.
* BGK relaxes the first moment to rho u* = j + F / 2 and the second to
  rho u* u* + rho / 3 I (the second-order D2Q9 equilibrium's moments).
* MRT leaves the momentum moments unrelaxed (rates 0, NodeAndLattice.cpp:306).

Together they fix each factor: a dropped or doubled 1/2 in the correction,
a missing or doubled Guo term, the wrong tau in it, or the uncorrected u in
it changes the first or second moment of the post-collision populations
at every forced node.
"""
import numpy as np
import pytest

EX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1], dtype=float)
EY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1], dtype=float)


def _moments(f):
    """mass, momentum and second moments of [9, ny, nx] populations"""
    m = f.sum(0)
    jx = np.tensordot(EX, f, 1)
    jy = np.tensordot(EY, f, 1)
    pxx = np.tensordot(EX * EX, f, 1)
    pxy = np.tensordot(EX * EY, f, 1)
    pyy = np.tensordot(EY * EY, f, 1)
    return m, jx, jy, pxx, pxy, pyy


def _one_forced_step(oracle, coll, tau0, les=False, cs=0.0, ud=(0.04, -0.03), u0=(0.05, 0.01), noise=2e-3):
    nx, ny = 48, 40
    s = oracle.Sim(nx, ny, coll, "periodic_bb", les=les, tau0=tau0, cs=cs)
    s.init_fields(17, 1.0, u0, noise)
    s.add_body(23.3, 19.6, 12.0, 38, ud=ud)
    pre = s.fnew()
    assert s.step(1) == 0
    post = s.fcol()
    fx, fy = s.force_field()
    tau = s.tau()
    mk = s.markers()
    s.close()
    return pre, post, fx, fy, tau, mk


@pytest.mark.parametrize("tau0", [0.55, 0.8, 1.3])
def test_bgk_forced_moments(oracle, tau0):
    """BGK with constant tau: per node, momentum gain = F exactly (F/(2 tau)
    from relaxing towards j + F/2, plus (1 - 1/(2 tau)) F from Guo), mass
    unchanged, and the second moment of the closed form."""
    pre, post, fx, fy, tau, mk = _one_forced_step(oracle, "bgk", tau0)
    assert np.all(tau == tau0)
    r, jx, jy, pxx, pxy, pyy = _moments(pre)
    r1, jx1, jy1, pxx1, pxy1, pyy1 = _moments(post)
    forced = (fx != 0) | (fy != 0)
    assert forced.sum() > 100  # the body's 4x4 supports
    sc = np.abs(fx).max()
    assert sc > 1e-3
    np.testing.assert_allclose(r1, r, rtol=0, atol=1e-15)
    np.testing.assert_allclose(jx1 - jx, fx, rtol=0, atol=1e-14 * max(1.0, sc))
    np.testing.assert_allclose(jy1 - jy, fy, rtol=0, atol=1e-14 * max(1.0, sc))
    # second moment: (1 - 1/tau) Pi + (1/tau) (rho u* u* + rho/3 I) + (1 - 1/(2 tau)) (u* F + F u*)
    ux, uy = jx / r + 0.5 * fx / r, jy / r + 0.5 * fy / r
    a, b, c = 1 - 1 / tau0, 1 / tau0, 1 - 0.5 / tau0
    exx = a * pxx + b * (r * ux * ux + r / 3) + c * 2 * ux * fx
    eyy = a * pyy + b * (r * uy * uy + r / 3) + c * 2 * uy * fy
    exy = a * pxy + b * (r * ux * uy) + c * (ux * fy + uy * fx)
    for got, exp in ((pxx1, exx), (pyy1, eyy), (pxy1, exy)):
        np.testing.assert_allclose(got, exp, rtol=0, atol=2e-15)
    # the Eulerian force field carries the marker forces (sum phi = 1 per marker)
    np.testing.assert_allclose([fx.sum(), fy.sum()], [mk[:, 4].sum(), mk[:, 5].sum()], rtol=1e-13, atol=1e-15)


def test_bgk_les_guo_uses_base_tau(oracle):
    """With the hybrid LES the collision tau differs from node to node, while
    the Guo term keeps TauAtBeginning: momentum gain F (1 + 1/(2 tau_n) -
    1/(2 tau0)) per node. A Guo term with the node's tau would give F."""
    tau0 = 0.52
    pre, post, fx, fy, tau, _ = _one_forced_step(oracle, "bgk", tau0, les=True, cs=0.3, noise=2e-2)
    forced = (fx != 0) | (fy != 0)
    assert np.abs(tau[forced] - tau0).max() > 1e-4  # LES tau really differs where the force acts
    _, jx, jy, *_ = _moments(pre)
    _, jx1, jy1, *_ = _moments(post)
    g = 1 + 0.5 / tau - 0.5 / tau0
    np.testing.assert_allclose(jx1 - jx, fx * g, rtol=0, atol=2e-15)
    np.testing.assert_allclose(jy1 - jy, fy * g, rtol=0, atol=2e-15)
    assert np.abs(fx * (g - 1)).max() > 1e-8  # the test tells the two forms apart


@pytest.mark.parametrize("tau0", [0.51, 0.9])
def test_mrt_forced_momentum_is_guo_only(oracle, tau0):
    """MRT does not relax momentum, so the momentum gain is the Guo term's
    first moment alone: (1 - 1/(2 tau0)) F, mass unchanged."""
    pre, post, fx, fy, tau, _ = _one_forced_step(oracle, "mrt", tau0)
    r, jx, jy, *_ = _moments(pre)
    r1, jx1, jy1, *_ = _moments(post)
    c = 1 - 0.5 / tau0
    np.testing.assert_allclose(r1, r, rtol=0, atol=2e-15)
    np.testing.assert_allclose(jx1 - jx, c * fx, rtol=0, atol=2e-15)
    np.testing.assert_allclose(jy1 - jy, c * fy, rtol=0, atol=2e-15)


def test_direct_forcing_drives_markers_towards_target(oracle):
    """One direct-forcing step (F = 2 (U_d - U_interp), mesh.cpp:514-532)
    moves the interpolated marker velocity towards the desired one: from a
    fluid at rest the marker error shrinks after one forced step."""
    nx, ny = 64, 48
    ud = np.array([0.03, -0.02])
    s = oracle.Sim(nx, ny, "bgk", "periodic_bb", tau0=0.7)
    s.init_fields(3, 1.0, (0.0, 0.0), 0.0)
    s.add_body(31.2, 23.7, 14.0, 44, ud=tuple(ud))
    s.step(1)
    e0 = np.abs(s.markers()[:, 2:4] - ud).max(axis=0)
    np.testing.assert_allclose(e0, np.abs(ud), rtol=0, atol=1e-16)  # fluid at rest: U_interp = 0
    s.step(1)
    mk = s.markers()
    e1 = np.abs(mk[:, 2:4] - ud)
    assert np.all(e1.max(axis=0) < 0.8 * e0)
    # and the interpolated velocity moved in the direction of U_d at every marker
    assert np.all(mk[:, 2] * ud[0] > 0) and np.all(mk[:, 3] * ud[1] > 0)
    s.close()
