"""Pins of the oracle's effective-relaxation-time pairs (stress matching,
oracle.stress_pairs) by closed forms a plausible mistake would fail:

* uniform flow and uniform non-equilibrium stress at a uniform fine tau_f:
  the label is tau_c = 1/2 + (tau_f - 1/2)/2 (the same viscosity on a grid
  twice as coarse at the same lattice velocity) -- a wrong strain factor
  or a wrong stress / strain relation moves it;
* a fine relaxation time that varies inside the block: the label is
  1/4 + 1/(2 <1/tau_l>) -- using the base tau instead of each node's moves it;
* a subgrid velocity pattern with <u'u'> known in closed form: the label
  grows by eps^2 tau_f / (2 p) -- dropping the subgrid stress or its sign
  moves it;
* features: those of the C oracle's or_mlp_features on the restricted
  populations feq(rho_c, u_c) + s <f - feq>;
* exclusion boxes get weight 0, zero-strain blocks a weight at rounding level.
"""
import numpy as np
import pytest

EX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1], dtype=float)
EY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1], dtype=float)
W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)


def feq(rho, ux, uy):
    eu = EX[:, None, None] * ux + EY[:, None, None] * uy
    return W[:, None, None] * rho * (1 + 3 * eu + 4.5 * eu * eu - 1.5 * (ux * ux + uy * uy))


def neq_of(pxx, pxy, pyy, shape):
    """Hermite non-equilibrium part with zero mass and momentum and second
    moment (pxx, pxy, pyy) - (trace) / 3 contributions of the lattice."""
    q = (EX * EX - 1 / 3) * pxx + 2 * EX * EY * pxy + (EY * EY - 1 / 3) * pyy
    return np.broadcast_to((4.5 * W * q)[:, None, None], (9,) + shape).copy()


def state(nx, ny, ux, uy, p):
    return feq(np.ones((ny, nx)), ux, uy) + neq_of(*p, (ny, nx))


def test_uniform_state_gives_coarse_tau(oracle):
    nx, ny = 8, 6
    for tau_f, p in ((0.52, (2e-4, 1e-4, -3e-4)), (0.8, (0.0, 5e-4, 0.0)), (0.5012, (1e-3, 0.0, -1e-3))):
        fp = state(nx, ny, np.full((ny, nx), 0.05), np.full((ny, nx), -0.01), p)
        feat, lab, w, tau = oracle.stress_pairs(fp, np.full((ny, nx), tau_f), tau_f, tau_f)
        tau_c = 0.5 + (tau_f - 0.5) / 2
        assert np.all(w > 0)
        np.testing.assert_allclose(tau, tau_c, rtol=0, atol=1e-12)
        np.testing.assert_allclose(lab, 1 / tau_c, rtol=1e-12)


def test_block_varying_fine_tau(oracle):
    nx, ny = 4, 4
    tl = np.array([[0.51, 0.53, 0.51, 0.53], [0.56, 0.6, 0.56, 0.6]] * 2)
    fp = state(nx, ny, np.zeros((ny, nx)), np.zeros((ny, nx)), (0.0, 4e-4, 0.0))
    _, _, _, tau = oracle.stress_pairs(fp, tl, 0.55, 0.55)
    m = np.mean(1 / np.array([0.51, 0.53, 0.56, 0.6]))
    np.testing.assert_allclose(tau, 0.25 + 0.5 / m, rtol=0, atol=1e-12)
    # the base-tau answer differs by far more than the tolerance
    assert abs((0.25 + 0.5 / m) - (0.5 + (0.55 - 0.5) / 2)) > 1e-4


@pytest.mark.parametrize("eps", [0.0, 3e-3, 1e-2])
def test_subgrid_stress(oracle, eps):
    nx, ny, tau_f, pxy = 8, 4, 0.51, 2e-4
    X = np.arange(nx)[None, :] * np.ones((ny, 1))
    s = (-1.0) ** X  # zero mean over each 2x2 block, <s s> = 1
    ux, uy = 0.04 + eps * s, eps * s
    fp = state(nx, ny, ux, uy, (0.0, pxy, 0.0))
    _, _, w, tau = oracle.stress_pairs(fp, np.full((ny, nx), tau_f), tau_f, tau_f)
    tau_c = 0.5 + (tau_f - 0.5) / 2
    # T_xy = -(1 - 1/(2 tau_f)) p - eps^2, C_xy = -3 p / tau_f, tau = 1/2 + 3 T_xy / (2 C_xy)
    expect = tau_c + eps * eps * tau_f / (2 * pxy)
    np.testing.assert_allclose(tau, expect, rtol=0, atol=1e-10)
    np.testing.assert_allclose(w, 2 * (3 * pxy / tau_f) ** 2, rtol=1e-10)


def test_features_match_c_oracle(oracle):
    rng = np.random.default_rng(4)
    nx, ny, tau_f, tau_c = 6, 4, 0.53, 0.515
    ux = 0.05 + 0.01 * rng.standard_normal((ny, nx))
    uy = 0.01 * rng.standard_normal((ny, nx))
    fp = feq(1 + 1e-3 * rng.standard_normal((ny, nx)), ux, uy) + 1e-4 * rng.standard_normal((9, ny, nx)) * W[:, None, None]
    feat, _, _, _ = oracle.stress_pairs(fp, np.full((ny, nx), tau_f), tau_c, tau_f)
    s = 2 * tau_c / tau_f
    for J in range(ny // 2):
        for I in range(nx // 2):
            blk = fp[:, 2 * J:2 * J + 2, 2 * I:2 * I + 2].reshape(9, 4)
            nq = []
            for m in range(4):
                r, vx, vy = oracle.macro(blk[:, m])
                nq.append(blk[:, m] - oracle.feq(r, vx, vy))
            fa = blk.mean(1)
            r, vx, vy = oracle.macro(fa)
            fc = oracle.feq(r, vx, vy) + s * np.mean(nq, axis=0)
            ref = oracle.mlp_features(fc, r, vx, vy)
            np.testing.assert_allclose(feat[J * (nx // 2) + I], ref, rtol=1e-10, atol=1e-14)


def test_exclusion_and_zero_strain(oracle):
    nx, ny = 8, 8
    fp = state(nx, ny, np.full((ny, nx), 0.03), np.zeros((ny, nx)), (1e-4, 0.0, -1e-4))
    _, lab, w, _ = oracle.stress_pairs(fp, np.full((ny, nx), 0.52), 0.51, 0.52, excl=[(1, 1, 2, 2)])
    w = w.reshape(4, 4)
    assert np.all(w[1:3, 1:3] == 0) and np.all(lab.reshape(4, 4)[1:3, 1:3] == 0)
    assert np.count_nonzero(w) == 12
    fp0 = state(nx, ny, np.full((ny, nx), 0.03), np.zeros((ny, nx)), (0.0, 0.0, 0.0))
    _, lab0, w0, _ = oracle.stress_pairs(fp0, np.full((ny, nx), 0.52), 0.51, 0.52)
    assert np.all(w0 < 1e-25)  # rounding of feq only: no strain to fit
