"""Pins of the oracle's relaxation-time label search (FindAndWriteTau,
IPPI-LBM/mesh.cpp:2667-2704) by closed forms, not by its own formula.

With one relaxation time for all nine sources, the candidate's sums are
exactly A + B / (tau + dtau) with A = sum fnew and B = sum (feq - fnew)
(weighted by 1, e_x, e_y), because (1 - 1/T) f + feq / T = f + (feq - f) / T.
A target built from a chosen candidate m* must therefore come back as m*:
a swapped fnew/feq, a sign error in dtau or a wrong velocity component would
move the optimum."""
import numpy as np
import pytest

from oracle import pyoracle as o

EX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1], dtype=float)
EY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1], dtype=float)
W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)


def make_case(rng, n, tau0=0.6, spread=0.0):
    feq = W * (1 + rng.uniform(-0.02, 0.02, (n, 9)))
    fnew = feq + W * rng.uniform(-0.05, 0.05, (n, 9))
    tau = np.full((n, 9), tau0) + rng.uniform(-spread, spread, (n, 9))
    return fnew, feq, tau


@pytest.mark.parametrize("mstar", [-9000, -137, 0, 1, 4242, 10000])
def test_closed_form_optimum(mstar):
    rng = np.random.default_rng(20000 + mstar)
    n = 5
    fnew, feq, tau = make_case(rng, n)
    T = 0.6 + mstar * 1e-5
    target = np.stack([(fnew.sum(1) + (feq - fnew).sum(1) / T),
                       (fnew @ EX + ((feq - fnew) @ EX) / T),
                       (fnew @ EY + ((feq - fnew) @ EY) / T)], axis=1)
    bd, be = o.tau_search(fnew, feq, tau, target)
    assert np.all(np.abs(bd - mstar * 1e-5) < 0.5e-5), (bd, mstar)
    assert np.all(be < 1e-13)


def test_score_is_v_shaped_around_the_optimum():
    """Away from m*, the score grows like |B| |dtau - dtau*| / T^2: the
    search over a narrow window off the optimum returns its nearer edge."""
    rng = np.random.default_rng(7)
    fnew, feq, tau = make_case(rng, 3)
    T = 0.6 + 500e-5
    target = np.stack([fnew.sum(1) + (feq - fnew).sum(1) / T, fnew @ EX + ((feq - fnew) @ EX) / T,
                       fnew @ EY + ((feq - fnew) @ EY) / T], axis=1)
    bd, _ = o.tau_search(fnew, feq, tau, target, m_lo=-2000, m_hi=100)
    assert np.allclose(bd, 100e-5)
    bd, _ = o.tau_search(fnew, feq, tau, target, m_lo=900, m_hi=3000)
    assert np.allclose(bd, 900e-5)


def test_single_candidate_and_no_candidate():
    rng = np.random.default_rng(3)
    fnew, feq, tau = make_case(rng, 4, spread=0.01)
    target = np.array([[1.0, 0.01, 0.0], [1.0, 0.0, 0.01], [np.nan, 0.0, 0.0], [1.0, np.inf, 0.0]])
    bd, be = o.tau_search(fnew, feq, tau, target, m_lo=12, m_hi=12)
    assert np.all(bd[:2] == 12 * 1e-5) and np.all(be[:2] < 1)
    # NaN and infinite scores never beat the initial 1e10 (the reference's sentinel)
    assert np.all(bd[2:] == 1e10) and np.all(be[2:] == 1e10)
