"""The GPU relaxation-time label search (ippi_tau_search, the candidate loop
of FindAndWriteTau, IPPI-LBM/mesh.cpp:2667-2704) against the oracle's
sequential loop: the chosen dtau and its score must be identical (FP64,
same operation order; the wave reduction keeps the first smallest score)."""
import numpy as np
import pytest

from oracle import pyoracle as o
from tests import _pkg
from tests.test_oracle_tausearch import EX, EY, make_case

pytestmark = pytest.mark.gpu


def test_bitwise_random_nodes_full_range():
    print("[tausearch] random 67 nodes, m in [-10000, 10000]", flush=True)
    ippi = _pkg.load()
    rng = np.random.default_rng(11)
    n = 67
    fnew, feq, tau = make_case(rng, n, tau0=0.55, spread=0.02)
    target = np.stack([1.0 + rng.uniform(-0.01, 0.01, n), rng.uniform(-0.05, 0.1, n), rng.uniform(-0.03, 0.03, n)], 1)
    bd_g, be_g = ippi.tau_search(fnew, feq, tau, target)
    bd_o, be_o = o.tau_search(fnew, feq, tau, target)
    assert np.array_equal(bd_g, bd_o) and np.array_equal(be_g, be_o)
    assert len(np.unique(bd_g)) > 5  # the optima are spread over the range


@pytest.mark.parametrize("mstar", [-9999, -64, 63, 9999])
def test_closed_form_optimum_gpu(mstar):
    print(f"[tausearch] closed form m*={mstar}", flush=True)
    ippi = _pkg.load()
    rng = np.random.default_rng(30000 + mstar)
    fnew, feq, tau = make_case(rng, 9)
    T = 0.6 + mstar * 1e-5
    target = np.stack([fnew.sum(1) + (feq - fnew).sum(1) / T, fnew @ EX + ((feq - fnew) @ EX) / T,
                       fnew @ EY + ((feq - fnew) @ EY) / T], axis=1)
    bd, be = ippi.tau_search(fnew, feq, tau, target)
    assert np.all(np.abs(bd - mstar * 1e-5) < 0.5e-5)
    bd_o, be_o = o.tau_search(fnew, feq, tau, target)
    assert np.array_equal(bd, bd_o) and np.array_equal(be, be_o)


def test_edges():
    print("[tausearch] edges", flush=True)
    ippi = _pkg.load()
    rng = np.random.default_rng(5)
    fnew, feq, tau = make_case(rng, 6, spread=0.01)
    target = np.array([[1.0, 0.01, 0.0], [1.0, 0.0, 0.01], [np.nan, 0.0, 0.0], [1.0, np.inf, 0.0],
                       [1.0, 0.0, 0.0], [0.99, 0.02, -0.01]])
    for lo, hi in ((12, 12), (-3, 70), (-10000, -9990)):
        g = ippi.tau_search(fnew, feq, tau, target, m_lo=lo, m_hi=hi)
        r = o.tau_search(fnew, feq, tau, target, m_lo=lo, m_hi=hi)
        assert np.array_equal(g[0], r[0]) and np.array_equal(g[1], r[1])
    assert np.all(g[0][2:4] == 1e10) and np.all(g[1][2:4] == 1e10)
    e = np.zeros((0, 9))
    bd, be = ippi.tau_search(e, e, e, np.zeros((0, 3)))
    assert bd.shape == (0,)
    with pytest.raises(ippi.IppiError):
        ippi.tau_search(fnew, feq, tau, target, m_lo=5, m_hi=4)


def test_both_division_paths():
    """Nodes whose candidate denominators tau + dtau or numerators feq leave
    2^-200 .. 2^200 (here: tau + dtau crossing zero, a denormal feq, a huge
    feq, a zero feq) divide plainly; the others share one refined
    reciprocal per population. Both must give the oracle's bits, wide
    sampling of tau and feq included."""
    print("[tausearch] division paths", flush=True)
    ippi = _pkg.load()
    rng = np.random.default_rng(17)
    n = 48
    fnew, feq, tau = make_case(rng, n, tau0=0.55, spread=0.02)
    tau[0, 3] = 0.04          # 0.04 + dtau crosses zero in [-0.1, 0.1]
    feq[1, 5] = 5e-310        # denormal numerator
    feq[2, 2] = 1e70          # huge numerator
    feq[3, 7] = 0.0           # zero numerator (the shared path)
    tau[4:12] = rng.uniform(0.15, 3.0, (8, 9))       # wide relaxation times
    feq[12:20] *= 10.0 ** rng.uniform(-30, 30, (8, 9))  # wide numerators
    target = np.stack([1.0 + rng.uniform(-0.01, 0.01, n), rng.uniform(-0.05, 0.1, n), rng.uniform(-0.03, 0.03, n)], 1)
    bd_g, be_g = ippi.tau_search(fnew, feq, tau, target)
    bd_o, be_o = o.tau_search(fnew, feq, tau, target)
    assert np.array_equal(bd_g, bd_o) and np.array_equal(be_g, be_o)
