"""Host side of the training workflow (training.py), no GPU: the label
transform (molecular floor, eddy-viscosity scale) and the pair
collection's subsampling, with a stub in place of the solver."""
import numpy as np
import pytest

from tests import _pkg


@pytest.fixture(scope="module")
def training():
    _pkg.load()
    from ippi_amd import training
    return training


def test_eddy_labels(training):
    tau_c = 0.504
    tau = np.array([0.5005, 0.503, 0.504, 0.505, 0.52])
    for alpha in (1, 2, 8):
        got = 1.0 / training.eddy_labels(1.0 / tau, tau_c, alpha)
        exp = np.array([tau_c, tau_c, tau_c, tau_c + alpha * 0.001, tau_c + alpha * 0.016])
        np.testing.assert_allclose(got, exp, rtol=0, atol=1e-14)


class StubSolver:
    """Returns weights 0..n-1 (so the threshold keeps a known set) and counts steps."""

    def __init__(self, n):
        self.n, self.steps, self.calls = n, 0, 0

    def coarse_pairs(self, tau_coarse, criterion="step"):
        assert criterion == "stress"
        self.calls += 1
        w = np.arange(self.n, dtype=np.float32)
        feat = np.tile(w[:, None], (1, 6))
        return feat, np.full(self.n, 1.0 / 0.51, np.float32), w

    def step(self, n):
        self.steps += n


def test_collect_pairs_threshold_and_subsample(training):
    s = StubSolver(1000)
    X, Y, W = training.collect_pairs(s, 0.505, snaps=3, every=7, per_snap=100, min_weight=0.5, seed=1)
    assert s.calls == 3 and s.steps == 14  # no step after the last snapshot
    assert len(Y) == 300 and X.shape == (300, 6)
    assert np.all(X[:, 0] > 0.5 * 999)  # only weights above half the largest
    assert np.all(W <= 10.0) and np.isclose(np.median(W), 1.0)
    s = StubSolver(1)  # a single zero weight: nothing usable
    with pytest.raises(ValueError):
        training.collect_pairs(s, 0.505, snaps=2, every=1)


def test_calibrate_picks_lowest_error(training, monkeypatch):
    seen = []

    def fake_train(ippi, X, omega, w, **kw):
        seen.append(1.0 / omega[0])
        return np.array([1.0 / omega[0]]), 0, 0, 0
    monkeypatch.setattr(training, "train", fake_train)
    # the validation error is smallest for the network whose label tau is 0.504 + 4 * 0.002
    alpha, blob, errs = training.calibrate(None, None, np.array([1 / 0.506]), None, 0.504,
                                           lambda b: abs(b[0] - 0.512), alphas=(1, 2, 4, 8))
    assert alpha == 4 and abs(blob[0] - 0.512) < 1e-12 and len(errs) == 4
    np.testing.assert_allclose(seen, [0.506, 0.508, 0.512, 0.52])
