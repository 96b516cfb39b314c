"""Grid-level pins of the oracle: streaming, walls, LES tau, Peskin delta,
immersed-boundary interpolation/spreading, decay of a shear wave."""
import math

import numpy as np
import pytest

EX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
EY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])
OPP = np.array([0, 3, 4, 1, 2, 7, 8, 5, 6])
W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)


def test_stream_periodic_bb_by_roll(oracle):
    """Pull streaming checked with numpy.roll (periodic x) and the halfway
    bounce-back rule at the two walls; also a permutation of fcol."""
    nx, ny = 12, 7
    s = oracle.Sim(nx, ny, "bgk", "periodic_bb", tau0=0.8)
    s.init_fields(11, 1.0, (0.03, 0.01), 1e-2)
    s.step(1)
    fcol, fnew = s.fcol(), s.fnew()
    for k in range(9):
        exp = np.roll(np.roll(fcol[k], EX[k], axis=1), EY[k], axis=0)
        if EY[k] == 1:
            exp[0] = fcol[OPP[k]][0]
        if EY[k] == -1:
            exp[ny - 1] = fcol[OPP[k]][ny - 1]
        assert np.array_equal(fnew[k], exp), k
    assert np.array_equal(np.sort(fcol.ravel()), np.sort(fnew.ravel()))
    s.close()


def test_mass_conserved_periodic_bb(oracle):
    s = oracle.Sim(32, 16, "mrt", "periodic_bb", les=True, tau0=0.52, cs=0.1)
    s.init_fields(5, 1.0, (0.08, 0.0), 1e-3)
    m0 = s.mass()
    assert s.step(200) == 0
    assert abs(s.mass() - m0) / m0 < 1e-13
    s.close()


def test_inflow_walls_and_nee(oracle):
    nx, ny, uin = 24, 10, 0.05
    s = oracle.Sim(nx, ny, "bgk", "inflow", tau0=0.7, u_in=(uin, 0.0))
    s.init_fields(3, 1.0, (uin, 0.0), 1e-3)
    s.step(5)
    rho, ux, uy = s.fields()
    f = s.fnew()
    # specular free-slip rows: no normal momentum at the wall nodes (up to
    # the rounding of the reference's summation order)
    assert np.abs(uy[0, 1:-1]).max() < 1e-16 and np.abs(uy[ny - 1, 1:-1]).max() < 1e-16
    # inlet / outlet macros are the imposed ones
    assert np.all(rho[:, 0] == 1.0) and np.all(ux[:, 0] == uin) and np.all(uy[:, 0] == 0.0)
    assert np.all(rho[:, -1] == 1.0) and np.array_equal(ux[:, -1], ux[:, -2])
    # non-equilibrium extrapolation: f_b - feq_b == f_n - feq_n
    for y in range(ny):
        fb = oracle.feq(1.0, uin, 0.0)
        fn = oracle.feq(rho[y, 1], ux[y, 1], uy[y, 1])
        assert np.allclose(f[:, y, 0] - fb, f[:, y, 1] - fn, atol=1e-16)
    s.close()


def test_les_tau_identity(oracle):
    """Hybrid Smagorinsky tau solves tau (tau - tau0) = 4.5 Cs^2 |S| / rho with
    |S| = sqrt(2 S:S), S = -3/(2 rho tau0) Pi_neq; Cs = 0 or Pi_neq = 0 -> tau0."""
    rng = np.random.default_rng(7)
    for _ in range(40):
        rho, ux, uy = rng.uniform(0.9, 1.1), rng.uniform(-.1, .1), rng.uniform(-.1, .1)
        feq = oracle.feq(rho, ux, uy)
        f = feq + W * rng.uniform(-1e-2, 1e-2, 9)
        tau0, cs = rng.uniform(0.5001, 0.9), rng.uniform(0.05, 0.2)
        t = oracle.les_tau(f, rho, ux, uy, tau0, cs)
        neq = f - feq
        pxx, pxy, pyy = (neq * EX * EX).sum(), (neq * EX * EY).sum(), (neq * EY * EY).sum()
        c = 3.0 / (2 * rho * tau0)
        Q = math.sqrt(2 * c * c * (pxx * pxx + 2 * pxy * pxy + pyy * pyy))
        assert abs(t * (t - tau0) - 4.5 * cs * cs * Q / rho) < 1e-14
        assert t >= tau0
        assert oracle.les_tau(f, rho, ux, uy, tau0, 0.0) == tau0
        assert oracle.les_tau(feq, rho, ux, uy, tau0, cs) == pytest.approx(tau0, abs=1e-15)


def test_peskin_identities(oracle):
    for r in np.linspace(0.0, 1.0, 37):
        js = np.arange(-3, 5)
        ph = np.array([oracle.delta(r - j) for j in js])
        assert abs(ph.sum() - 1) < 1e-15
        assert abs(((r - js) * ph).sum()) < 1e-15
        assert abs((ph * ph).sum() - 3 / 8) < 1e-15
    assert oracle.delta(2.0) == 0.0 and oracle.delta(-2.5) == 0.0
    assert oracle.delta(0.0) == 0.5


def _ib_sim(oracle, ux, uy):
    ny, nx = ux.shape
    s = oracle.Sim(nx, ny, "bgk", "periodic_bb", tau0=0.8)
    rho = np.ones_like(ux)
    f = np.stack([np.array([oracle.feq(1.0, a, b)[k] for a, b in zip(ux.ravel(), uy.ravel())])
                  for k in range(9)])
    s.set_state(f, rho, ux, uy)
    s.add_body(20.3, 17.6, 12.0, 37, ud=(0.01, -0.02))
    s.step(1)
    return s


def test_ib_interpolates_linear_field(oracle):
    ny, nx = 36, 40
    Y, X = np.mgrid[0:ny, 0:nx].astype(float)
    ux = 0.01 + 1e-3 * X - 2e-3 * Y
    uy = -0.02 + 5e-4 * X + 1e-3 * Y
    s = _ib_sim(oracle, ux, uy)
    m = s.markers()
    assert np.allclose(m[:, 2], 0.01 + 1e-3 * m[:, 0] - 2e-3 * m[:, 1], atol=1e-15)
    assert np.allclose(m[:, 3], -0.02 + 5e-4 * m[:, 0] + 1e-3 * m[:, 1], atol=1e-15)
    # direct forcing: F = 2 (U_d - U_interp)
    assert np.allclose(m[:, 4], 2 * (0.01 - m[:, 2]), atol=0)
    assert np.allclose(m[:, 5], 2 * (-0.02 - m[:, 3]), atol=0)
    s.close()


def test_ib_spread_is_transpose(oracle):
    rng = np.random.default_rng(9)
    ny, nx = 36, 40
    ux = rng.uniform(-0.05, 0.05, (ny, nx))
    uy = rng.uniform(-0.05, 0.05, (ny, nx))
    s = _ib_sim(oracle, ux, uy)
    m = s.markers()
    fx, fy = s.force_field()
    # <I u, F> == <u, S F>
    lhs = (m[:, 2] * m[:, 4]).sum() + (m[:, 3] * m[:, 5]).sum()
    rhs = (ux * fx).sum() + (uy * fy).sum()
    assert abs(lhs - rhs) < 1e-15
    # total Eulerian force equals total marker force; drag record is minus it
    assert abs(fx.sum() - m[:, 4].sum()) < 1e-14 and abs(fy.sum() - m[:, 5].sum()) < 1e-14
    fr = s.forces()
    assert fr.shape == (1, 1, 2)
    assert abs(fr[0, 0, 0] + m[:, 4].sum()) < 1e-15
    s.close()


@pytest.mark.parametrize("coll", ["bgk", "mrt", "rr"])
def test_shear_wave_decay(oracle, coll):
    """u_x = U sin(pi (y + 1/2) / H) between halfway bounce-back walls decays
    at nu (pi/H)^2 with nu = (tau - 1/2)/3."""
    nx, ny, tau, U, steps = 4, 32, 0.8, 1e-4, 400
    Y = np.arange(ny)[:, None] * np.ones((1, nx))
    ux = U * np.sin(np.pi * (Y + 0.5) / ny)
    uy = np.zeros_like(ux)
    f = np.stack([np.array([oracle.feq(1.0, a, 0.0)[k] for a in ux.ravel()]) for k in range(9)])
    s = oracle.Sim(nx, ny, coll, "periodic_bb", tau0=tau)
    s.set_state(f, np.ones_like(ux), ux, uy)
    s.step(steps)
    _, ux2, _ = s.fields()
    nu = (tau - 0.5) / 3
    amp = (ux2[:, 0] * np.sin(np.pi * (np.arange(ny) + 0.5) / ny)).sum() / \
        (np.sin(np.pi * (np.arange(ny) + 0.5) / ny) ** 2).sum()
    expect = U * math.exp(-nu * (math.pi / ny) ** 2 * steps)
    assert abs(amp / expect - 1) < 5e-3
    s.close()


def _flatplate(oracle, nx=40, ny=12, plate_x=15, steps=7, u_in=0.06, noise=1e-3, coll="bgk"):
    s = oracle.Sim(nx, ny, coll, "flatplate", les=True, tau0=0.56, cs=0.173, u_in=(u_in, 0.0), plate_x=plate_x)
    s.init_fields(4, 1.0, (u_in, 0.0), noise)
    return s


def test_flatplate_uniform_inflow_is_steady(oracle):
    """No plate (plate_x = nx) and no noise: the uniform inflow is a fixed
    point of every rule (NEE inlet/outlet/top, specular bottom, untouched
    corners that hold feq of the same flow)."""
    nx, ny = 30, 10
    s = _flatplate(oracle, nx, ny, plate_x=nx, noise=0.0)
    s.step(40)
    rho, ux, uy = s.fields()
    assert np.abs(rho - 1).max() < 1e-14
    assert np.abs(ux - 0.06).max() < 1e-14 and np.abs(uy).max() < 1e-15
    s.close()


def test_flatplate_rules(oracle):
    """Each boundary rule of the reference's flat plate (mesh.cpp:228-356)
    checked on the oracle's state after a few steps of a noisy flow."""
    nx, ny, px, uin = 40, 12, 15, 0.06
    s = _flatplate(oracle, nx, ny, plate_x=px)
    s.step(1)
    f0, (r0, x0, y0) = s.fnew(), s.fields()
    corners = [(0, 0), (0, nx - 1)]
    c0 = {c: (f0[:, c[0], c[1]].copy(), r0[c], x0[c], y0[c]) for c in corners}
    s.step(6)
    f, fc = s.fnew(), s.fcol()
    rho, ux, uy = s.fields()

    def nee_ok(yb, xb, yn, xn, rb, ubx, uby):
        fb = oracle.feq(rb, ubx, uby)
        fq = oracle.feq(rho[yn, xn], ux[yn, xn], uy[yn, xn])
        assert np.allclose(f[:, yb, xb], fb + f[:, yn, xn] - fq, rtol=0, atol=1e-16), (yb, xb)
        assert rho[yb, xb] == rb and ux[yb, xb] == ubx and uy[yb, xb] == uby, (yb, xb)

    for y in range(1, ny - 1):
        nee_ok(y, 0, y, 1, 1.0, uin, 0.0)                               # inlet
        nee_ok(y, nx - 1, y, nx - 2, 1.0, ux[y, nx - 2], uy[y, nx - 2])  # outlet
    for x in range(nx):
        nee_ok(ny - 1, x, ny - 2, x, 1.0, ux[ny - 2, x], uy[ny - 2, x])  # top
    for x in range(px, nx - 1):
        nee_ok(0, x, 1, x, rho[1, x], 0.0, 0.0)                         # no-slip plate ('d')
    # specular free slip ('m') from the post-collision populations
    for x in range(1, px):
        exp = [fc[0, 0, x], fc[1, 0, x - 1], fc[4, 1, x], fc[3, 0, x + 1], fc[4, 1, x],
               fc[8, 1, x - 1], fc[7, 1, x + 1], fc[7, 1, x + 1], fc[8, 1, x - 1]]
        assert np.array_equal(f[:, 0, x], np.array(exp)), x
        assert abs(uy[0, x]) < 1e-16
    # interior pull is a plain shift of the post-collision populations
    for k in range(9):
        exp = fc[k, 1 - EY[k]:ny - 1 - EY[k], 1 - EX[k]:nx - 1 - EX[k]]
        assert np.array_equal(f[k, 1:ny - 1, 1:nx - 1], exp), k
    # the two bottom corners keep their state
    for c in corners:
        fc0, rc, xc, yc = c0[c]
        assert np.array_equal(f[:, c[0], c[1]], fc0)
        assert rho[c] == rc and ux[c] == xc and uy[c] == yc
    s.close()


def test_flatplate_boundary_layer_grows(oracle):
    """Physics: far upstream of the leading edge the free-slip wall leaves
    the inflow nearly undisturbed (the plate's upstream influence decays);
    behind it the no-slip plate slows the first row of fluid, more so
    further downstream (boundary-layer growth)."""
    nx, ny, px = 120, 24, 30
    s = oracle.Sim(nx, ny, "bgk", "flatplate", tau0=0.6, u_in=(0.05, 0.0), plate_x=px)
    s.init_fields(1, 1.0, (0.05, 0.0), 0.0)
    assert s.step(600) == 0
    _, ux, _ = s.fields()
    assert np.abs(ux[1, 3:9] - 0.05).max() < 1e-3
    assert np.all(np.diff(ux[1, 3:px]) < 0)              # slowing towards the leading edge
    assert ux[1, px + 20] < 0.04 and ux[1, px + 60] < ux[1, px + 20] + 1e-3
    assert np.all(ux[0, px:nx - 1] == 0.0)
    s.close()
