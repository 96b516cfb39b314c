"""Running means on the device (ippi_accumulate_mean, ComputeAvgVelocity of
IPPI-LBM/mesh.cpp:1954-1974) against the same recurrence evaluated on the
host from per-step read-backs; and the facade's writers (WriteDatFluid,
WriteAvgu/MergeAvgu, WriteDatParticle, OutputCd; mesh.cpp:1180-1243,
2009-2099, 1283-1319, 1461-1526) against the binding's state."""
import os
import subprocess

import numpy as np
import pytest

from tests import _pkg

pytestmark = pytest.mark.gpu
BIN = os.path.join(_pkg.PKG_DIR, "facade", "ippi_main")


@pytest.mark.parametrize("precision,bc", [("fp64", "flatplate"), ("fp32", "inflow"), ("fp64", "periodic_bb")])
def test_running_mean_matches_host_recurrence(precision, bc):
    print(f"[mean] {precision} {bc}", flush=True)
    ippi = _pkg.load()
    with ippi.Solver(96, 40, "bgk", precision, bc, les=True, tau0=0.55, cs=0.17, u_in=(0.06, 0.0),
                     plate_x=30) as s:
        s.init_fields(9, 1.0, (0.06, 0.0), 1e-3)
        s.step(5)
        acc = [np.zeros((40, 96)) for _ in range(3)]
        for k in range(12):
            s.step(1)
            s.accumulate_mean(k)
            cur = s.fields()
            acc = [(a * k + c) / (k + 1) for a, c in zip(acc, cur)]
        for a, m in zip(acc, s.mean()):
            assert np.array_equal(a, m)


def test_facade_writers(tmp_path):
    assert os.path.exists(BIN), "facade driver not built (run __graft_entry__.build())"
    print("[facade] ippi_main check", flush=True)
    r = subprocess.run([BIN, "check"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    ippi = _pkg.load()
    nx, ny = 128, 64
    with ippi.Solver(nx, ny, "bgk", "fp64", "periodic_bb", tau0=0.6) as s:
        s.init_fields(7, 1.0, (0.05, 0.0), 1e-3)
        s.add_body(40.0, 32.0, 16.0, 51)
        acc = None
        for t in range(1, 51):
            s.step(1)
            if t >= 20:
                s.accumulate_mean(t - 20)
        rho, ux, uy = s.fields()
        ar, au, av = s.mean()
        tau = s.tau()
        mk = s.markers()
        fx = s.body_forces()[-1, 0]
    # fluid dump: X Y Density U V avgu avgv tau avgRho, x-major, y fastest
    d = np.loadtxt(tmp_path / "fluid" / "fluid_Merged_50.dat", skiprows=3)
    X, Y = d[:, 0].astype(int), d[:, 1].astype(int)
    for col, ref in ((2, rho), (3, ux), (4, uy), (5, au), (6, av), (7, tau), (8, ar)):
        np.testing.assert_allclose(d[:, col], ref[Y, X], rtol=1e-5, atol=1e-10)
    # averaged fields: rows 1 .. ny-2, Style 0 (periodic kind: no inlet/outlet)
    a = np.loadtxt(tmp_path / "fluid" / "fluid_avg_50.dat", skiprows=3)
    assert a.shape == (nx * (ny - 2), 8)
    X, Y = a[:, 0].astype(int), a[:, 1].astype(int)
    assert Y.min() == 1 and Y.max() == ny - 2 and np.all(a[:, 7] == 0)
    np.testing.assert_allclose(a[:, 5], au[Y, X], rtol=1e-5, atol=1e-10)
    # marker polyline: header, 51 nodes, 51 line elements closing on node 1
    with open(tmp_path / "particle" / "particle_t50.dat") as f:
        lines = f.read().splitlines()
    assert lines[2].startswith('ZONE T="Particle", N=51, E=51') and lines[3] == "Nodes=51, Elements=51"
    nodes = np.array([[float(v) for v in ln.split()] for ln in lines[4:55]])
    np.testing.assert_allclose(nodes[:, :2], mk[:, :2], rtol=1e-5)
    assert lines[55] == "1 2" and lines[104] == "50 51" and lines[105] == "1 51" and len(lines) == 106
    # drag / lift coefficients of the last force record
    t, cd, cl = np.loadtxt(tmp_path / "particle" / "CD_CL.dat")
    assert t == 50
    assert cd == pytest.approx(ippi.drag_coefficient(fx[0], 0.05, 8.0), rel=1e-5)
    assert cl == pytest.approx(ippi.drag_coefficient(fx[1], 0.05, 8.0), rel=1e-5, abs=1e-9)
