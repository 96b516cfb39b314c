"""The C++ Mesh facade (reference method names over the C ABI) driven by a
main.cpp written like the reference's, against the Python binding of the
same library: the drag/lift coefficients that OutputCd appends to
particle/CD_CL.dat (format of IPPI-LBM/mesh.cpp:1504-1514) must equal the
binding's force records normalised as mesh.cpp:1501-1502."""
import os
import subprocess

import numpy as np
import pytest

from tests import _pkg

pytestmark = pytest.mark.gpu
BIN = os.path.join(_pkg.PKG_DIR, "facade", "ippi_main")


def test_facade_driver_matches_binding(tmp_path):
    assert os.path.exists(BIN), "facade driver not built (run __graft_entry__.build())"
    print("[facade] ippi_main 600 200", flush=True)
    r = subprocess.run([BIN, "600", "200"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Initial Completed" in r.stdout
    rows = np.loadtxt(tmp_path / "particle" / "CD_CL.dat", ndmin=2)
    assert rows.shape == (3, 3) and list(rows[:, 0]) == [200, 400, 600]

    ippi = _pkg.load()
    U, R = 0.08, 40.0
    tau = 0.5 + 3.0 * U * (2 * R) / 1e4
    with ippi.Solver(2048, 1024, "mrt", "fp32", "inflow", les=True, tau0=tau, cs=0.1, u_in=(U, 0.0)) as s:
        s.init_fields(20240601, 1.0, (U, 0.0), 1e-3)
        s.add_body(512.0, 512.0, 2 * R, 251)
        s.step(600)
        rec = s.body_forces()
    for t, cd, cl in rows:
        fx, fy = rec[int(t) - 1, 0]
        assert cd == pytest.approx(ippi.drag_coefficient(fx, U, R), rel=1e-5, abs=1e-9)
        assert cl == pytest.approx(ippi.drag_coefficient(fy, U, R), rel=1e-5, abs=1e-9)
