"""High-fidelity (XFlow) data and relaxation-time labels through the Mesh
facade (IPPI-LBM/mesh.cpp:2189-2766), on synthetic CSV files in the
reference's column layout (the reference's own data files are not part of
it): ReadXflowdata's parsing, unit conversion, duplicate merging and crop;
ChangeTauFromXflow's tau = 0.5 + 3 nu; FindAndWriteTau's label (the GPU
search against the oracle's loop on the same inputs) and its tau update.
Also the search inputs themselves: ippi_get_stencil against the oracle's
post-streaming state."""
import math
import os
import subprocess

import numpy as np
import pytest

from oracle import pyoracle as oracle
from tests import _pkg

pytestmark = pytest.mark.gpu
BIN = os.path.join(_pkg.PKG_DIR, "facade", "ippi_main")
DV = 340.112


def row(X, Y, nu, rho, u, v, idx):
    """One data line: id, x, y, z, VelocityModule, StaticPressure, -, TotalPressure, -, viscosity, u, v, w, |u|
    with node (X, Y) at x = X mm - 0.4 m, y = Y mm."""
    xm, ym = X * 0.001 - 0.4, Y * 0.001
    sp = (rho - 1.0) * 101325
    ud, vd = u * DV * math.sqrt(3.0), v * DV * math.sqrt(3.0)
    return f"{idx},{xm:.3f},{ym:.3f},0.0,{math.hypot(ud, vd):.9g},{sp:.9g},7.0,{sp + 1:.9g},0.5,{nu:.9g}," \
           f"{ud:.9g},{vd:.9g},0.0,{math.hypot(ud, vd):.9g}"


def round_half_away(x):
    return math.copysign(math.floor(abs(x) + 0.5), x)


def parse_like_reference(path):
    pts = []
    with open(path) as f:
        next(f)
        for line in f:
            tok = line.rstrip("\n").split(",")
            if len(tok) < 14:  # the 13th number must be followed by a comma
                continue
            try:
                v = [float(t) for t in tok[:13]]
            except ValueError:
                continue
            pts.append({"vx": round_half_away(v[1] * 1000.0) / 1000.0, "vy": round_half_away(v[2] * 1000.0) / 1000.0,
                        "VM": v[4], "SP": v[5], "nu": v[9], "ux": v[10], "uy": v[11]})
    out, seen = [], {}
    for p in pts:
        vx = (p["vx"] + 0.4) / 0.001
        vy = p["vy"] / 0.001
        q = {"vx": vx, "vy": vy, "ux": p["ux"] / DV / math.sqrt(3.0), "uy": p["uy"] / DV / math.sqrt(3.0),
             "rho": (p["SP"] + 101325) / 101325, "nu": p["nu"], "VM": p["VM"] / DV / math.sqrt(3.0)}
        key = (vx, vy)
        if key not in seen:
            seen[key] = len(out)
            out.append(q)
        else:
            c = out[seen[key]]
            for k in ("ux", "uy", "rho", "nu", "VM"):
                c[k] = (c[k] + q[k]) / 2.0
    return [q for q in out if -2.5 <= q["vx"] <= 64.5]


def nu_of(X, Y):
    return 0.02 + 0.001 * X + 0.0005 * Y


def test_xflow_labels(tmp_path):
    assert os.path.exists(BIN), "facade driver not built"
    d = tmp_path / "data"
    d.mkdir()
    lines = ["id,x,y,z,VelocityModule,StaticPressure,p,TotalPressure,q,viscosity,u,v,w,speed"]
    idx = 0
    for X in range(16, 25):
        for Y in range(18, 23):
            lines.append(row(X, Y, nu_of(X, Y), 1.0 + 1e-4 * X, 0.05 + 1e-4 * Y, 1e-4 * X, idx))
            idx += 1
    lines.append(row(17, 19, nu_of(17, 19) + 0.004, 1.002, 0.06, 0.001, idx))  # duplicate: merged
    lines.append("x,not,a,number")
    lines.append(row(18, 18, 0.5, 1.0, 0.0, 0.0, idx + 1).rsplit(",", 1)[0])  # no field after w: skipped
    lines.append(row(200, 5, 0.1, 1.0, 0.0, 0.0, idx + 2))  # outside the columns: cropped
    (d / "pb_10024.csv").write_text("\n".join(lines) + "\n")
    target = (1.0012, 0.0480, 0.0011)
    (d / "pb_10030.csv").write_text("\n".join(lines[:1] + [row(20, 20, 0.03, *target, 0)]) + "\n")
    print("[xflow] facade run", flush=True)
    r = subprocess.run([BIN, "xflow", str(d / "pb_")], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    # ReadXflowdata
    got = np.loadtxt(tmp_path / "xflow_points.txt", ndmin=2)
    ref = parse_like_reference(d / "pb_10024.csv")
    assert got.shape == (len(ref), 7) and len(ref) == 45
    want = np.array([[q["vx"], q["vy"], q["ux"], q["uy"], q["rho"], q["nu"], q["VM"]] for q in ref])
    assert np.array_equal(got, want)
    # FindAndWriteTau: the label from the same inputs, and tau = 0.5 + 3 nu + label on the nine nodes
    sten = np.loadtxt(tmp_path / "xflow_stencil.txt")
    tgt = parse_like_reference(d / "pb_10030.csv")[0]
    bd, _ = oracle.tau_search(sten[:, 0][None], sten[:, 1][None], sten[:, 2][None],
                              np.array([[tgt["rho"], tgt["ux"], tgt["uy"]]]))
    line = (tmp_path / "cp" / "BestTau.dat").read_text().split()
    assert int(line[0]) == 15 and float(line[1]) == pytest.approx(bd[0], abs=1e-12)
    base = {}
    for q in ref:
        base[(round_half_away(q["vx"]), round_half_away(q["vy"]))] = 0.5 + 3.0 * q["nu"]
    EX = [0, 1, 0, -1, 0, 1, -1, -1, 1]
    EY = [0, 0, 1, 0, -1, 1, 1, -1, -1]
    taus = [float(t) for t in line[2:11]]
    for k in range(9):
        assert taus[k] == pytest.approx(base[(20 + EX[k], 20 + EY[k])] + bd[0], abs=5e-6)
    assert "BestTau:" in r.stderr


@pytest.mark.parametrize("les", [False, True])
def test_stencil_matches_oracle_fp64(les):
    """The search inputs after t steps are the oracle's post-streaming state
    after t - 1 steps (the fused kernel streams step t - 1 and collides step
    t): pulled populations, their equilibrium and relaxation time, bitwise."""
    print(f"[stencil] les={les}", flush=True)
    ippi = _pkg.load()
    nx, ny, t = 48, 32, 12
    kw = dict(les=les, tau0=0.57, cs=0.15, u_in=(0.05, 0.0))
    g = ippi.Solver(nx, ny, "bgk", "fp64", "inflow", **kw)
    o = oracle.Sim(nx, ny, "bgk", "inflow", **kw)
    g.init_fields(9, 1.0, (0.05, 0.0), 1e-3)
    o.init_fields(9, 1.0, (0.05, 0.0), 1e-3)
    g.step(t)
    o.step(t - 1)
    nodes = np.array([[2, 2], [10, 7], [24, 16], [45, 29]])
    fn, fq, tk = g.stencil(nodes)
    F = o.fnew()
    EX = [0, 1, 0, -1, 0, 1, -1, -1, 1]
    EY = [0, 0, 1, 0, -1, 1, 1, -1, -1]
    for i, (x, y) in enumerate(nodes):
        for k in range(9):
            sx, sy = x - EX[k], y - EY[k]
            f = F[:, sy, sx].copy()
            rho, ux, uy = oracle.macro(f)
            q = oracle.feq(rho, ux, uy)
            assert fn[i, k] == f[k] and fq[i, k] == q[k]
            tau = oracle.les_tau(f, rho, ux, uy, 0.57, 0.15) if les else 0.57
            assert tk[i, k] == tau
    with pytest.raises(ippi.IppiError):
        g.stencil([[1, 5]])
    g.close()


def test_stencil_fp32_and_state():
    """FP32 shifted populations: the stencil adds the weights back in FP64
    and agrees with the FP64 oracle to the FP32 path's budget; before any
    pull step there is no collision input to report."""
    print("[stencil] fp32", flush=True)
    ippi = _pkg.load()
    nx, ny, t = 48, 32, 20
    kw = dict(les=True, tau0=0.56, cs=0.12, u_in=(0.05, 0.0))
    g = ippi.Solver(nx, ny, "bgk", "fp32", "inflow", **kw)
    g.init_fields(4, 1.0, (0.05, 0.0), 1e-3)
    with pytest.raises(ippi.IppiError):
        g.stencil([[10, 10]])
    o = oracle.Sim(nx, ny, "bgk", "inflow", **kw)
    o.init_fields(4, 1.0, (0.05, 0.0), 1e-3)
    g.step(t)
    o.step(t - 1)
    nodes = np.array([[x, y] for x in (3, 17, 40) for y in (2, 15, 29)])
    fn, fq, tk = g.stencil(nodes)
    F = o.fnew()
    EX = [0, 1, 0, -1, 0, 1, -1, -1, 1]
    EY = [0, 0, 1, 0, -1, 1, 1, -1, -1]
    want = np.array([[F[k, y - EY[k], x - EX[k]] for k in range(9)] for x, y in nodes])
    assert np.linalg.norm(fn - want) / np.linalg.norm(want) < 1e-6
    assert np.linalg.norm(fq - want) / np.linalg.norm(want) < 1e-2  # near equilibrium
    assert np.all(np.abs(tk - 0.56) < 0.02) and np.all(tk >= 0.56)
    g.close()
