"""Benchmark of the MI355X D2Q9 time step of IPPI-LBM (one JSON line).

Workload (the headline configuration): MRT collision + hybrid Smagorinsky
LES (Cs = 0.1), 2048 x 1024 nodes, one immersed-boundary cylinder of
diameter 80 lu (251 markers) at (512, 512), Re = 1e4 (tau0 = 0.5 + 3 U D/Re
= 0.50192), inflow U = 0.08 with seeded +-1e-3 velocity noise, non-
equilibrium-extrapolation inlet/outlet and free-slip walls; populations
stored in FP32 as f - w. A "step" is one full time step (IB marker kernel,
IB spread kernel, fused stream-collide kernel).

Reported:
  value         MLUPS of the FP32 path = nx*ny*steps / seconds / 1e6, with
                the state resident in HBM (stream synchronised on both
                sides of the timed region).
  fp64_mlups    the same workload on the plain FP64 path, and the ratio.
  accuracy      relative L2 error of velocity and density of the FP32 path
                against the FP64 GPU path after the timed steps (both start
                from the same seeded state), and of the drag record.
  roofline      achieved HBM bandwidth of the fused stream-collide kernel:
                72 algorithmic bytes per node (9 FP32 loads + 9 FP32
                stores) x nodes / average launch time, from HIP events on
                the library's stream; `traffic` from rocprof counters when
                available (IPPI_TRAFFIC_BYTES), else null.
  cpu_baseline  the FP64 CPU restatement (oracle/, OpenMP) on a bounded
                sample of the same workload on this host's cores.
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "machine-learning-based-implicit-physical-parametric-inference-for-cfd-simulation-promotion_amd")
METRIC = "MLUPS + % HBM roofline, FP32/NN-tau vs plain FP64 at ≤1e-5 rel L2 velocity"
HBM_PEAK_GBS = 8000.0

NX, NY = 2048, 1024
U0 = 0.08
DIAM, NMARK = 80.0, 251
CX, CY = 512.0, 512.0
RE = 1e4
TAU0 = 0.5 + 3.0 * U0 * DIAM / RE
CS = 0.1
SEED = 20240601


def load_pkg():
    if "ippi_amd" in sys.modules:
        return sys.modules["ippi_amd"]
    spec = importlib.util.spec_from_file_location("ippi_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    m = importlib.util.module_from_spec(spec)
    sys.modules["ippi_amd"] = m
    spec.loader.exec_module(m)
    return m


def make_solver(ippi, precision):
    s = ippi.Solver(NX, NY, "mrt", precision, "inflow", les=True, tau0=TAU0, cs=CS, u_in=(U0, 0.0))
    s.init_fields(SEED, 1.0, (U0, 0.0), 1e-3)
    s.add_body(CX, CY, DIAM, NMARK)
    return s


def timed_run(s, steps, warmup):
    s.step(warmup)
    s.synchronize()
    t0 = time.perf_counter()
    s.step(steps)
    s.synchronize()
    return time.perf_counter() - t0


def cpu_baseline(seconds_target=12.0):
    """FP64 oracle (OpenMP) on the same workload, bounded number of steps."""
    sys.path.insert(0, ROOT)
    from oracle import pyoracle
    threads = min(16, os.cpu_count() or 1)
    o = pyoracle.Sim(NX, NY, "mrt", "inflow", les=True, tau0=TAU0, cs=CS, u_in=(U0, 0.0), threads=threads)
    o.init_fields(SEED, 1.0, (U0, 0.0), 1e-3)
    o.add_body(CX, CY, DIAM, NMARK)
    t0 = time.perf_counter()
    o.step(1)
    one = time.perf_counter() - t0
    n = max(2, min(200, int(seconds_target / max(one, 1e-3))))
    t0 = time.perf_counter()
    o.step(n)
    dt = time.perf_counter() - t0
    o.close()
    return {"value": round(NX * NY * n / dt / 1e6, 2), "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": f"oracle/ FP64 MRT+LES+IB, {NX}x{NY}, {n} steps after 1 warm-up step, {threads} OpenMP threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--profile-steps", type=int, default=500,
                    help="steps of the event-timed pass used for the roofline figure")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the D2Q9 step runs the whole domain on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU x-slab design is described in DESIGN.md but not built).", file=sys.stderr)
        sys.exit(2)
    ippi = load_pkg()
    K, W = args.steps, args.warmup

    # FP32 shifted populations (the measured path)
    s32 = make_solver(ippi, "fp32")
    print(f"[bench] fp32 MRT+LES+IB {NX}x{NY}: warmup {W}, timed {K} steps", file=sys.stderr, flush=True)
    dt32 = timed_run(s32, K, W)
    mlups32 = NX * NY * K / dt32 / 1e6

    # plain FP64 path, same workload
    s64 = make_solver(ippi, "fp64")
    print(f"[bench] fp64 MRT+LES+IB {NX}x{NY}: warmup {W}, timed {K} steps", file=sys.stderr, flush=True)
    dt64 = timed_run(s64, K, W)
    mlups64 = NX * NY * K / dt64 / 1e6

    # accuracy of FP32 against FP64 after the same W + K steps
    r32, x32, y32 = s32.fields()
    r64, x64, y64 = s64.fields()
    eu = float(np.sqrt(np.sum((x32 - x64) ** 2 + (y32 - y64) ** 2) / np.sum(x64 ** 2 + y64 ** 2)))
    er = float(np.linalg.norm(r32 - r64) / np.linalg.norm(r64))
    d32, d64 = s32.body_forces(), s64.body_forces()
    n = min(len(d32), len(d64))
    drag_err = float(np.abs(d32[-n:, 0, 0] - d64[-n:, 0, 0]).mean() / np.abs(d64[-n:, 0, 0]).mean())
    s64.close()

    # event-timed pass for the fused kernel's average launch time
    P = args.profile_steps
    s32.set_timing(True)
    s32.step(P)
    step_ms, nstep, ib_ms, nib = s32.timing()
    s32.set_timing(False)
    s32.close()
    avg_ms = step_ms / max(nstep, 1)
    algo_bytes = 72 * NX * NY
    achieved = algo_bytes / (avg_ms * 1e-3) / 1e9
    traffic = os.environ.get("IPPI_TRAFFIC_BYTES")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": float(traffic) if traffic else None,
                "kernel": "k_step_f32<MRT,INFLOW,pull>", "avg_launch_us": round(avg_ms * 1e3, 3),
                "algorithmic_bytes_per_launch": algo_bytes,
                "ib_kernels_avg_us": round(ib_ms / max(nib, 1) * 1e3, 3)}

    cpu = None
    if not args.no_cpu:
        print("[bench] cpu baseline (oracle, OpenMP)", file=sys.stderr, flush=True)
        cpu = cpu_baseline()

    line = {
        "metric": METRIC, "value": round(mlups32, 1), "unit": "MLUPS", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": round(dt32 / K * 1e3, 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "MRT+LES(Cs=0.1)+IB cylinder D=80 (251 markers), 2048x1024, inflow 0.08 +-1e-3 "
                               "noise, Re=1e4, FP32 shifted populations (f-w)",
                   "nx": NX, "ny": NY, "collision": "mrt", "les_cs": CS, "tau0": TAU0, "u_in": U0,
                   "bodies": 1, "markers": NMARK, "bc": "inflow"},
        "fp64_mlups": round(mlups64, 1), "speedup_vs_fp64": round(mlups32 / mlups64, 3),
        "rel_l2_u_vs_fp64": eu, "rel_l2_rho_vs_fp64": er, "drag_rel_err_vs_fp64": drag_err,
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
