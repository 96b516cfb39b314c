"""Benchmark of the MI355X D2Q9 time step of IPPI-LBM (one JSON line).

Headline workload (`value`): the FP32 fused MRT + hybrid Smagorinsky LES
(Cs = 0.1) + immersed-boundary step with the relaxation time inferred by the
fused network (6 -> 32 -> 32 -> 1, SiLU, softplus + 0.5, bf16 MFMA with
FP32 accumulation, guard >= 0.501) at every node, on 4096 x 2048 nodes: one
cylinder of diameter 160 lu (503 markers) at (1024, 1024), Re = 1e4
(tau0 = 0.5 + 3 U D / Re = 0.50384), inflow U = 0.08 with seeded +-1e-3
velocity noise, non-equilibrium-extrapolation inlet/outlet and free-slip
walls, populations stored in FP32 as deviations f - feq(1, (0.08, 0)) from
the free-stream equilibrium. The network's weights are the seeded Xavier
draw of workloads.xavier_blob(7) (W1, W2 rounded to bf16). A "step" is one full time
step: one fused stream-collide kernel, which also computes the immersed
boundary and the network.

Reported beside it:
  fp64_mlups        the plain FP64 path on the same grid and body (MRT + LES +
                    IB, no network), and speedup_vs_fp64 = value / it;
  fp64_nn_mlups     the FP64 path with the same network (evaluated in FP64);
  plain_fp32        the FP32 path without the network on the same grid, with
                    and without the body: MLUPS, fused-kernel time and its
                    fraction of the HBM roofline (72 B/node);
  rel_l2_u_vs_fp64  FP32 against the FP64 path with the network after the
                    timed steps (same seeded start); accuracy_gate_met
                    uses the 10,000-step figure of the accuracy leg when
                    that leg ran, else this one (accuracy_gate_steps says which);
  roofline          achieved HBM bandwidth of the fused kernel: 72 algorithmic
                    bytes per node (9 FP32 loads + 9 FP32 stores) x nodes /
                    average launch time, from one HIP event pair on the
                    library's stream around the timed steps; `traffic` = HBM
                    bytes per launch from the committed rocprofv3 FETCH_SIZE /
                    WRITE_SIZE summary, scaled by the same counters on a plain
                    copy of known size;
  cpu_baseline      the FP64 CPU restatement (oracle/, OpenMP) on a bounded
                    sample of the same workload on this host's cores, and
                    (other_workloads) on C1, C2 and the flat plate;
  legs              the other workloads, each measured the same way:
                    accuracy_nn_4096x2048  the headline workload over 10,000
                                   steps: FP32 vs FP64 (both with the
                                   network), FP64 over FP32-rounded storage,
                                   tau error, and the plain-bf16 network;
                    c3_rr_trained_nn_4096x2048  RR with the network trained
                                   on the GPU on pairs from a 2x finer run
                                   (8192 x 4096): pairs, training, FP32
                                   MLUPS, FP32 vs FP64 over 10,000 steps,
                                   and the a-posteriori gain over the
                                   constant-tau and Smagorinsky fallbacks;
                    mrt_les_ib_2048x1024   MRT + LES + IB without the network,
                                   D = 80, Re = 1e4: FP32 and FP64 MLUPS,
                                   roofline, and the FP32-vs-FP64 error over
                                   20,000 steps;
                    adjoint_1024x512  BGK + IB discrete adjoint of the mean drag
                                   over steps 2001-4000, checkpoints every 100,
                                   tau field 0.6 +- 0.02;
                    multibody_8192x4096  MRT + LES + network tau, 16 cylinders
                                   D ~ U[40, 120] (seeded placement), FP32 vs
                                   FP64, 5,000 steps, per-body drag error;
                    flatplate_2401x501  the reference's own run (flat plate,
                                   BGK + LES, Re = 5e5), also with its
                                   boundary-row / column networks set;
                    c1_bgk_channel_256x128  BGK channel with a cylinder,
                                   FP64 and FP32, 10,000 steps;
                    tausearch      FindAndWriteTau's candidate search.

The headline's W warm-up and K timed steps run after the legs, right after
untimed steps of a second solver of the same workload (100-step chunks
until the per-step time stops improving, clock_warm_steps in the line):
the network kernel is issue-bound and runs ~25% slow for its first ~150
steps after an idle gap while the shader clock ramps up.

`python bench.py --copy-probe MB` runs only a plain device copy of MB
megabytes (the calibration program for the rocprofv3 traffic counters).
"""
import argparse
import importlib.util
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "machine-learning-based-implicit-physical-parametric-inference-for-cfd-simulation-promotion_amd")
METRIC = "MLUPS + % HBM roofline, FP32/NN-tau vs plain FP64 at ≤1e-5 rel L2 velocity"
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_traffic_k_step_f32_ib_nn_4096x2048.md")

# headline: MRT + LES + IB + network tau, 4096 x 2048
NX, NY = 4096, 2048
U0 = 0.08
DIAM = 160.0
NMARK = int(round(3.141592653589793 * DIAM))  # 503 markers, ~1 lu apart
CX, CY = NX / 4.0, NY / 2.0
RE = 1e4
TAU0 = 0.5 + 3.0 * U0 * DIAM / RE
CS = 0.1
SEED = 20240601
NET_SEED = 7

# MRT + LES + IB without the network, 2048 x 1024 (leg mrt_les_ib_2048x1024)
NX2, NY2, DIAM2, NMARK2, CX2, CY2 = 2048, 1024, 80.0, 251, 512.0, 512.0
TAU02 = 0.5 + 3.0 * U0 * DIAM2 / RE


def load_pkg():
    if "ippi_amd" in sys.modules:
        return sys.modules["ippi_amd"]
    spec = importlib.util.spec_from_file_location("ippi_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    m = importlib.util.module_from_spec(spec)
    sys.modules["ippi_amd"] = m
    spec.loader.exec_module(m)
    return m


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def timed_run(s, steps, warmup):
    s.step(warmup)
    s.synchronize()
    t0 = time.perf_counter()
    s.step(steps)
    s.synchronize()
    return time.perf_counter() - t0


def vel_err(a, b):
    return float(np.sqrt(np.sum((a[1] - b[1]) ** 2 + (a[2] - b[2]) ** 2) / np.sum(b[1] ** 2 + b[2] ** 2)))


def kernel_avg_us(s, steps):
    """Average fused-kernel and IB-kernel launch times over `steps` event-timed steps
    (HIP events on the solver's stream around every launch)."""
    s.set_timing(True)
    s.step(steps)
    step_ms, nstep, ib_ms, nib = s.timing()
    s.set_timing(False)
    return step_ms / max(nstep, 1) * 1e3, ib_ms / max(nib, 1) * 1e3


def timed_run_events(s, steps, warmup):
    """timed_run with one HIP event pair on the solver's stream around the
    timed steps (recorded inside the call, so the launches stay back to
    back): (wall seconds, average launch time in us of the step kernel, or
    None when a step holds more than that one kernel)."""
    s.step(warmup)
    s.synchronize()
    s.set_timing("calls")
    t0 = time.perf_counter()
    s.step(steps)
    s.synchronize()
    dt = time.perf_counter() - t0
    step_ms, nstep, _, nib = s.timing()
    s.set_timing(False)
    return dt, (None if nib else step_ms / max(nstep, 1) * 1e3)


# ------------------------------------------------------------------ headline
def blob():
    from ippi_amd import workloads
    return workloads.xavier_blob(NET_SEED)


def make_headline(ippi, precision, nn=True, tau_mode="nn", noise=1e-3, force_history=0):
    """The headline workload; nn=False gives the plain path (no network)."""
    s = ippi.Solver(NX, NY, "mrt", precision, "inflow", les=True, tau_mode=tau_mode if nn else "const",
                    tau0=TAU0, cs=CS, u_in=(U0, 0.0), force_history=force_history)
    s.init_fields(SEED, 1.0, (U0, 0.0), noise)
    s.add_body(CX, CY, DIAM, NMARK)
    if nn:
        s.set_mlp(blob())
    return s


def read_traffic():
    """HBM bytes per launch of the headline fused kernel from the committed
    rocprofv3 counter summary (None when it is absent)."""
    try:
        with open(TRAFFIC_FILE) as f:
            m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", f.read())
        return float(m.group(1)) if m else None
    except OSError:
        return None


def cpu_baseline(seconds_target=15.0):
    """FP64 oracle (OpenMP) on the headline workload, bounded number of steps."""
    sys.path.insert(0, ROOT)
    from oracle import pyoracle
    threads = min(16, os.cpu_count() or 1)
    o = pyoracle.Sim(NX, NY, "mrt", "inflow", les=True, tau_mode="nn", tau0=TAU0, cs=CS, u_in=(U0, 0.0),
                     threads=threads)
    o.init_fields(SEED, 1.0, (U0, 0.0), 1e-3)
    o.add_body(CX, CY, DIAM, NMARK)
    o.set_mlp(blob())
    t0 = time.perf_counter()
    o.step(1)
    one = time.perf_counter() - t0
    n = max(2, min(200, int(seconds_target / max(one, 1e-3))))
    t0 = time.perf_counter()
    o.step(n)
    dt = time.perf_counter() - t0
    o.close()
    out = {"value": round(NX * NY * n / dt / 1e6, 2), "unit": "MLUPS", "cores": threads, "kind": "port",
           "sample": f"oracle/ FP64 MRT+LES+IB+network tau, {NX}x{NY}, {n} steps after 1 warm-up step, "
                     f"{threads} OpenMP threads"}
    # the same oracle on the other CPU-sized workloads (a few seconds each)
    others = {}
    for name, (nx, ny, coll, bc, kw, body) in {
        "bgk_channel_cyl_256x128": (256, 128, "bgk", "periodic_bb", dict(tau0=0.6, u_in=(0.05, 0.0)), (64.0, 64.0, 20.0, 63)),
        "mrt_les_ib_2048x1024": (2048, 1024, "mrt", "inflow", dict(les=True, tau0=0.5 + 3 * 0.08 * 80 / 1e4, cs=0.1,
                                                                    u_in=(0.08, 0.0)), (512.0, 512.0, 80.0, 251)),
        "flatplate_2401x501": (2401, 501, "bgk", "flatplate", dict(les=True, tau0=0.5 + 6.0 * (501 * 0.1) / 5e5,
                                                                    cs=0.173, u_in=(0.1, 0.0), plate_x=400), None),
    }.items():
        o = pyoracle.Sim(nx, ny, coll, bc, threads=threads, **kw)
        o.init_fields(1, 1.0, kw["u_in"], 1e-3)
        if body:
            o.add_body(*body)
        t0 = time.perf_counter()
        o.step(1)
        one = time.perf_counter() - t0
        m = max(2, min(2000, int(4.0 / max(one, 1e-4))))
        t0 = time.perf_counter()
        o.step(m)
        dt = time.perf_counter() - t0
        o.close()
        others[name] = {"mlups": round(nx * ny * m / dt / 1e6, 2), "steps": m}
    out["other_workloads"] = others
    return out


# ---------------------------------------------------------------- other legs
def clock_warm(ippi, min_steps=300, max_steps=3000, chunk=100):
    """A second headline solver stepped in `chunk`-step calls (HIP events
    around each) until two chunks in a row come within 0.5% of the fastest
    so far (at least `min_steps`, at most `max_steps`): (solver, steps)."""
    sw = make_headline(ippi, "fp32")
    best, calm, done = None, 0, 0
    while done < max_steps:
        sw.set_timing("calls")
        sw.step(chunk)
        sw.synchronize()
        ms, n, _, _ = sw.timing()
        sw.set_timing(False)
        done += chunk
        t = ms / max(n, 1)
        calm = calm + 1 if best is not None and t >= 0.995 * best else 0
        best = t if best is None else min(best, t)
        if done >= min_steps and calm >= 2:
            break
    return sw, done


def leg_accuracy_nn(ippi, steps):
    """The headline workload over `steps` steps: FP32 against FP64 (both
    with the network), beside FP64 arithmetic over FP32-rounded storage
    (IPPI_FP64_ROUND_STORE=1: the storage format alone), the relaxation
    time's error, the network's acceptance, and the plain-bf16 network
    variant after 1,000 steps."""
    out = {"workload": "headline workload (MRT+LES+IB+network tau, 4096x2048)", "steps": steps}
    runs = {"fp32": make_headline(ippi, "fp32"), "fp64": make_headline(ippi, "fp64")}
    os.environ["IPPI_FP64_ROUND_STORE"] = "1"  # read when the handle is created
    try:
        runs["fp64_fp32_storage"] = make_headline(ippi, "fp64")
    finally:
        del os.environ["IPPI_FP64_ROUND_STORE"]
    marks = [m for m in (1000, 2000, 5000, 10000, 20000) if m < steps] + [steps]
    out["marks"] = marks
    out["fp32_vs_fp64"], out["fp64_fp32_storage_vs_fp64"], out["tau_fp32_vs_fp64"] = [], [], []
    done = 0
    for m in marks:
        log(f"accuracy (network, {NX}x{NY}): to step {m}")
        for s in runs.values():
            s.step(m - done)
        done = m
        f = {k: s.fields() for k, s in runs.items()}
        out["fp32_vs_fp64"].append(vel_err(f["fp32"], f["fp64"]))
        out["fp64_fp32_storage_vs_fp64"].append(vel_err(f["fp64_fp32_storage"], f["fp64"]))
        t32, t64 = runs["fp32"].tau(), runs["fp64"].tau()
        out["tau_fp32_vs_fp64"].append(float(np.linalg.norm(t32 - t64) / np.linalg.norm(t64)))
    out["network_tau_accepted_frac"] = float(np.mean(runs["fp64"].tau() > 0.501))
    r32, r64 = runs["fp32"].body_forces(), runs["fp64"].body_forces()
    n = min(len(r32), len(r64)) // 2
    out["mean_drag_rel_err_last_half"] = float(abs(r32[-n:, 0, 0].mean() - r64[-n:, 0, 0].mean()) /
                                               abs(r64[-n:, 0, 0].mean()))
    for s in runs.values():
        s.close()
    out["rel_l2_u_vs_fp64"] = out["fp32_vs_fp64"][-1]
    # the plain-bf16 network (IPPI_TAU_NN_BF16): speed and error after 1,000 steps
    a, c = make_headline(ippi, "fp32", tau_mode="nn_bf16"), make_headline(ippi, "fp64")
    a.step(1000)
    c.step(1000)
    e_u, e_t = vel_err(a.fields(), c.fields()), float(np.linalg.norm(a.tau() - c.tau()) / np.linalg.norm(c.tau()))
    dtb, kb = timed_run_events(a, 500, 10)
    a.close()
    c.close()
    out["bf16_plain"] = {"rel_l2_u_vs_fp64_at_1000": e_u, "rel_l2_tau_vs_fp64_at_1000": e_t,
                         "mlups_fp32": round(NX * NY * 500 / dtb / 1e6, 1),
                         "fused_kernel_avg_us": round(kb, 2) if kb else None}
    return out


def leg_c3(ippi, steps=10000, val_steps=500, gain_steps=1000, pre=6000, snaps=16, per_snap=131072, iters=3000):
    """RR collision with the network trained on the GPU on pairs the solver
    generates from a 2x finer run (training.py): the fine run (8192 x 4096,
    D = 320, RR + LES, tau_f = 1/2 + 2 (tau0 - 1/2)) gives stress-matching
    pairs at `snaps` instants 50 fine steps apart after `pre` steps (up to
    `per_snap` each); the trainer fits the network (Adam, W1 / W2 rounded
    to bf16 after) for eddy-viscosity scales 1, 2, 4, 8, 16, the scale
    being chosen on a validation window (`val_steps` coarse steps from a
    restricted fine state); then, on a later evaluation window, the
    a-posteriori velocity error against the restricted fine solution after
    `gain_steps` coarse steps beside the constant-tau and Smagorinsky
    fallbacks, and the 4096 x 2048 run itself from the seeded start: FP32
    MLUPS and FP32 vs FP64 after `steps` steps."""
    from ippi_amd import training
    from ippi_amd import workloads as wl
    out = {"workload": "C3: RR + network tau trained on the GPU on stress-matching pairs from a 2x finer run "
                       "(8192x4096, D=320), applied on 4096x2048, cylinder D=160, inflow 0.08, Re=1e4, FP32",
           "steps": steps}
    tau_f = 0.5 + 2.0 * (TAU0 - 0.5)
    fine = ippi.Solver(2 * NX, 2 * NY, "rr", "fp32", "inflow", les=True, tau0=tau_f, cs=CS, u_in=(U0, 0.0))
    fine.init_fields(SEED + 1, 1.0, (U0, 0.0), 1e-3)
    fine.add_body(2 * CX, 2 * CY, 2 * DIAM, wl.markers_for(2 * DIAM))
    t0 = time.perf_counter()
    log(f"C3: fine run {2 * NX}x{2 * NY}, {pre} steps, then {snaps} pair snapshots")
    fine.step(pre)
    X, Y, W = training.collect_pairs(fine, TAU0, snaps, 50, per_snap=per_snap, seed=11)
    out["pairs"] = int(len(Y))
    out["label_tau_pct_5_50_95"] = [float(v) for v in np.percentile(1.0 / Y, [5, 50, 95])]
    out["pair_generation_s"] = round(time.perf_counter() - t0, 2)

    def coarse(prec, mode, les, fc=None, net=None, noise=1e-3):
        s = ippi.Solver(NX, NY, "rr", prec, "inflow", les=les, tau_mode=mode, tau0=TAU0, cs=CS, u_in=(U0, 0.0))
        if fc is not None:
            s.set_populations(fc)
        else:
            s.init_fields(SEED, 1.0, (U0, 0.0), noise)
        s.add_body(CX, CY, DIAM, NMARK)
        if net is not None:
            s.set_mlp(net)
        return s

    def err(fc, target, mode, les, net, n):
        s = coarse("fp32", mode, les, fc, net)
        try:
            s.step(n)
            return training.velocity_error(s.fields(), target)
        except ippi.IppiError:
            return float("inf")
        finally:
            s.close()

    def window(n):
        fc, _ = fine.restrict(TAU0)
        fine.step(2 * n)
        _, (_, xf, yf) = fine.restrict(TAU0, populations=False)
        return fc, (xf, yf)

    fine.step(400)
    fc_v, tgt_v = window(val_steps)
    t0 = time.perf_counter()
    log(f"C3: training on {len(Y)} pairs, eddy scale chosen on a {val_steps}-step validation window")
    alpha, net, verr = training.calibrate(ippi, X, Y, W, TAU0, lambda b: err(fc_v, tgt_v, "nn", True, b, val_steps),
                                          seed=3, iters=iters)
    del fc_v
    _, l0, l1, l2 = training.train(ippi, X, training.eddy_labels(Y, TAU0, alpha), W, seed=3, iters=iters)
    out["training"] = {"adam_iters": iters, "batch": 8192, "eddy_scale_validation_error": verr,
                       "eddy_scale": alpha, "loss_start": l0, "loss_trained": l1, "loss_bf16_weights": l2,
                       "seconds_all_scales": round(time.perf_counter() - t0, 2)}
    fine.step(400)
    log(f"C3: evaluation window, {gain_steps} coarse steps")
    fc, tgt = window(gain_steps)
    fine.close()
    e = {"const": err(fc, tgt, "const", False, None, gain_steps), "les": err(fc, tgt, "const", True, None, gain_steps),
         "nn+const": err(fc, tgt, "nn", False, net, gain_steps), "nn+les": err(fc, tgt, "nn", True, net, gain_steps)}
    del fc
    out["aposteriori"] = {"coarse_steps": gain_steps, "velocity_error_vs_restricted_fine":
                          {k: (None if not np.isfinite(v) else v) for k, v in e.items()},
                          "gain_vs_const": e["const"] / e["nn+const"], "gain_vs_les": e["les"] / e["nn+les"]}
    # the C3 run itself: seeded start, FP32 vs FP64, both with the trained network (Smagorinsky fallback),
    # beside FP64 from a start perturbed by 1 part in 1e6 of the noise (how fast the flow amplifies any
    # perturbation) and FP64 arithmetic over FP32-rounded storage (the storage format alone)
    r32, r64 = coarse("fp32", "nn", True, net=net), coarse("fp64", "nn", True, net=net)
    rp = coarse("fp64", "nn", True, net=net, noise=1e-3 * (1 + 1e-6))
    os.environ["IPPI_FP64_ROUND_STORE"] = "1"  # read when the handle is created
    try:
        rs = coarse("fp64", "nn", True, net=net)
    finally:
        del os.environ["IPPI_FP64_ROUND_STORE"]
    marks = [m for m in (1000, 5000) if m < steps] + [steps]
    out["marks"], out["fp32_vs_fp64"], out["tau_fp32_vs_fp64"] = marks, [], []
    out["fp64_perturbed_vs_fp64"], out["fp64_fp32_storage_vs_fp64"] = [], []
    done = 0
    for m in marks:
        log(f"C3: to step {m}")
        for s in (r32, r64, rp, rs):
            s.step(m - done)
        done = m
        f64 = r64.fields()
        out["fp32_vs_fp64"].append(vel_err(r32.fields(), f64))
        out["fp64_perturbed_vs_fp64"].append(vel_err(rp.fields(), f64))
        out["fp64_fp32_storage_vs_fp64"].append(vel_err(rs.fields(), f64))
        t32, t64 = r32.tau(), r64.tau()
        out["tau_fp32_vs_fp64"].append(float(np.linalg.norm(t32 - t64) / np.linalg.norm(t64)))
    rp.close()
    rs.close()
    out["network_tau_accepted_frac"] = float(np.mean(t64 > 0.501))
    d32, d64 = r32.body_forces(), r64.body_forces()
    n = min(len(d32), len(d64)) // 2
    out["mean_drag_rel_err_last_half"] = float(abs(d32[-n:, 0, 0].mean() - d64[-n:, 0, 0].mean()) /
                                               abs(d64[-n:, 0, 0].mean()))
    dt, k_us = timed_run_events(r32, 500, 10)
    out["mlups_fp32"] = round(NX * NY * 500 / dt / 1e6, 1)
    out["fused_kernel_avg_us"] = round(k_us, 2) if k_us else None
    r64.close()
    r32.close()
    out["rel_l2_u_vs_fp64"] = out["fp32_vs_fp64"][-1]
    return out


def leg_mrt_les_ib(ippi, steps, warmup, profile_steps, err_steps=20000):
    """MRT + LES + IB without the network on 2048 x 1024 (D = 80, Re = 1e4):
    FP32 and FP64 MLUPS, the fused kernel's roofline, and the FP32-vs-FP64
    velocity error over 20,000 steps beside FP64 over FP32-rounded storage
    and FP64 from a start perturbed by 1 part in 1e6 of the noise."""
    def make(prec, noise=1e-3):
        s = ippi.Solver(NX2, NY2, "mrt", prec, "inflow", les=True, tau0=TAU02, cs=CS, u_in=(U0, 0.0))
        s.init_fields(SEED, 1.0, (U0, 0.0), noise)
        s.add_body(CX2, CY2, DIAM2, NMARK2)
        return s
    out = {"workload": "MRT+LES(Cs=0.1)+IB cylinder D=80 (251 markers), 2048x1024, inflow 0.08 +-1e-3, Re=1e4",
           "steps": steps}
    for prec in ("fp32", "fp64"):
        s = make(prec)
        log(f"mrt+les+ib 2048x1024 {prec}: {steps} steps")
        dt, k_timed = timed_run_events(s, steps, warmup)
        out[f"mlups_{prec}"] = round(NX2 * NY2 * steps / dt / 1e6, 1)
        if prec == "fp32":
            k_each, _ = kernel_avg_us(s, profile_steps)
            k_us = k_timed if k_timed is not None else k_each
            gbs = 72 * NX2 * NY2 / (k_us * 1e-6) / 1e9
            out.update({"fused_kernel_avg_us": round(k_us, 3), "fused_kernel_gbs": round(gbs, 1),
                        "fused_kernel_roofline_frac": round(gbs / HBM_PEAK_GBS, 4)})
        s.close()
    out["speedup_vs_fp64"] = round(out["mlups_fp32"] / out["mlups_fp64"], 3)
    marks = [m for m in (1000, 2000, 5000, 10000, 20000) if m <= err_steps]
    if not marks:
        return out
    runs = {"fp32": make("fp32"), "fp64": make("fp64"), "fp64_perturbed": make("fp64", 1e-3 * (1 + 1e-6))}
    os.environ["IPPI_FP64_ROUND_STORE"] = "1"  # read when the handle is created
    try:
        runs["fp64_fp32_storage"] = make("fp64")
    finally:
        del os.environ["IPPI_FP64_ROUND_STORE"]
    out["error_steps"] = list(marks)
    for k in ("fp32_vs_fp64", "fp64_fp32_storage_vs_fp64", "fp64_perturbed_vs_fp64"):
        out[k] = []
    done = 0
    for m in marks:
        log(f"accuracy (2048x1024): to step {m}")
        for s in runs.values():
            s.step(m - done)
        done = m
        f = {k: s.fields() for k, s in runs.items()}
        out["fp32_vs_fp64"].append(vel_err(f["fp32"], f["fp64"]))
        out["fp64_fp32_storage_vs_fp64"].append(vel_err(f["fp64_fp32_storage"], f["fp64"]))
        out["fp64_perturbed_vs_fp64"].append(vel_err(f["fp64_perturbed"], f["fp64"]))
    for s in runs.values():
        s.close()
    return out


def leg_adjoint(ippi):
    from ippi_amd import adjoint
    nx, ny = 1024, 512
    nu = 0.1 / 3
    tau = 0.6 + np.random.default_rng(4).uniform(-0.02, 0.02, (ny, nx))
    c = (256.47, 256.27)
    out = {"workload": "BGK+IB adjoint, 1024x512, cylinder D=40 (126 markers), tau 0.6+-0.02 per node, Re=200, "
                       "4000 steps, checkpoints every 100, J = mean drag over steps 2001-4000"}
    for name, prec, f64 in (("fp32", "fp32", False), ("fp32_fwd_fp64_adj", "fp32", True), ("fp64", "fp64", False)):
        case = adjoint.DragCase(nx=nx, ny=ny, diameter=40.0, markers=126, nsteps=4000, avg_from=2001, ckpt=100,
                                precision=prec, fp64_adjoint=f64, seed=20240607, u0=(200 * nu / 40.0, 0.0))
        log(f"adjoint {name}")
        with case.solver(tau, c) as w:  # warm-up: first launches of every kernel
            w.adjoint_drag(20, 1, 10, fp64_adjoint=f64)
        with case.solver(tau, c) as s:
            t0 = time.perf_counter()
            J, gt, gc = s.adjoint_drag(case.nsteps, case.avg_from, case.ckpt, fp64_adjoint=f64)
            wall = time.perf_counter() - t0
        # the all-node adjoint kernel timed alone: the recompute of the next segment, which otherwise
        # runs beside the sweep on a second stream, runs between the sweeps (IPPI_ADJ_OVERLAP=0)
        os.environ["IPPI_ADJ_OVERLAP"] = "0"  # read when the handle is created
        try:
            s = case.solver(tau, c)
        finally:
            del os.environ["IPPI_ADJ_OVERLAP"]
        with s:
            s.set_timing(True)
            t0 = time.perf_counter()
            J1, gt1, gc1 = s.adjoint_drag(case.nsteps, case.avg_from, case.ckpt, fp64_adjoint=f64)
            wall_serial = time.perf_counter() - t0
            ms, n = s.adjoint_timing()
            s.set_timing(False)
        same = bool(J1 == J and np.array_equal(gt1, gt) and np.array_equal(gc1, gc))
        per = 36 if prec == "fp32" else 72  # forward state read per node
        lam = 72 if (prec == "fp64" or f64) else 36
        tau_b = 4 if prec == "fp32" else 8
        acc = 8 if (prec == "fp64" or f64) else 4
        bytes_node = per + 2 * lam + tau_b + 2 * acc
        us = ms / max(n, 1) * 1e3
        gbs = bytes_node * nx * ny / (us * 1e-6) / 1e9
        out[name] = {"J": J, "dJ_dcenter": [float(gc[0][0]), float(gc[0][1])], "wall_s": round(wall, 3),
                     "wall_s_recompute_not_overlapped": round(wall_serial, 3), "overlap_bitwise_equal": same,
                     "adjoint_kernel_avg_us": round(us, 2), "adjoint_kernel_bytes_per_node": bytes_node,
                     "adjoint_kernel_gbs": round(gbs, 1), "adjoint_kernel_roofline_frac": round(gbs / HBM_PEAK_GBS, 4),
                     "_gt": gt}
    g64 = out["fp64"].pop("_gt")
    for k in ("fp32", "fp32_fwd_fp64_adj"):
        gt = out[k].pop("_gt")
        out[k]["tau_grad_rel_l2_vs_fp64"] = float(np.linalg.norm(gt - g64) / np.linalg.norm(g64))
        out[k]["center_grad_rel_err_vs_fp64"] = [abs(a - b) / abs(b) for a, b in
                                                 zip(out[k]["dJ_dcenter"], out["fp64"]["dJ_dcenter"])]
    return out


def leg_multibody(ippi, steps, warmup):
    from ippi_amd import workloads
    nx, ny, U = 8192, 4096, 0.06
    bodies = workloads.place_cylinders(nx, ny, 16, 40.0, 120.0, seed=11)
    nm = sum(b[3] for b in bodies)
    dmean = float(np.mean([b[2] for b in bodies]))
    tau0 = 0.5 + 3 * U * dmean / 1e4
    blob = workloads.xavier_blob(7)
    out = {"workload": f"MRT+LES(Cs=0.1)+network tau, 8192x4096, 16 cylinders D~U[40,120] ({nm} markers), "
                       f"inflow 0.06 +-1e-3, tau0={tau0:.5f}", "steps": steps, "markers": nm}
    res = {}
    for prec in ("fp32", "fp64"):
        s = ippi.Solver(nx, ny, "mrt", prec, "inflow", les=True, tau_mode="nn", tau0=tau0, cs=CS, u_in=(U, 0.0),
                        force_history=steps + warmup)
        s.init_fields(SEED, 1.0, (U, 0.0), 1e-3)
        for (cx, cy, d, m) in bodies:
            s.add_body(cx, cy, d, m)
        s.set_mlp(blob)
        log(f"multibody 8192x4096 {prec}: {steps} steps")
        dt = timed_run(s, steps, warmup)
        res[prec] = (nx * ny * steps / dt / 1e6, s.fields(), s.body_forces())
        s.close()
    f32, f64 = res["fp32"][2], res["fp64"][2]
    n = min(len(f32), len(f64)) // 2
    d32, d64 = f32[-n:, :, 0].mean(0), f64[-n:, :, 0].mean(0)
    err = np.abs(d32 - d64) / np.abs(d64)
    out.update({"mlups_fp32": round(res["fp32"][0], 1), "mlups_fp64": round(res["fp64"][0], 1),
                "speedup_vs_fp64": round(res["fp32"][0] / res["fp64"][0], 3),
                "rel_l2_u_vs_fp64": vel_err(res["fp32"][1], res["fp64"][1]),
                "drag_rel_err_vs_fp64_per_body": [float(e) for e in err], "drag_rel_err_max": float(err.max()),
                "drag_window": f"mean over the last {n} records"})
    return out


def leg_tausearch(ippi, n=65536, n_cpu=48):
    """FindAndWriteTau's candidate search (20,001 candidates, FP64) for n
    nodes on the GPU, next to the oracle's sequential loop on a sample of
    n_cpu nodes (one core), with the chosen candidates compared."""
    sys.path.insert(0, ROOT)
    from oracle import pyoracle
    rng = np.random.default_rng(SEED)
    w = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
    feq = w * (1 + rng.uniform(-0.02, 0.02, (n, 9)))
    fnew = feq + w * rng.uniform(-0.05, 0.05, (n, 9))
    tau = 0.55 + rng.uniform(-0.02, 0.02, (n, 9))
    target = np.stack([1.0 + rng.uniform(-0.01, 0.01, n), rng.uniform(-0.05, 0.1, n),
                       rng.uniform(-0.03, 0.03, n)], 1)
    ippi.tau_search(fnew[:256], feq[:256], tau[:256], target[:256])  # warm-up
    t0 = time.perf_counter()
    bd, be = ippi.tau_search(fnew, feq, tau, target)
    gpu_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    bo, eo = pyoracle.tau_search(fnew[:n_cpu], feq[:n_cpu], tau[:n_cpu], target[:n_cpu])
    cpu_s = time.perf_counter() - t0
    cand = 20001
    log(f"tau search: {n} nodes on the GPU in {gpu_s * 1e3:.1f} ms; {n_cpu} nodes on one CPU core in {cpu_s:.2f} s")
    return {"workload": f"FindAndWriteTau candidate loop (m = -10000..10000, step 1e-5, FP64), {n} nodes, "
                        "seeded populations and targets",
            "gpu_ms_incl_transfers": round(gpu_s * 1e3, 2),
            "gpu_candidates_per_s": round(n * cand / gpu_s, 1),
            "cpu_candidates_per_s_one_core": round(n_cpu * cand / cpu_s, 1),
            "speedup_vs_one_core": round((n * cand / gpu_s) / (n_cpu * cand / cpu_s), 1),
            "cpu_sample_bitwise_equal": bool(np.array_equal(bd[:n_cpu], bo) and np.array_equal(be[:n_cpu], eo))}


def leg_c1(ippi, steps=10000):
    """C1: BGK (tau = 0.6) channel, 256 x 128, periodic in x with halfway
    bounce-back walls, cylinder D = 20 (63 markers) at (64, 64), U = 0.05,
    10,000 steps: FP64 (the reference's precision) and FP32 MLUPS, and the
    FP32 velocity against FP64 at the end."""
    nx, ny = 256, 128
    out = {"workload": "BGK tau=0.6 channel 256x128, periodic x + bounce-back walls, cylinder D=20 (63 markers) "
                       "at (64, 64), U=0.05 +-1e-3", "steps": steps}
    runs = {}
    for prec in ("fp64", "fp32"):
        s = ippi.Solver(nx, ny, "bgk", prec, "periodic_bb", tau0=0.6, u_in=(0.05, 0.0))
        s.init_fields(SEED, 1.0, (0.05, 0.0), 1e-3)
        s.add_body(64.0, 64.0, 20.0, 63)
        log(f"C1 {prec}: {steps} steps")
        dt = timed_run(s, steps, 0)
        out[f"mlups_{prec}"] = round(nx * ny * steps / dt / 1e6, 1)
        runs[prec] = s
    out["rel_l2_u_fp32_vs_fp64"] = vel_err(runs["fp32"].fields(), runs["fp64"].fields())
    for s in runs.values():
        s.close()
    return out


def leg_flatplate(ippi, steps=5000, marks=(1000, 5000)):
    """The reference's own run (IPPI-LBM/main.cpp): flat plate on
    2401 x 501, BGK + hybrid LES (Cs = 0.173), U = 0.1, Re = 5e5
    (tau = 0.5 + 6 Ny U / Re = 0.5006), free slip ahead of x = 400 and no
    slip beyond, uniform start. MLUPS of FP32 and FP64 over `steps`
    steps, and the FP32-vs-FP64 velocity error next to FP64 vs FP64 from a
    start perturbed by 1e-9 noise: the flow amplifies either perturbation."""
    nx, ny, u = 2401, 501, 0.1
    tau = 0.5 + 6.0 * (ny * u) / 5e5
    out = {"workload": f"flat plate 2401x501 (the reference's main.cpp run), BGK+LES(Cs=0.173), U=0.1, Re=5e5, "
                       f"tau={tau:.7f}, no-slip from x=400; with and without the boundary-row / column "
                       "networks (types 1, 3, 5, 7; seeded weights)", "steps": steps}

    def make(prec, noise):
        s = ippi.Solver(nx, ny, "bgk", prec, "flatplate", les=True, tau0=tau, cs=0.173, u_in=(u, 0.0), plate_x=400)
        s.init_fields(1, 1.0, (u, 0.0), noise)
        return s
    from ippi_amd import workloads
    for prec in ("fp32", "fp64"):
        s = make(prec, 0.0)
        log(f"flat plate {prec}: {steps} steps")
        dt = timed_run(s, steps, 20)
        out[f"mlups_{prec}"] = round(nx * ny * steps / dt / 1e6, 1)
        s.close()
        # the reference's boundary-row / column networks (types 1, 3, 5, 7) at its 5,800 edge nodes
        s = make(prec, 0.0)
        for t in (1, 3, 5, 7):
            s.set_boundary_mlp(t, workloads.boundary_blob(t))
        log(f"flat plate {prec} with the boundary networks: {steps // 5} steps")
        dtb = timed_run(s, steps // 5, 20)
        out[f"mlups_{prec}_boundary_networks"] = round(nx * ny * (steps // 5) / dtb / 1e6, 1)
        out[f"boundary_networks_us_per_step_{prec}"] = round((dtb / (steps // 5) - dt / steps) * 1e6, 2)
        s.close()
    out["speedup_vs_fp64"] = round(out["mlups_fp32"] / out["mlups_fp64"], 3)
    runs = {"fp32": make("fp32", 0.0), "fp64": make("fp64", 0.0), "fp64_noise1e-9": make("fp64", 1e-9)}
    out["error_steps"] = list(marks)
    out["rel_l2_u_fp32_vs_fp64"], out["rel_l2_u_fp64_noise1e-9_vs_fp64"] = [], []
    done = 0
    for mk in marks:
        for r in runs.values():
            r.step(mk - done)
        done = mk
        f = {k: r.fields() for k, r in runs.items()}
        out["rel_l2_u_fp32_vs_fp64"].append(vel_err(f["fp32"], f["fp64"]))
        out["rel_l2_u_fp64_noise1e-9_vs_fp64"].append(vel_err(f["fp64_noise1e-9"], f["fp64"]))
    for r in runs.values():
        r.close()
    return out


def copy_probe(mb):
    import torch
    n = int(mb * (1 << 20)) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(10):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": 4 * n, "launches": 10}), flush=True)


def pattern_probe(ippi):
    """Bandwidth ceiling of the fused step's memory pattern: the same nine
    pulled loads and nine stores with no arithmetic, in the plain and the
    XCD-grouped block order, next to a plain nine-array copy, at the
    headline and network-leg sizes and the reference's flat plate."""
    out = {}
    for (nx, ny) in ((2048, 1024), (4096, 2048), (8192, 4096), (2401, 501)):
        for prec in ("fp32", "fp64"):
            bpn = 72 if prec == "fp32" else 144
            for mode in ("pull", "pull_xcd", "copy", "copy16"):
                us = ippi.pattern_probe(nx, ny, prec, mode, reps=300)
                gbs = bpn * nx * ny / (us * 1e-6) / 1e9
                out[f"{nx}x{ny}_{prec}_{mode}"] = {"us": round(us, 3), "gbs": round(gbs, 1),
                                                   "frac_of_8tbs": round(gbs / 8000.0, 4)}
                log(f"pattern probe {nx}x{ny} {prec} {mode}: {us:.2f} us, {gbs:.0f} GB/s")
    print(json.dumps({"pattern_probe": out}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparisons of the headline")
    ap.add_argument("--legs", default="accuracy,c3,mrtles,adjoint,multibody,flatplate,c1,tausearch",
                    help="comma list of extra workloads (accuracy, c3, mrtles, adjoint, multibody, flatplate, "
                         "c1, tausearch) or 'none'")
    ap.add_argument("--c3-steps", type=int, default=10000, help="steps of the C3 run (FP32 vs FP64)")
    ap.add_argument("--acc-steps", type=int, default=10000, help="steps of the headline accuracy leg")
    ap.add_argument("--mrtles-steps", type=int, default=2000)
    ap.add_argument("--mrtles-err-steps", type=int, default=20000,
                    help="last step of the MRT+LES+IB leg's FP32-vs-FP64 error run (0: none)")
    ap.add_argument("--mb-steps", type=int, default=5000)
    ap.add_argument("--plate-steps", type=int, default=5000,
                    help="steps of the flat-plate leg (its FP32-vs-FP64 errors at 1,000 / 5,000 when reached)")
    ap.add_argument("--profile-steps", type=int, default=200,
                    help="steps of the per-launch event-timed pass reported beside the roofline figure")
    ap.add_argument("--copy-probe", type=float, default=0.0, metavar="MB",
                    help="only run a plain device copy of MB megabytes (counter calibration)")
    ap.add_argument("--pattern-probe", action="store_true",
                    help="only time the fused step's memory pattern with no arithmetic (bandwidth ceiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the D2Q9 step runs the whole domain on one GPU; --gpus {args.gpus} is not supported "
              "(a multi-GPU x-slab split is a design, not built).", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe > 0:
        copy_probe(args.copy_probe)
        return
    ippi = load_pkg()
    ippi.library()
    if args.pattern_probe:
        pattern_probe(ippi)
        return
    K, W = args.steps, args.warmup

    # the other workloads first: the headline's warm-up and timed steps then run on a GPU at its
    # steady clocks (a short timed region right after start-up reads the ramp-up, not the kernel)
    legs = {}
    names = [x for x in args.legs.split(",") if x and x != "none"]
    if "accuracy" in names:
        legs["accuracy_nn_4096x2048"] = leg_accuracy_nn(ippi, args.acc_steps)
    if "c3" in names:
        legs["c3_rr_trained_nn_4096x2048"] = leg_c3(ippi, args.c3_steps)
    if "mrtles" in names:
        legs["mrt_les_ib_2048x1024"] = leg_mrt_les_ib(ippi, args.mrtles_steps, W, args.profile_steps,
                                                       args.mrtles_err_steps)
    if "adjoint" in names:
        legs["adjoint_1024x512"] = leg_adjoint(ippi)
    if "multibody" in names:
        legs["multibody_8192x4096"] = leg_multibody(ippi, args.mb_steps, W)
    if "flatplate" in names:
        legs["flatplate_2401x501"] = leg_flatplate(ippi, args.plate_steps,
                                                   tuple(m for m in (1000, 5000) if m <= args.plate_steps))
    if "c1" in names:
        legs["c1_bgk_channel_256x128"] = leg_c1(ippi)
    if "tausearch" in names:
        legs["tausearch"] = leg_tausearch(ippi)

    s32 = make_headline(ippi, "fp32")
    # shader-clock warm-up: the issue-bound network kernel runs ~25% slow for its first ~150 steps after
    # the host-side set-up leaves the GPU idle (the memory-bound plain step does not), so a second
    # solver of the same workload runs untimed steps right before the W warm-up steps
    sw, warm_steps = clock_warm(ippi)
    log(f"fp32 MRT+LES+IB+network {NX}x{NY}: warmup {W}, timed {K} steps (after {warm_steps} clock warm-up steps)")
    dt32, k_us_timed = timed_run_events(s32, K, W)
    sw.close()
    mlups32 = NX * NY * K / dt32 / 1e6
    extra = {}
    if not args.no_fp64:
        p64 = make_headline(ippi, "fp64", nn=False)
        log(f"fp64 plain MRT+LES+IB {NX}x{NY}: warmup {W}, timed {K} steps")
        dtp = timed_run(p64, K, W)
        p64.close()
        mlups_p64 = NX * NY * K / dtp / 1e6
        s64 = make_headline(ippi, "fp64")
        log(f"fp64 MRT+LES+IB+network {NX}x{NY}: warmup {W}, timed {K} steps")
        dt64 = timed_run(s64, K, W)
        mlups64 = NX * NY * K / dt64 / 1e6
        r32, r64 = s32.fields(), s64.fields()
        d32, d64 = s32.body_forces(), s64.body_forces()
        n = min(len(d32), len(d64))
        extra = {"fp64_mlups": round(mlups_p64, 1), "speedup_vs_fp64": round(mlups32 / mlups_p64, 3),
                 "fp64_nn_mlups": round(mlups64, 1), "speedup_vs_fp64_nn": round(mlups32 / mlups64, 3),
                 "rel_l2_u_vs_fp64": vel_err(r32, r64), "rel_l2_u_vs_fp64_steps": K + W,
                 "rel_l2_rho_vs_fp64": float(np.linalg.norm(r32[0] - r64[0]) / np.linalg.norm(r64[0])),
                 "rel_l2_tau_vs_fp64": float(np.linalg.norm(s32.tau() - s64.tau()) / np.linalg.norm(s64.tau())),
                 "drag_rel_err_vs_fp64": float(np.abs(d32[-n:, 0, 0] - d64[-n:, 0, 0]).mean() /
                                               np.abs(d64[-n:, 0, 0]).mean())}
        s64.close()
    # the plain FP32 step (no network) on the same grid, with the body (IB inside the step kernel) and
    # without it: MLUPS and the fused kernel's HBM fraction at 72 algorithmic bytes per node
    plain = {}
    for tag, body in (("with_body", True), ("no_body", False)):
        sp = ippi.Solver(NX, NY, "mrt", "fp32", "inflow", les=True, tau0=TAU0, cs=CS, u_in=(U0, 0.0))
        sp.init_fields(SEED, 1.0, (U0, 0.0), 1e-3)
        if body:
            sp.add_body(CX, CY, DIAM, NMARK)
        log(f"fp32 plain MRT+LES{'+IB' if body else ''} {NX}x{NY}: warmup {W}, timed {K} steps")
        dtq, kq = timed_run_events(sp, K, W)
        sp.close()
        plain[tag] = {"mlups": round(NX * NY * K / dtq / 1e6, 1), "kernel_avg_us": round(kq, 2) if kq else None,
                      "roofline_frac": round(72 * NX * NY / (kq * 1e-6) / 8e12, 4) if kq else None}
    extra["plain_fp32"] = plain

    # the roofline uses the timed steps' own launches (HIP events around the timed call, the kernels back
    # to back); per-launch events (each launch bracketed) are reported beside it
    k_us_each, ib_us = kernel_avg_us(s32, args.profile_steps)
    k_us = k_us_timed if k_us_timed is not None else k_us_each
    s32.close()
    algo_bytes = 72 * NX * NY
    achieved = algo_bytes / (k_us * 1e-6) / 1e9
    # the same memory pattern with no arithmetic (nine pulled loads, nine non-temporal stores): the
    # ceiling the kernel's traffic can reach at this size
    probe_us = ippi.pattern_probe(NX, NY, "fp32", "pull", reps=200)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": read_traffic(),
                "kernel": "k_step_f32_ib<MRT,INFLOW,NN> (fused step, IB and network inside)" if ib_us == 0.0
                          else "k_step_f32<MRT,INFLOW,pull,NN>", "avg_launch_us": round(k_us, 3),
                "avg_launch_us_timed_individually": round(k_us_each, 3),
                "algorithmic_bytes_per_launch": algo_bytes, "ib_kernel_avg_us": round(ib_us, 3),
                "pattern_probe_us": round(probe_us, 3), "frac_of_pattern_probe": round(probe_us / k_us, 4)}

    if "accuracy_nn_4096x2048" in legs:
        gate_err, gate_steps = legs["accuracy_nn_4096x2048"]["rel_l2_u_vs_fp64"], args.acc_steps
    else:
        gate_err, gate_steps = extra.get("rel_l2_u_vs_fp64"), K + W
    extra["accuracy_gate"] = {"rel_l2_u_vs_fp64": gate_err, "steps": gate_steps, "bar": 1e-5,
                              "met": None if gate_err is None else bool(gate_err <= 1e-5)}

    cpu = None
    if not args.no_cpu:
        log("cpu baseline (oracle, OpenMP)")
        cpu = cpu_baseline()

    line = {
        "metric": METRIC, "value": round(mlups32, 1), "unit": "MLUPS", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": round(dt32 / K * 1e3, 5), "higher_is_better": True, "scaling": "weak",
        "clock_warm_steps": warm_steps,
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "FP32 fused MRT+LES(Cs=0.1)+IB step with network tau at every node (6-32-32-1 "
                               "SiLU, bf16 MFMA, seeded Xavier weights, guard >= 0.501), cylinder D=160 (503 "
                               "markers), 4096x2048, inflow 0.08 +-1e-3 noise, Re=1e4, FP32 populations stored as f - feq(1, (0.08, 0))",
                   "nx": NX, "ny": NY, "collision": "mrt", "les_cs": CS, "tau0": TAU0, "tau_mode": "nn",
                   "u_in": U0, "bodies": 1, "markers": NMARK, "bc": "inflow"},
        **extra,
        "roofline": roofline, "cpu_baseline": cpu, "legs": legs,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
