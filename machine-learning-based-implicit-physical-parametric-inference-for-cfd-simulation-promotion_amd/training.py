"""Training the relaxation-time network on pairs the solver generates at
twice the resolution (host-side workflow over the C ABI).

The reference has no counterpart: its networks are trained offline and
loaded as TorchScript files (mesh.cpp:2767-3260), and its high-fidelity
relaxation times come from FindAndWriteTau's one-step search
(mesh.cpp:2625-2715), which `Solver.coarse_pairs(criterion="step")` keeps.

The workflow here:

1. `collect_pairs`: stress-matching pairs (`Solver.coarse_pairs(criterion=
   "stress")`, the effective relaxation time of each coarse node from the
   restricted fine momentum flux) at several instants of the fine run,
   subsampled where the coarse strain is not negligible;
2. `eddy_labels`: the labels with a molecular floor (tau >= tau_c: no
   negative eddy viscosity, which makes the coarse run locally inviscid)
   and the eddy part scaled by alpha:
   tau = tau_c + alpha * max(tau_label - tau_c, 0);
3. `train`: Adam on the GPU trainer (ippi_trainer_*), W1 / W2 rounded to
   bf16 afterwards (the FP32 step then multiplies them exactly);
4. `calibrate`: alpha picked a posteriori, by the velocity error of short
   coarse runs against the restricted fine solution on a validation window
   (the coarse scheme's own numerical error is not in the a-priori labels:
   an under-resolved coarse run needs more dissipation than the filtered
   fine stress carries).
"""
import numpy as np

from . import workloads as wl


def collect_pairs(fine, tau_coarse, snaps, every, per_snap=None, min_weight=1e-3, seed=0):
    """Stress-matching pairs at `snaps` instants `every` fine steps apart:
    (features [n, 6], omega [n], weights [n] capped at 10x their median).
    Per instant, nodes with weight above `min_weight` of the largest, at
    most `per_snap` of them (uniform random subset)."""
    rng = np.random.default_rng(seed)
    X, Y, W = [], [], []
    for i in range(snaps):
        f, om, w = fine.coarse_pairs(tau_coarse, criterion="stress")
        keep = np.nonzero(w > min_weight * w.max())[0] if w.max() > 0 else np.zeros(0, dtype=np.int64)
        if per_snap is not None and len(keep) > per_snap:
            keep = rng.choice(keep, per_snap, replace=False)
        X.append(f[keep])
        Y.append(om[keep])
        W.append(w[keep])
        if i + 1 < snaps:
            fine.step(every)
    X, Y, W = np.concatenate(X), np.concatenate(Y), np.concatenate(W)
    if len(Y) == 0:
        raise ValueError("no usable pairs (no strain above the weight threshold)")
    return X, Y, np.minimum(W / np.median(W), 10.0)


def eddy_labels(omega, tau_coarse, alpha):
    """omega labels -> 1 / (tau_c + alpha max(1/omega - tau_c, 0))."""
    tau = 1.0 / np.asarray(omega, dtype=np.float64)
    return 1.0 / (tau_coarse + alpha * np.maximum(tau - tau_coarse, 0.0))


def train(ippi, X, omega, w, seed=3, iters=3000, batch=8192, lr=1e-3, bf16=True):
    """GPU training from workloads.init_blob(X, omega, seed): (blob, loss
    at the start, loss trained, loss with bf16-rounded W1 / W2)."""
    tr = ippi.Trainer(len(omega))
    try:
        tr.set_data(X, omega, w)
        tr.set_params(wl.init_blob(X, omega, seed))
        l0, _ = tr.loss_grad()
        tr.adam(iters, batch, lr, seed=seed)
        l1, _ = tr.loss_grad()
        blob = tr.params()
        l2 = l1
        if bf16:
            blob = wl.bf16_weights(blob)
            tr.set_params(blob)
            l2, _ = tr.loss_grad()
    finally:
        tr.close()
    return blob, l0, l1, l2


def velocity_error(fields, target):
    """Relative L2 of (ux, uy) against target (ux, uy)."""
    _, a, b = fields
    xf, yf = target
    return float(np.sqrt(np.sum((a - xf) ** 2 + (b - yf) ** 2) / np.sum(xf ** 2 + yf ** 2)))


def calibrate(ippi, X, omega, w, tau_coarse, evaluate, alphas=(1, 2, 4, 8, 16), **train_kw):
    """Train for every alpha, keep the network whose evaluate(blob) (an
    a-posteriori error on a validation window; inf for a diverged run) is
    lowest: (alpha, blob, {alpha: error})."""
    best, errs = None, {}
    for a in alphas:
        blob, _, _, _ = train(ippi, X, eddy_labels(omega, tau_coarse, a), w, **train_kw)
        e = evaluate(blob)
        errs[a] = e
        if best is None or e < best[1]:
            best = (a, e, blob)
    return best[0], best[2], errs
