"""Gradient of time-averaged drag behind torch.autograd.

`DragCase` fixes everything of a BGK + immersed-boundary run except the
per-node relaxation-time field and the body centre; `drag(tau, center,
case)` returns the mean drag as a differentiable torch scalar. The forward
pass runs the HIP solver; backward returns the gradients that the library's
discrete adjoint (ippi_adjoint_drag: checkpoint every `ckpt` steps,
recomputation per segment) computed along with it. There is no counterpart
in the reference (IPPI-LBM has no gradient code): the differentiated step
is the one of mesh.cpp:357-513 with the IB wiring of mesh.cpp:514-741,
1244-1281 and the Guo source of NodeAndLattice.cpp:162-195.
"""
from dataclasses import dataclass

import numpy as np
import torch

from . import Solver


@dataclass(frozen=True)
class DragCase:
    nx: int
    ny: int
    diameter: float
    markers: int
    nsteps: int
    avg_from: int
    ckpt: int = 100
    precision: str = "fp64"
    fp64_adjoint: bool = False  # FP64 adjoint over an FP32 forward
    seed: int = 1
    u0: tuple = (0.05, 0.0)
    noise: float = 1e-3
    device: int = 0

    def solver(self, tau, center):
        """A fresh solver in the case's initial state."""
        s = Solver(self.nx, self.ny, "bgk", self.precision, "periodic_bb", tau_mode="field",
                   tau0=float(np.mean(tau)), u_in=(self.u0[0], 0.0), device=self.device,
                   force_history=max(self.nsteps, 1))  # u_in: the FP32 storage's reference velocity
        s.set_tau_field(tau)
        s.init_fields(self.seed, 1.0, self.u0, self.noise)
        s.add_body(float(center[0]), float(center[1]), self.diameter, self.markers)
        return s

    def objective(self, tau, center):
        """Mean drag from a forward run only (no adjoint)."""
        with self.solver(tau, center) as s:
            s.step(self.nsteps)
            rec = s.body_forces()
        return float(np.mean(rec[self.avg_from - 1:self.nsteps, 0, 0]))

    def gradient(self, tau, center):
        """(J, dJ/dtau [ny, nx], dJ/dcenter [2]) from the discrete adjoint."""
        with self.solver(tau, center) as s:
            J, gt, gc = s.adjoint_drag(self.nsteps, self.avg_from, self.ckpt, fp64_adjoint=self.fp64_adjoint)
        return J, gt, gc[0]


class _Drag(torch.autograd.Function):
    @staticmethod
    def forward(ctx, tau, center, case):
        J, gt, gc = case.gradient(tau.detach().cpu().double().numpy(), center.detach().cpu().double().numpy())
        ctx.save_for_backward(torch.from_numpy(gt).to(tau.dtype), torch.from_numpy(gc).to(center.dtype))
        ctx.devices = (tau.device, center.device)
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gJ):
        gt, gc = ctx.saved_tensors
        g = gJ.to(torch.float64).item()
        return (gt * g).to(ctx.devices[0]), (gc * g).to(ctx.devices[1]), None


def drag(tau, center, case):
    """Mean drag of `case` as a differentiable function of tau [ny, nx] and center [2]."""
    return _Drag.apply(tau, center, case)
