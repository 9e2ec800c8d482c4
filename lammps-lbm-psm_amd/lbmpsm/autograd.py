"""torch.autograd.Function over the lbmpsm adjoint.

`HydroRun.apply(F_lbm, v, omega, tau, spec)` runs `spec.steps` FP64 LBM
steps on the GPU with the particles of `spec` frozen (same list every step)
and returns the last step's per-entry force [n,3], torque [n,3] and
stresslet [n,6].  Its backward is the library's discrete adjoint
(checkpoints every `spec.checkpoint_every` steps, recomputation in between),
giving gradients with respect to the lattice body force F_lbm [3], the entry
velocities v [n,3] and angular velocities omega [n,3] (units of the list)
and tau (unit conversion factors held fixed).

Tensors are CPU float64; the lattice itself lives on the GPU for the
duration of one forward/backward pair.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from .lattice import Lattice, lattice_units_config


@dataclass
class RunSpec:
    dim: int
    n: tuple            # nodes per axis (fully periodic)
    tag: np.ndarray     # [m] int32, list order
    x: np.ndarray       # [m,3] positions (lattice units)
    radius: np.ndarray  # [m]
    steps: int
    checkpoint_every: int
    device: int = 0


class HydroRun(torch.autograd.Function):
    @staticmethod
    def forward(ctx, F_lbm, v, omega, tau, spec: RunSpec):
        cfg = lattice_units_config(spec.dim, spec.n, (1, 1, 1), float(tau),
                                   F_lbm=tuple(float(a) for a in F_lbm.tolist()),
                                   device=spec.device)
        lat = Lattice(cfg)
        try:
            F, T, S = lat.adjoint_forward(spec.steps, spec.checkpoint_every, spec.tag, spec.x,
                                          v.detach().numpy(), omega.detach().numpy(),
                                          spec.radius)
        except Exception:
            lat.close()
            raise
        ctx.lat = lat
        return torch.from_numpy(F), torch.from_numpy(T), torch.from_numpy(S)

    @staticmethod
    def backward(ctx, gF, gT, gS):
        lat = ctx.lat
        try:
            dF, dtau, dv, dw = lat.adjoint_backward(gF.numpy(), gT.numpy(), gS.numpy())
        finally:
            lat.close()
            ctx.lat = None
        return (torch.from_numpy(dF), torch.from_numpy(dv), torch.from_numpy(dw),
                torch.tensor(dtau, dtype=torch.float64), None)


def run_losses(F_lbm, v, omega, tau, spec: RunSpec):
    """Forward only (no graph): the same run as HydroRun, for finite differences."""
    cfg = lattice_units_config(spec.dim, spec.n, (1, 1, 1), float(tau),
                               F_lbm=tuple(float(a) for a in F_lbm), device=spec.device)
    with Lattice(cfg) as lat:
        return lat.adjoint_forward(spec.steps, spec.checkpoint_every, spec.tag, spec.x, v, omega,
                                   spec.radius)
