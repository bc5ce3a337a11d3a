"""Differentiable Brownian-dynamics trajectory: K FP64 Ermak-McCammon steps with given
Gaussians as a torch.autograd.Function, backed by the device's reverse-mode step
(bdh_step_vjp).  The reference has no derivatives; the step differentiated is its
comp_step_ermak with Cholesky noise (src/diff_alg/diff_comp.c:152-179), refactored at
every step.

The forward pass keeps the positions of every `checkpoint`-th step only; the backward pass
re-runs each segment from its checkpoint, saving the factor of each step on the device, and
walks it back step by step with those factors.
"""
from __future__ import annotations

import numpy as np
import torch

from .system import BDSystem


class Trajectory(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xyz0, xi, system: BDSystem, radii, checkpoint: int):
        xi_np = np.ascontiguousarray(xi.detach().cpu().numpy(), dtype=np.float64)
        K = xi_np.shape[0]
        coord = np.concatenate([xyz0.detach().cpu().numpy().reshape(-1, 3),
                                np.asarray(radii, float).reshape(-1, 1)], axis=1)
        checkpoints = [coord.copy()]
        system.set_coords(coord)
        for k in range(K):
            system.compute_hydro(1, xi_np[k])
            if (k + 1) % checkpoint == 0 and k + 1 < K:
                checkpoints.append(system.coords())
        out = system.coords()
        ctx.system = system
        ctx.xi = xi_np
        ctx.checkpoints = checkpoints
        ctx.every = checkpoint
        return torch.from_numpy(out[:, :3].copy())

    @staticmethod
    def backward(ctx, g_out):
        system, xi, every = ctx.system, ctx.xi, ctx.every
        K = xi.shape[0]
        g = g_out.detach().cpu().numpy().astype(np.float64).reshape(-1)
        system.factor_cache(every)
        for seg in range(len(ctx.checkpoints) - 1, -1, -1):
            t0 = seg * every
            t1 = min(K, t0 + every)
            # positions at the start of each step of the segment, and the step's factor
            starts = [ctx.checkpoints[seg]]
            system.set_coords(starts[0])
            for t in range(t0, t1):
                system.compute_hydro(1, xi[t])
                system.save_factor(t - t0)
                if t + 1 < t1:
                    starts.append(system.coords())
            for t in range(t1 - 1, t0 - 1, -1):
                g = system.step_vjp(starts[t - t0], xi[t], g, slot=t - t0)
        system.factor_cache(0)
        return torch.from_numpy(g.reshape(-1, 3)), None, None, None, None


def trajectory(system: BDSystem, xyz0: torch.Tensor, xi, radii, checkpoint: int = 10) -> torch.Tensor:
    """Positions (N, 3) after len(xi) steps from xyz0 (N, 3, float64) with Gaussians
    xi (K, 3N); differentiable with respect to xyz0.  `system` must be an FP64 BDSystem
    created with refactor_every=1 (its state is overwritten)."""
    if system.cfg.precision != 0 or system.cfg.refactor_every != 1:
        raise ValueError("trajectory: the system must be FP64 with refactor_every=1 (the "
                         "derivative is that of a step refactored at its own positions)")
    if checkpoint < 1:
        raise ValueError("trajectory: checkpoint must be positive")
    if not torch.is_tensor(xi):
        xi = torch.as_tensor(np.asarray(xi, dtype=np.float64))
    return Trajectory.apply(xyz0, xi, system, radii, checkpoint)
