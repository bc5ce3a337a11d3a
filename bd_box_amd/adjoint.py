"""Differentiable Brownian-dynamics trajectories: K Ermak-McCammon steps with given
Gaussians as torch.autograd.Functions, backed by the device's reverse-mode step
(bdh_step_vjp, bdh_step_vjp_params).  The reference has no derivatives; the step
differentiated is its comp_step_ermak with Cholesky noise (src/diff_alg/diff_comp.c:152-179),
refactored at every step.

trajectory()        final positions, differentiable in the initial positions
param_trajectory()  the positions after every step, differentiable in the initial positions
                    and the force-field parameters: per-bead LJ epsilon and radius
                    (LJ.c:37-64, data.c:97-102), charges (electro.c:36-56) and per-bond force
                    constants (bond.c:43-85)
rg_loss()           (mean squared radius of gyration over the frames - target)^2

The forward pass keeps the positions of every `checkpoint`-th step only; the backward pass
re-runs each segment from its checkpoint, saving the factor of each step on the device, and
walks it back step by step with those factors.

FP64 systems differentiate the FP64 step.  Mixed-precision systems (FP32 tensor, forces and
trailing updates) differentiate the step they take: its FP32 factor, promoted to FP64, enters
the reverse step, which runs in FP64; the factorizations of the forward pass and the
re-runs, the bulk of the time, run in mixed precision.
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


def check_differentiable(cfg) -> None:
    """Raise ValueError unless the configuration's step is the one differentiated: Cholesky
    noise refactored at every step, free space, ermak_const, no overlap check (any
    precision).  With check_overlap a rejected step is redrawn from the device stream (and
    ermak_var halves dt), so the given-Gaussian step would not be the step taken."""
    if cfg.refactor_every != 1:
        raise ValueError("trajectory: the system must have refactor_every=1 (the derivative "
                         "is that of a step refactored at its own positions)")
    if cfg.hydro != 1 or cfg.bc != 0:
        raise ValueError("trajectory: free-space Cholesky noise only (hydro cholesky, bc none)")
    if cfg.check_overlap or cfg.algorithm != 0:
        raise ValueError("trajectory: ermak_const without check_overlap only (a rejected step "
                         "is redrawn from the stream, which the given-Gaussian step does not "
                         "follow)")


def trajectory(system: BDSystem, xyz0: torch.Tensor, xi, radii, checkpoint: int = 10) -> torch.Tensor:
    """Positions (N, 3) after len(xi) steps from xyz0 (N, 3, float64) with Gaussians
    xi (K, 3N); differentiable with respect to xyz0.  `system` must be a BDSystem created
    with refactor_every=1 (its state is overwritten)."""
    check_differentiable(system.cfg)
    if checkpoint < 1:
        raise ValueError("trajectory: checkpoint must be positive")
    if not torch.is_tensor(xi):
        xi = torch.as_tensor(np.asarray(xi, dtype=np.float64))
    return Trajectory.apply(xyz0, xi, system, radii, checkpoint)


class ParamTrajectory(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xyz0, eps, R, Q, H, xi, system: BDSystem, radii, bond_r0_rmax, checkpoint: int):
        xi_np = np.ascontiguousarray(xi.detach().cpu().numpy(), dtype=np.float64)
        K = xi_np.shape[0]
        LJ = np.stack([R.detach().cpu().numpy(), eps.detach().cpu().numpy()], axis=1)
        bp = np.concatenate([np.asarray(bond_r0_rmax, float).reshape(-1, 2),
                             H.detach().cpu().numpy().reshape(-1, 1)], axis=1)
        system.set_params(Q=Q.detach().cpu().numpy(), LJ=LJ, bond_parms=bp)
        coord = np.concatenate([xyz0.detach().cpu().numpy().reshape(-1, 3),
                                np.asarray(radii, float).reshape(-1, 1)], axis=1)
        checkpoints = [coord.copy()]
        frames = np.zeros((K, system.N, 3))
        system.set_coords(coord)
        for k in range(K):
            system.compute_hydro(1, xi_np[k])
            c = system.coords()
            frames[k] = c[:, :3]
            if (k + 1) % checkpoint == 0 and k + 1 < K:
                checkpoints.append(c)
        ctx.system = system
        ctx.xi = xi_np
        ctx.checkpoints = checkpoints
        ctx.every = checkpoint
        return torch.from_numpy(frames)

    @staticmethod
    def backward(ctx, g_out):
        system, xi, every = ctx.system, ctx.xi, ctx.every
        K = xi.shape[0]
        G = g_out.detach().cpu().numpy().astype(np.float64).reshape(K, -1)
        bars = [np.zeros(system.N), np.zeros(system.N), np.zeros(system.N), np.zeros(system.nbonds)]
        g = G[K - 1].copy()
        system.factor_cache(every)
        for seg in range(len(ctx.checkpoints) - 1, -1, -1):
            t0 = seg * every
            t1 = min(K, t0 + every)
            starts = [ctx.checkpoints[seg]]
            system.set_coords(starts[0])
            for t in range(t0, t1):
                system.compute_hydro(1, xi[t])
                system.save_factor(t - t0)
                if t + 1 < t1:
                    starts.append(system.coords())
            for t in range(t1 - 1, t0 - 1, -1):
                g, eb, Rb, Qb, Hb = system.step_vjp_params(starts[t - t0], xi[t], g, slot=t - t0)
                for acc, part in zip(bars, (eb, Rb, Qb, Hb)):
                    acc += part
                if t > 0:
                    g = g + G[t - 1]
        system.factor_cache(0)
        t = torch.from_numpy
        return (t(g.reshape(-1, 3)), t(bars[0]), t(bars[1]), t(bars[2]), t(bars[3]),
                None, None, None, None, None)


def param_trajectory(system: BDSystem, xyz0: torch.Tensor, eps: torch.Tensor, R: torch.Tensor,
                     Q: torch.Tensor, H: torch.Tensor, xi, radii, bond_r0_rmax,
                     checkpoint: int = 10) -> torch.Tensor:
    """Positions (K, N, 3) after each of the K = len(xi) steps from xyz0 (N, 3) with
    Gaussians xi (K, 3N), under per-bead LJ epsilon eps (N), LJ radius R (N; sigma_ij =
    R_i + R_j), charges Q (N) and per-bond force constants H (nbonds; bond_r0_rmax holds the
    bonds' (r0, r_max)); differentiable with respect to xyz0, eps, R, Q and H (float64
    tensors).  `system` must be a BDSystem created with refactor_every=1 over the same beads
    and bonds; its parameters and state are overwritten."""
    check_differentiable(system.cfg)
    if checkpoint < 1:
        raise ValueError("param_trajectory: checkpoint must be positive")
    if not torch.is_tensor(xi):
        xi = torch.as_tensor(np.asarray(xi, dtype=np.float64))
    return ParamTrajectory.apply(xyz0, eps, R, Q, H, xi, system, radii, bond_r0_rmax, checkpoint)


def rg_loss(frames: torch.Tensor, target: float = 400.0) -> torch.Tensor:
    """(mean over the K frames of the squared radius of gyration - target)^2, frames (K, N, 3)."""
    c = frames - frames.mean(dim=1, keepdim=True)
    rg2 = (c * c).sum(dim=(1, 2)) / frames.shape[1]
    return (rg2.mean() - target) ** 2
