"""Trajectory-loss gradients as a torch.autograd.Function over the C ABI.

The loss of one engine call (lima_adjoint_run) over the frames x_0..x_{K-1}
that start at the engine's current state is

    L_r = sum_k alpha_r (E_k,r - E*_k,r)^2 + beta_r sum_i |F_k,i - F*_k,i|^2

per system r, with E the potential energy and F the per-particle force of the
LIMA_MD step (Engine.cu:400-451, integrated with the plain Verlet update
Engine.cu:849-869; the thermostat must be off).  Its gradient with respect to
the parameter vector of each system

    [sigma_0..31, epsilon_0..31, kb, b0, k_theta, theta_0]

(ForceField, Forcefield.cuh:12-27; PairBond/AngleBond, Bodies.cuh:61-88, the
bonded constants shared by every term of the system) comes from the
hand-written adjoint on the GPU in the same call.  The autograd node therefore
evaluates value and gradient together in `forward` and only scales the stored
gradient in `backward`.
"""
from __future__ import annotations

import numpy as np
import torch

from .abi import N_PARAMETERS
from .engine import Engine


class TrajectoryLoss(torch.autograd.Function):
    """loss[R] = L(params[R, 68]) for the engine's systems from its current state."""

    @staticmethod
    def forward(ctx, params, engine: Engine, n_frames: int, energy_weight, force_weight, energy_target,
                chain_force_target, solvent_force_target, checkpoint_interval: int):
        R = len(engine.simulation.systems)
        p = params.detach().to("cpu", torch.float64).reshape(R, N_PARAMETERS).numpy()
        for r in range(R):
            engine.set_parameters(p[r], r)
        need = bool(params.requires_grad)
        out = engine.adjoint(n_frames, energy_weight, force_weight, energy_target, chain_force_target,
                             solvent_force_target, checkpoint_interval=checkpoint_interval, gradient=need)
        loss, grad = out if need else (out, np.zeros((R, N_PARAMETERS)))
        ctx.save_for_backward(torch.from_numpy(grad).to(params.device, params.dtype).reshape(params.shape))
        return torch.from_numpy(loss).to(params.device, params.dtype)

    @staticmethod
    def backward(ctx, grad_loss):
        (grad,) = ctx.saved_tensors
        g = grad_loss.reshape(-1, *([1] * (grad.dim() - 1))) if grad.dim() > 1 else grad_loss
        return g * grad, None, None, None, None, None, None, None, None


def trajectory_loss(engine: Engine, params: torch.Tensor, n_frames: int, energy_weight=None, force_weight=None,
                    energy_target=None, chain_force_target=None, solvent_force_target=None,
                    checkpoint_interval: int = 0) -> torch.Tensor:
    """Differentiable loss per system; `params` is [68] (one system) or [R, 68]."""
    R = len(engine.simulation.systems)
    if params.numel() != R * N_PARAMETERS:
        raise ValueError(f"params must hold {R} x {N_PARAMETERS} values")
    loss = TrajectoryLoss.apply(params, engine, n_frames, energy_weight, force_weight, energy_target,
                                chain_force_target, solvent_force_target, checkpoint_interval)
    return loss[0] if params.dim() == 1 else loss
