"""Parameter sensitivities of a shear-droplet run as a torch.autograd.Function.

    theta = torch.tensor([M, kappa_g, U], dtype=torch.float64, requires_grad=True)
    J = droplet_loss(theta, cfg, drops, nsteps=400, checkpoint_every=20)
    J.backward()            # theta.grad = dJ/d(M, kappa_g, U)

M is the phi-population mobility of the Inamuro model in lattice units,
M = tau_f - 1/2 (the order-parameter equation is dphi/dt + div(phi u) =
M lap A to leading order, A the F-bracket of feq), kappa_g sets the surface
tension (sigma = 1.7039 kappa_g^0.9991, program/lbm_drop_xy.f90:51) and U is
the wall speed (walls at -U / +U, shear rate 2U/H,
prototype_program/two-phase/dropch.f90:666-698).

The forward pass runs the FP64 HIP step through a drp_tape (checkpoint every
K steps) and returns J = w0 D + w1 A, D the second-moment Taylor deformation
and A = sum |grad phi|; the backward pass is the hand-written discrete adjoint
(droplet_amd/csrc/adjoint.hip) with each K-step segment recomputed from its
checkpoint.  There is no CPU path: without the HIP library this raises.
"""
from __future__ import annotations

import torch

from .abi import Config
from .lattice import Lattice, Tape

PARAMS = ("mobility", "kappa_g", "wall_u")


def config_with(cfg: Config, theta) -> Config:
    """Copy of cfg with (M, kappa_g, U) = theta applied (tau_f = M + 1/2)."""
    M, kg, U = (float(v) for v in theta)
    c = Config.from_buffer_copy(cfg)
    c.tau_f = M + 0.5
    c.kappa_g = kg
    c.wall_u = U
    return c


class DropletLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, cfg, drops, nsteps, checkpoint_every, w):
        lat = Lattice(config_with(cfg, theta.detach().cpu().tolist()))
        lat.init_droplets(drops)
        tape = Tape(lat, checkpoint_every)
        tape.step(nsteps)
        J, _, _ = tape.loss(w)
        ctx.tape = tape
        ctx.w = w
        return theta.new_tensor(J)

    @staticmethod
    def backward(ctx, gJ):
        _, g = ctx.tape.backward(ctx.w)
        lat = ctx.tape.lattice
        ctx.tape.close()
        lat.close()
        ctx.tape = None
        grad = gJ * torch.tensor(g, dtype=gJ.dtype, device=gJ.device)
        return grad, None, None, None, None, None


def droplet_loss(theta: torch.Tensor, cfg: Config, drops, nsteps: int, checkpoint_every: int = 20,
                 w=(1.0, 1.0)) -> torch.Tensor:
    """J(theta) of an nsteps run from the droplet initial state (differentiable in theta)."""
    if theta.shape != (3,):
        raise ValueError("theta must hold (mobility, kappa_g, wall_u)")
    return DropletLoss.apply(theta, cfg, [tuple(d) for d in drops], int(nsteps), int(checkpoint_every),
                             tuple(float(v) for v in w))
