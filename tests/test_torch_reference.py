"""Pins of the PyTorch FP64 restatement (oracle/torch_step.py), the reference
gradient for the hand-written adjoint.

Its forward is checked against the C restatement (itself bit-exact with the
compiled reference on the golden fixtures) to rounding, its loss pieces against
the C restatement's deformation mode 2 and interface area, and its autograd
gradient against central finite differences of its own forward.  CPU only.
"""
import numpy as np
import pytest
import torch

from droplet_amd.abi import DRP_WALL_Y, DRP_WALL_Z, POP_MOM, POP_PHI, make_config
from oracle import torch_step as ts
from oracle.pyoracle import OracleLattice


def _case(N=12, axis=DRP_WALL_Y, U=0.02, kg=0.01):
    cfg = make_config(N, N, N, nu1=0.1, nu2=0.1, kappa_g=kg, wall_axis=axis, wall_u=U, wall_start=2)
    c = 0.5 * (N - 1)
    return cfg, [(c, c + 0.3, c - 0.2, 0.55 * N)]


def _torch_run(cfg, drops, nsteps):
    f, g, mac = ts.initial_state(cfg, drops, cfg.kappa_g)
    for n in range(1, nsteps + 1):
        f, g, mac = ts.step(f, g, mac, cfg, cfg.tau_f, cfg.kappa_g, cfg.wall_u, n)
    return f, g, mac


@pytest.mark.parametrize("axis", [DRP_WALL_Y, DRP_WALL_Z])
def test_forward_matches_c_restatement(axis):
    cfg, drops = _case(axis=axis)
    o = OracleLattice(cfg, threads=2)
    o.init_droplets(drops)
    o.step(15)
    f, g, mac = _torch_run(cfg, drops, 15)
    phi_o = o.macro(0)
    assert np.abs(mac[0].numpy() - phi_o).max() / np.abs(phi_o).max() < 1e-11
    for k in range(3):
        assert np.abs(mac[2][k].numpy() - o.macro(2 + k)).max() < 1e-13
    assert np.abs(f.numpy() - o.populations(POP_PHI)).max() < 1e-11
    assert np.abs(g.numpy() - o.populations(POP_MOM)).max() < 1e-13
    # a planted sign error in the wall term is visible at this tolerance
    _, g_bad, _ = _torch_run(make_config(cfg.nx, **{k: getattr(cfg, k) for k in ("nu1", "nu2", "kappa_g",
                                                                                    "wall_axis", "wall_start")},
                                         wall_u=-cfg.wall_u), drops, 15)
    assert np.abs(g_bad.numpy() - o.populations(POP_MOM)).max() > 1e-6


def test_loss_pieces_match_c_restatement():
    cfg, drops = _case(N=16)
    o = OracleLattice(cfg, threads=2)
    o.init_droplets(drops)
    o.step(30)
    phi = torch.from_numpy(o.macro(0))
    _, _, D_o = o.deformation(2)
    assert abs(ts.deformation_inertia(phi, cfg).item() - D_o) < 1e-13
    A_o = o.reduce()["interface_area"]
    assert abs(ts.interface_area(phi).item() - A_o) / A_o < 1e-13


def test_deformation_of_an_ellipse():
    """Uniform elliptic disc with semi-axes a > b: D -> (a - b)/(a + b) as the grid refines."""
    N = 200
    cfg = make_config(N, N, 3, wall_axis=DRP_WALL_Y)
    z, y, x = torch.meshgrid(torch.arange(3.0, dtype=torch.float64), torch.arange(N, dtype=torch.float64),
                             torch.arange(N, dtype=torch.float64), indexing="ij")
    a, b, c = 60.0, 40.0, 99.5
    inside = ((x - c) / a) ** 2 + ((y - c) / b) ** 2 <= 1.0
    phi = torch.where(inside, torch.full_like(x, cfg.phi2), torch.full_like(x, cfg.phi1))
    D = ts.deformation_inertia(phi, cfg).item()
    assert abs(D - (a - b) / (a + b)) < 2e-3


@pytest.mark.parametrize("axis", [DRP_WALL_Y, DRP_WALL_Z])
def test_autograd_matches_central_differences(axis):
    cfg, drops = _case(N=12, axis=axis)
    nsteps = 12
    theta = torch.tensor([cfg.tau_f, cfg.kappa_g, cfg.wall_u], dtype=torch.float64, requires_grad=True)
    for w, tol in (((1.0, 1.0), (1e-6, 1e-6, None)), ((1.0, 0.0), (1e-5, 1e-5, 1e-5))):
        theta.grad = None
        J, _, _, _ = ts.run_loss(cfg, drops, nsteps, theta[0], theta[1], theta[2], w=w)
        J.backward()
        for k in range(3):
            if tol[k] is None:   # dJ/dU of the area-dominated loss is below the FD noise floor here
                continue
            h = 1e-6 * abs(theta[k].item())
            vals = []
            for s in (1.0, -1.0):
                tt = theta.detach().clone()
                tt[k] += s * h
                with torch.no_grad():
                    vals.append(ts.run_loss(cfg, drops, nsteps, tt[0], tt[1], tt[2], w=w)[0].item())
            fd = (vals[0] - vals[1]) / (2 * h)
            rel = abs(fd - theta.grad[k].item()) / abs(fd)
            assert rel < tol[k], f"w={w} param {k}: autograd {theta.grad[k].item()} vs FD {fd} (rel {rel:.2e})"
