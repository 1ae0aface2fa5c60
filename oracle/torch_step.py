"""FP64 PyTorch restatement of the two-phase step, differentiable by autograd.

TEST INFRASTRUCTURE ONLY: the reference gradient the hand-written HIP adjoint
is checked against (tests/, CPU, small grids).  Never imported by the product.

It restates the same step as oracle/lbm_oracle.c, vectorised over nodes:
  ini / ini_op            program/lbm_drop_xy.f90:224-251  (sharp-step droplet, f = feq, g = geq)
  various_cal, gphi_cal   :590-666  (lap phi, grad phi, Korteweg tensor)
  equilibrium_cal         :421-476  (feq, geq with Anu = 0)
  fg_cal                  :478-515  (collide at the source node, stream by pull)
  bounce_back_LBM         prototype_program/two-phase/dropch.f90:666-698 (walls normal
                          to y or z, -U on the bottom layer, +U on the top layer,
                          from step wall_start on)
  physics                 program/lbm_drop_xy.f90:541-553 (moments)
with equal viscosities (tau_g constant) and no body force or LES, which is
the scope of the adjoint.  Operation order differs from the reference, so
results agree with the C restatement to rounding (~1e-13), not bitwise; the
forward is pinned against the C restatement in tests/test_torch_reference.py.

Loss (drp_tape_loss): J = w0 D + w1 A, D the second-moment Taylor deformation of
the search plane (drp_deformation mode 2) and A = sum |grad phi|.
"""
from __future__ import annotations

import torch

CX = [0, 1, 0, 0, -1, 0, 0, 1, -1, 1, 1, -1, 1, -1, -1]
CY = [0, 0, 1, 0, 0, -1, 0, 1, 1, -1, 1, -1, -1, 1, -1]
CZ = [0, 0, 0, 1, 0, 0, -1, 1, 1, 1, -1, -1, -1, -1, 1]
OPP = [0, 4, 5, 6, 1, 2, 3, 11, 12, 13, 14, 7, 8, 9, 10]
E = [2 / 9] + [1 / 9] * 6 + [1 / 72] * 8
F = [-7 / 3] + [1 / 3] * 6 + [1 / 24] * 8
WALL_Y, WALL_Z = 2, 3


def _col(v):
    return torch.tensor(v, dtype=torch.float64).view(15, 1, 1, 1)


def _shift(a, dx, dy, dz):
    """a(x + c) for a [nz, ny, nx] field, periodic."""
    return torch.roll(a, shifts=(-dz, -dy, -dx), dims=(0, 1, 2))


def stencils(phi):
    """lap phi (:608-613) and grad phi (:620-630) with ds = 1."""
    lap = -14.0 * phi
    g = [torch.zeros_like(phi) for _ in range(3)]
    for q in range(1, 15):
        v = _shift(phi, CX[q], CY[q], CZ[q])
        lap = lap + v
        g[0] = g[0] + CX[q] * v
        g[1] = g[1] + CY[q] * v
        g[2] = g[2] + CZ[q] * v
    return lap / 5.0, [gi / 10.0 for gi in g]


def equilibria(phi, p, u, cfg, kappa_g):
    lap, g = stencils(phi)
    cx, cy, cz = _col(CX), _col(CY), _col(CZ)
    Ew, Fw = _col(E), _col(F)
    H = torch.zeros(15, 1, 1, 1, dtype=torch.float64)
    H[0] = 1.0
    gsq = g[0] * g[0] + g[1] * g[1] + g[2] * g[2]
    p0 = phi * cfg.T / (1.0 - cfg.b * phi) - cfg.a * phi * phi
    A = p0 - cfg.kappa_f * phi * lap - (cfg.kappa_f / 6.0) * gsq
    cu = cx * u[0] + cy * u[1] + cz * u[2]
    cg = cx * g[0] + cy * g[1] + cz * g[2]
    c2 = cx * cx + cy * cy + cz * cz
    gt = 4.5 * cg * cg - 1.5 * gsq * c2
    gt = torch.cat([torch.zeros_like(gt[:1]), gt[1:]])
    usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2]
    feq = H * phi + Fw * A + 3.0 * Ew * phi * cu + Ew * cfg.kappa_f * gt
    geq = Ew * (3.0 * p + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq) + Ew * kappa_g * gt
    return feq, geq


def moments(f, g):
    phi = f.sum(0)
    p = g.sum(0) / 3.0
    u = [sum(c * g[q] for q, c in enumerate(C) if c != 0) for C in (CX, CY, CZ)]
    return phi, p, u


def step(f, g, mac, cfg, tau_f, kappa_g, wall_u, nstep):
    """One step; mac = (phi, p, [u1, u2, u3]); nstep is the reference loop counter."""
    phi, p, u = mac
    feq, geq = equilibria(phi, p, u, cfg, kappa_g)
    tau_g = 0.5 + 3.0 * cfg.nu1
    fs = f - (f - feq) / tau_f
    gs = g - (g - geq) / tau_g
    fn = torch.stack([_shift(fs[q], -CX[q], -CY[q], -CZ[q]) for q in range(15)])
    gn = [_shift(gs[q], -CX[q], -CY[q], -CZ[q]) for q in range(15)]
    if cfg.wall_axis in (WALL_Y, WALL_Z):
        # bounce-back every step; the walls move from step wall_start on (:670-674)
        if nstep < cfg.wall_start:
            wall_u = 0.0 * wall_u
        axis = 1 if cfg.wall_axis == WALL_Y else 0          # dims of [nz, ny, nx]
        nax = phi.shape[axis]
        idx = torch.arange(nax).view([nax, 1, 1] if axis == 0 else [1, nax, 1])
        bottom = (idx == 0)
        top = (idx == nax - 1)
        ca = CY if cfg.wall_axis == WALL_Y else CZ
        old = list(gn)
        for q in range(1, 15):
            o = OPP[q]
            if ca[q] == 1:
                gn[q] = torch.where(bottom, old[o] - 6.0 * E[o] * (-wall_u * CX[o]), old[q])
            elif ca[q] == -1:
                gn[q] = torch.where(top, old[o] - 6.0 * E[o] * (wall_u * CX[o]), old[q])
    gn = torch.stack(gn)
    return fn, gn, moments(fn, gn)


def initial_state(cfg, drops, kappa_g, p_init=1.0 / 3.0):
    nz, ny, nx = cfg.nz, cfg.ny, cfg.nx
    z, y, x = torch.meshgrid(torch.arange(nz, dtype=torch.float64), torch.arange(ny, dtype=torch.float64),
                             torch.arange(nx, dtype=torch.float64), indexing="ij")
    phi = torch.full((nz, ny, nx), cfg.phi1, dtype=torch.float64)
    for (xc, yc, zc, D) in drops:
        dif = (x - xc) ** 2 + (y - yc) ** 2 + (z - zc) ** 2
        r = 0.5 * D
        phi = torch.where(dif <= r * r, torch.full_like(phi, cfg.phi2), phi)
    p = torch.full_like(phi, p_init)
    u = [torch.zeros_like(phi) for _ in range(3)]
    f, g = equilibria(phi, p, u, cfg, kappa_g)
    return f, g, (phi, p, u)


def deformation_inertia(phi, cfg):
    """Second-moment D of the search plane (drp_deformation mode 2)."""
    if cfg.wall_axis == WALL_Y:
        S = phi[cfg.nz // 2]                     # [y, x]: j = y
    else:
        S = phi[:, cfg.ny // 2, :]               # [z, x]: j = z
    n2, n1 = S.shape
    j, i = torch.meshgrid(torch.arange(n2, dtype=torch.float64), torch.arange(n1, dtype=torch.float64), indexing="ij")
    sn = (S - cfg.phi1) / (cfg.phi2 - cfg.phi1)
    w = sn * sn
    W = w.sum()
    xc = (w * i).sum() / W
    zc = (w * j).sum() / W
    dx, dz = i - xc, j - zc
    ixx, izz, ixz = (w * dx * dx).sum(), (w * dz * dz).sum(), (w * dx * dz).sum()
    m = 0.5 * (ixx + izz)
    r = torch.sqrt(0.25 * (ixx - izz) ** 2 + ixz * ixz)
    L = 4.0 * torch.sqrt((m + r) / W)
    B = 4.0 * torch.sqrt(torch.clamp(m - r, min=0.0) / W)
    return (L - B) / (L + B)


def interface_area(phi):
    _, g = stencils(phi)
    return torch.linalg.vector_norm(torch.stack(g), dim=0).sum()


def run_loss(cfg, drops, nsteps, tau_f, kappa_g, wall_u, w=(1.0, 1.0), p_init=1.0 / 3.0):
    """J(tau_f, kappa_g, wall_u) after nsteps from the droplet initial state; returns (J, D, A, phi)."""
    f, g, mac = initial_state(cfg, drops, kappa_g, p_init)
    for n in range(1, nsteps + 1):
        f, g, mac = step(f, g, mac, cfg, tau_f, kappa_g, wall_u, n)
    phi = mac[0]
    D = deformation_inertia(phi, cfg)
    A = interface_area(phi)
    return w[0] * D + w[1] * A, D, A, phi
