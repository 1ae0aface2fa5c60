"""Hand-written adjoint (drp_tape_*) against the FP64 autograd reference and
central finite differences.

- small grids: gradient of J = D + A with respect to (M, kappa_g, U) against
  autograd through oracle/torch_step.py (FP64, CPU), both wall orientations,
  several checkpoint intervals (recomputation is exact, so the gradient must not
  depend on K at all), and a tape started mid-run;
- the 96^3 shear droplet (radius 16, 400 steps, checkpoint every 20) through the
  torch.autograd.Function, against central differences with h = 1e-6 |theta|
  (for the wall speed on the deformation term; see the test's docstring).
"""
import numpy as np
import pytest
import torch

from droplet_amd.abi import DRP_FP32_SHIFTED, DRP_WALL_Y, DRP_WALL_Z, make_config, sigma_from_kappa_g
from oracle import torch_step as ts
from tests.recording import record

pytestmark = pytest.mark.gpu


def _dl():
    from droplet_amd import lattice as dl
    return dl


def _case(N, axis, U=0.02, kg=0.01, wall_start=2):
    cfg = make_config(N, N, N, nu1=0.1, nu2=0.1, kappa_g=kg, wall_axis=axis, wall_u=U, wall_start=wall_start)
    c = 0.5 * (N - 1)
    return cfg, [(c, c + 0.3, c - 0.2, 0.55 * N)]


def _torch_grad(cfg, drops, nsteps, w, skip=0):
    """Autograd gradient of J after nsteps; with skip > 0 the first skip steps are
    taken as constants (the tape starts there)."""
    theta = torch.tensor([cfg.tau_f, cfg.kappa_g, cfg.wall_u], dtype=torch.float64, requires_grad=True)
    f, g, mac = ts.initial_state(cfg, drops, cfg.kappa_g if skip else theta[1])
    for n in range(1, nsteps + 1):
        if skip and n == skip + 1:
            f, g = f.detach(), g.detach()
            mac = (mac[0].detach(), mac[1].detach(), [u.detach() for u in mac[2]])
        tf, kg, U = (cfg.tau_f, cfg.kappa_g, cfg.wall_u) if n <= skip else (theta[0], theta[1], theta[2])
        f, g, mac = ts.step(f, g, mac, cfg, tf, kg, U, n)
    J = w[0] * ts.deformation_inertia(mac[0], cfg) + w[1] * ts.interface_area(mac[0])
    J.backward()
    return J.item(), theta.grad.numpy().copy()


def _gpu_grad(cfg, drops, nsteps, K, w, skip=0):
    dl = _dl()
    lat = dl.Lattice(cfg)
    lat.init_droplets(drops)
    if skip:
        lat.step(skip)
    tape = dl.Tape(lat, K)
    tape.step(nsteps - skip)
    J, g = tape.backward(w)
    tape.close()
    lat.close()
    return J, np.array(g)


@pytest.mark.parametrize("axis", [DRP_WALL_Y, DRP_WALL_Z])
@pytest.mark.parametrize("w", [(1.0, 1.0), (1.0, 0.0)], ids=["D+A", "D"])
def test_adjoint_matches_autograd_reference(axis, w):
    cfg, drops = _case(16, axis)
    nsteps = 20
    J_ref, g_ref = _torch_grad(cfg, drops, nsteps, w)
    J, g = _gpu_grad(cfg, drops, nsteps, 6, w)   # 6 does not divide 20: a short last segment
    print(f"axis {axis} w {w}: J {J:.12e} (ref {J_ref:.12e}) grad {g} (ref {g_ref})")
    assert abs(J - J_ref) <= 1e-11 * max(abs(J_ref), 1.0)
    rel = np.abs(g - g_ref) / np.abs(g_ref)
    assert (rel < 1e-9).all(), rel


def test_gradient_independent_of_checkpoint_interval():
    cfg, drops = _case(12, DRP_WALL_Z)
    nsteps, w = 15, (1.0, 0.5)
    grads = [_gpu_grad(cfg, drops, nsteps, K, w)[1] for K in (1, 4, 15, 40)]
    for g in grads[1:]:
        assert np.array_equal(g, grads[0]), (g, grads[0])
    _, g_ref = _torch_grad(cfg, drops, nsteps, w)
    assert (np.abs(grads[0] - g_ref) / np.abs(g_ref) < 1e-9).all()


def test_tape_started_mid_run():
    """No initial-equilibrium term: the populations at the tape start are constants."""
    cfg, drops = _case(12, DRP_WALL_Y, wall_start=1)
    J_ref, g_ref = _torch_grad(cfg, drops, 14, (1.0, 1.0), skip=5)
    J, g = _gpu_grad(cfg, drops, 14, 3, (1.0, 1.0), skip=5)
    assert abs(J - J_ref) / abs(J_ref) < 1e-12
    assert (np.abs(g - g_ref) / np.abs(g_ref) < 1e-9).all(), (g, g_ref)


def test_tape_rejects_unsupported_runs():
    dl = _dl()
    cfg, drops = _case(12, DRP_WALL_Y)
    for bad in (dict(precision=DRP_FP32_SHIFTED), dict(nu2=0.2)):
        c = make_config(12, **{k: getattr(cfg, k) for k in ("nu1", "nu2", "kappa_g", "wall_axis", "wall_u")})
        for k, v in bad.items():
            setattr(c, k, v)
        lat = dl.Lattice(c)
        lat.init_droplets(drops)
        with pytest.raises(dl.DropletError):
            dl.Tape(lat, 4)
        lat.close()
    lat = dl.Lattice(cfg)
    lat.init_droplets(drops)
    tape = dl.Tape(lat, 4)
    tape.step(3)
    lat.step(1)                       # stepping outside the tape is detected
    with pytest.raises(dl.DropletError):
        tape.backward()
    tape.close()
    lat.close()


def c5_config():
    """96^3 shear droplet, radius 16, walls normal to y, Ca = 0.2, nu = 0.1, kappa_g = 0.002."""
    N, kg, nu, D, Ca = 96, 0.002, 0.1, 32.0, 0.2
    U = Ca * (N - 1) * sigma_from_kappa_g(kg) / (nu * D)
    cfg = make_config(N, nu1=nu, nu2=nu, kappa_g=kg, wall_axis=DRP_WALL_Y, wall_u=U, wall_start=1)
    c = 0.5 * (N - 1)
    return cfg, [(c, c, c, D)]


def test_adjoint_c5_against_central_differences():
    """C5: J = D + A after 400 steps at 96^3, checkpoint every 20.

    dJ/dM and dJ/dkappa_g are checked with h = 1e-6 |theta|.  The wall speed is
    ill-conditioned for that step: after 400 steps the wall shear has diffused
    only sqrt(nu t) ~ 6 cells into the 95-cell gap, so dJ/dU ~ 0.02 against
    J ~ 8.6e3 (J / (U dJ/dU) ~ 2e7), and dD/dU ~ 4e-4 with D formed from the
    difference of two ~1e6 second moments.  Rounding noise of ~1e-11 in J (1e-14
    in D) over 2h = 4e-8 swamps the difference, so dJ/dU and dD/dU are checked
    against the Richardson extrapolation (4 d(h) - d(2h)) / 3 of central
    differences with h = 1e-2 |U| (truncation O(h^4), noise ~1e-11 / 4e-4); the
    h = 1e-6 and 1e-3 differences are printed for the record."""
    from droplet_amd.adjoint import config_with, droplet_loss
    dl = _dl()
    cfg, drops = c5_config()
    nsteps, K = 400, 20
    theta0 = [cfg.tau_f - 0.5, cfg.kappa_g, cfg.wall_u]
    theta = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    J = droplet_loss(theta, cfg, drops, nsteps, K, (1.0, 1.0))
    J.backward()
    g = theta.grad.numpy()

    lat = dl.Lattice(config_with(cfg, theta0))
    lat.init_droplets(drops)
    tape = dl.Tape(lat, K)
    tape.step(nsteps)
    J_D, gD = tape.backward((1.0, 0.0))
    J_A, gA = tape.backward((0.0, 1.0))
    tape.close()
    lat.close()
    # the two weightings add up to the full loss and gradient
    assert abs(J_D + J_A - J.item()) <= 1e-12 * abs(J.item())
    assert np.allclose(np.array(gD) + np.array(gA), g, rtol=1e-10, atol=0)

    def loss_at(th, w):
        lat = dl.Lattice(config_with(cfg, th))
        lat.init_droplets(drops)
        lat.step(nsteps)
        tape = dl.Tape(lat, nsteps)   # only to evaluate the same loss
        val = tape.loss(w)[0]
        tape.close()
        lat.close()
        return val

    def fd(k, rel_h, w):
        h = rel_h * abs(theta0[k])
        tp, tm = list(theta0), list(theta0)
        tp[k] += h
        tm[k] -= h
        return (loss_at(tp, w) - loss_at(tm, w)) / (2 * h)

    def richardson(k, rel_h, w):
        return (4.0 * fd(k, rel_h, w) - fd(k, 2 * rel_h, w)) / 3.0

    assert loss_at(theta0, (1.0, 1.0)) == J.item()
    print(f"J = {J.item():.10e}: D = {J_D:.6e}, A = {J_A:.6e}")
    checks = [(0, 1e-6, (1.0, 1.0), g[0], "central", True), (1, 1e-6, (1.0, 1.0), g[1], "central", True),
              (2, 1e-2, (1.0, 1.0), g[2], "richardson", True), (2, 1e-2, (1.0, 0.0), gD[2], "richardson", True),
              (2, 1e-3, (1.0, 1.0), g[2], "central", False), (2, 1e-6, (1.0, 1.0), g[2], "central", False),
              (2, 1e-6, (1.0, 0.0), gD[2], "central", False)]
    for k, rel_h, w, adj, kind, enforced in checks:
        ref = fd(k, rel_h, w) if kind == "central" else richardson(k, rel_h, w)
        rel = abs(adj - ref) / abs(ref)
        print(f"param {k} w={w} {kind} h={rel_h:g}|theta|: adjoint {adj:.10e}  FD {ref:.10e}  rel {rel:.2e}"
              f"{'' if enforced else '  (record only)'}")
        if enforced:
            assert rel < 1e-5


def test_adjoint_deformed_drop_central_differences():
    """The C5 droplet (96^3, radius 16, Ca 0.2) run until the wall shear has
    reached and deformed it (8000 steps, sqrt(nu t) ~ 28 cells), where J = D + A
    depends on the wall speed well above rounding: all three components of the
    adjoint gradient against central differences with h = 1e-6 |theta|, within
    1e-5 relative (checkpoint every 80 steps through the torch.autograd.Function)."""
    from droplet_amd.adjoint import config_with, droplet_loss
    dl = _dl()
    cfg, drops = c5_config()
    nsteps, K = 8000, 80
    theta0 = [cfg.tau_f - 0.5, cfg.kappa_g, cfg.wall_u]
    theta = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    J = droplet_loss(theta, cfg, drops, nsteps, K, (1.0, 1.0))
    J.backward()
    g = theta.grad.numpy()

    def loss_at(th):
        lat = dl.Lattice(config_with(cfg, th))
        lat.init_droplets(drops)
        lat.step(nsteps)
        tape = dl.Tape(lat, nsteps)
        val = tape.loss((1.0, 1.0))[0]
        tape.close()
        lat.close()
        return val

    assert loss_at(theta0) == J.item()
    D = dl.Lattice(config_with(cfg, theta0))
    D.init_droplets(drops)
    D.step(nsteps)
    deformation = D.deformation(2)[2]
    D.close()
    print(f"J = {J.item():.10e}, D = {deformation:.4e}")
    assert deformation > 0.01
    rows = []
    for k in range(3):
        h = 1e-6 * abs(theta0[k])
        tp, tm = list(theta0), list(theta0)
        tp[k] += h
        tm[k] -= h
        ref = (loss_at(tp) - loss_at(tm)) / (2 * h)
        rel = abs(g[k] - ref) / abs(ref)
        print(f"param {k}: adjoint {g[k]:.10e}  central difference {ref:.10e}  rel {rel:.2e}")
        rows.append({"param": ("M", "kappa_g", "wall_u")[k], "adjoint": float(g[k]), "central_difference": ref,
                     "rel": rel})
    record("adjoint_deformed_drop", {"steps": nsteps, "K": K, "J": J.item(), "D": deformation, "grad": rows})
    for r in rows:
        assert r["rel"] < 1e-5
