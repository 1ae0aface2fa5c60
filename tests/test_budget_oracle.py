"""Pins of the restated derived fields and energy budget (oracle/budget_oracle.c).

Derived fields (program/drop_analysis.f90:400-441):
* the bulk chemical potential is the derivative of the free-energy density,
  mu_bulk(phi) = d psi / d phi, for psi = phi T ln(phi/(1 - b phi)) - a phi^2
  (central differences of psi, several phi);
* the interface part is -kappa_g times the D3Q15 Laplacian, which for a field
  that varies along x only is the second difference: for phi = phi0 + A sin(kx),
  lap phi = -2 A (1 - cos k) sin(kx);
* u . grad phi with the D3Q15 gradient, the central difference: U A sin(k) cos(kx).

Energy budget (program/energy_budget.f90:1251-1700), terms evaluated on given
fields with old = new (stated per test): a uniform state and single Fourier modes
whose sums have closed forms.  Planted mistakes 6 (sign of right1) and 7 (a term
of the chemical potential dropped) are caught.
"""
import math

import numpy as np
import pytest

from droplet_amd.abi import make_config
from oracle import pyoracle

N = 12
K = 2 * math.pi / N


def _grid():
    return np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")  # z, y, x


def _mu_is_psi_derivative(libpath):
    cfg = make_config(N, kappa_g=0.01)
    ok = True
    for ph in (2.3, 3.1, 4.0, 4.8):
        mu = pyoracle.derived(cfg, (ph, 1 / 3, 0, 0, 0), 0, libpath)
        h = 1e-5
        psi = [pyoracle.derived(cfg, (ph + d, 1 / 3, 0, 0, 0), 1, libpath)[0, 0, 0] for d in (h, -h)]
        ok &= abs(mu[0, 0, 0] - (psi[0] - psi[1]) / (2 * h)) <= 1e-8 * abs(mu[0, 0, 0])
        ok &= np.ptp(mu) == 0.0
    return ok


def test_chemical_potential_is_free_energy_derivative():
    assert _mu_is_psi_derivative(None)


def test_chemical_potential_laplacian_and_advection():
    z, y, x = _grid()
    phi0, A, U = 3.2, 0.4, 0.03
    phi = phi0 + A * np.sin(K * x)
    kg = 0.013
    f = (phi, 1 / 3, U, 0.0, 0.0)
    mu = pyoracle.derived(make_config(N, kappa_g=kg), f, 0)
    mu0 = pyoracle.derived(make_config(N, kappa_g=0.0), f, 0)
    lap = -2 * A * (1 - math.cos(K)) * np.sin(K * x)
    np.testing.assert_allclose(mu - mu0, -kg * lap, rtol=0, atol=1e-13)
    adv = pyoracle.derived(make_config(N), f, 2)
    np.testing.assert_allclose(adv, U * A * math.sin(K) * np.cos(K * x), rtol=0, atol=1e-15)


def _uniform_budget(P):
    cfg = make_config(N, kappa_g=0.01, nu1=0.1, nu2=0.1)
    u, F = (0.01, -0.02, 0.005), (1e-4, 2e-4, -3e-4)
    f = (3.0, P) + u
    return pyoracle.budget_terms(cfg, f, f, F), cfg, u, F


def test_budget_uniform_state():
    """Uniform phi, p, u and force: every stencil term vanishes; left1 is
    N^3 |p - p0(phi)| with the van der Waals p0 = phi T/(1 - b phi) - a phi^2,
    right5 is N^3 u . F."""
    t_hi, cfg, u, F = _uniform_budget(5.0)
    t_lo, _, _, _ = _uniform_budget(-5.0)
    n = N ** 3
    assert math.isclose(t_hi[0] + t_lo[0], n * 10.0, rel_tol=1e-14)
    p0 = 3.0 * cfg.T / (1 - cfg.b * 3.0) - cfg.a * 9.0
    assert math.isclose((t_hi[0] - t_lo[0]) / (2 * n), -p0, rel_tol=1e-12)
    # the stencils of a uniform field vanish up to the rounding of their partial sums
    # (the z sum passes through 3 v), as in the reference's own order
    for k in (1, 2, 3, 4, 5):
        assert abs(t_hi[k]) <= 1e-15, k
    assert math.isclose(t_hi[6], n * sum(a * b for a, b in zip(u, F)), rel_tol=1e-13)


def _right1_closed_form(libpath):
    z, y, x = _grid()
    U, P = 0.02, 0.003
    f = (3.0, 1 / 3 + P * np.cos(K * y), 0.0, U * np.sin(K * y), 0.0)
    t = pyoracle.budget_terms(make_config(N), f, f, libpath=libpath)
    return math.isclose(t[2], U * P * math.sin(K) * N ** 3 / 2, rel_tol=1e-12)


def test_budget_single_modes():
    z, y, x = _grid()
    s, n = math.sin(K), N ** 3
    assert _right1_closed_form(None)
    # right3: u2 = U sin(ky), S22 = U sin k cos(ky): -2 nu sum S:S = -nu U^2 sin^2 k N^3
    nu, U = 0.07, 0.02
    cfg = make_config(N, nu1=nu, nu2=nu)
    f = (3.0, 1 / 3, 0.0, U * np.sin(K * y), 0.0)
    t = pyoracle.budget_terms(cfg, f, f)
    assert math.isclose(t[4], -nu * U * U * s * s * n, rel_tol=1e-12)
    assert abs(t[3]) <= 1e-15        # 2 nu div(u S) with uniform nu sums to zero
    # left2: u1 = U0 + U sin(ky), u2 = V cos(ky): sum u . grad |u|^2/2 = U0 U V sin k N^3 / 2
    U0, V = 0.01, 0.015
    f = (3.0, 1 / 3, U0 + U * np.sin(K * y), V * np.cos(K * y), 0.0)
    t = pyoracle.budget_terms(cfg, f, f)
    assert math.isclose(t[1], U0 * U * V * s * n / 2, rel_tol=1e-11)
    # right2 with nu(phi): phi = phi0 + A cos(2ky) gives nu = c0 + c1 cos(2ky), c1 = A dnu / (max - min);
    # u2 = U sin(ky): 2 sum nu d_y(u2 S22) = U^2 sin k sin 2k c1 N^3 / 2
    nu1, nu2, phi0, A = 0.05, 0.2, 3.5, 0.3
    cfg = make_config(N, nu1=nu1, nu2=nu2)
    phi = phi0 + A * np.cos(2 * K * y)
    f = (phi, 1 / 3, 0.0, U * np.sin(K * y), 0.0)
    lo, hi = 2.0, 5.0
    t = pyoracle.budget_terms(cfg, f, f, phi_range=(lo, hi))
    c1 = A * (nu2 - nu1) / (hi - lo)
    assert math.isclose(t[3], U * U * s * math.sin(2 * K) * c1 * n / 2, rel_tol=1e-11)
    # right4: phi = phi0 + A sin(kx), u1 = U sin(2kx):
    # -(2/9) kappa_g sum u1 d_x G11, G11 = 3 (d_x phi)^2  ->  kappa_g U A^2 sin^2 k sin 2k N^3 / 6
    kg = 0.02
    cfg = make_config(N, kappa_g=kg)
    f = (phi0 + A * np.sin(K * x), 1 / 3, U * np.sin(2 * K * x), 0.0, 0.0)
    t = pyoracle.budget_terms(cfg, f, f)
    assert math.isclose(t[5], kg * U * A * A * s * s * math.sin(2 * K) * n / 6, rel_tol=1e-11)
    # right5: u1 = U sin(kx), F1 = F sin(kx): U F N^3 / 2
    Fa = 3e-4
    f = (3.0, 1 / 3, U * np.sin(K * x), 0.0, 0.0)
    t = pyoracle.budget_terms(make_config(N), f, f, force=(Fa * np.sin(K * x), 0.0, 0.0))
    assert math.isclose(t[6], U * Fa * n / 2, rel_tol=1e-12)


def test_budget_time_levels():
    """p0, lap phi and grad phi come from the old phi, the phi multiplying lap phi
    in left1 from the new one (as the reference reads phi_procs after physics),
    the strain from the old velocity and every velocity factor from the new one."""
    z, y, x = _grid()
    U = 0.02
    cfg = make_config(N, nu1=0.1, nu2=0.1)
    old = (3.0, 1 / 3, 0.0, U * np.sin(K * y), 0.0)
    new = (3.0, 1 / 3, 0.0, 0.0, 0.0)
    t = pyoracle.budget_terms(cfg, old, new)
    assert math.isclose(t[4], -0.1 * U * U * math.sin(K) ** 2 * N ** 3, rel_tol=1e-12)  # S from old u
    t = pyoracle.budget_terms(cfg, new, old)
    assert abs(t[4]) <= 1e-30
    # left1 with a sinusoidal old phi: the Laplacian term carries the new phi
    A, kg = 0.2, 0.01
    cfg = make_config(N, kappa_g=kg)
    phio = 3.0 + A * np.sin(K * x)
    t1 = pyoracle.budget_terms(cfg, (phio, 50.0, 0, 0, 0), (3.0, 50.0, 0, 0, 0))
    t2 = pyoracle.budget_terms(cfg, (phio, 50.0, 0, 0, 0), (3.0 + np.sin(K * x), 50.0, 0, 0, 0))
    # sum (phi_new2 - phi_new1) lap phi_old = sum sin(kx) (-2 A (1 - cos k) sin(kx))
    want = kg * (-2 * A * (1 - math.cos(K))) * N ** 3 / 2
    assert abs((t2[0] - t1[0]) - want) <= 1e-9 * abs(want) + 1e-12


@pytest.mark.parametrize("plant,check", [(6, _right1_closed_form), (7, _mu_is_psi_derivative)])
def test_planted_budget_mistakes_are_caught(tmp_path, plant, check):
    assert check(None)
    assert not check(pyoracle.build_planted(plant, str(tmp_path)))
