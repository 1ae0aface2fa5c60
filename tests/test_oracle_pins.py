"""Pins of the CPU restatement that do not need the reference binary.

Each test checks something the mathematics fixes, chosen so that a plausible
slip in the restatement (dropped term, wrong sign or index, transposed operand)
fails at least one of them; test_planted_mistakes_are_caught demonstrates that
with deliberately broken builds of the restatement.
"""
import math
import os

import numpy as np
import pytest

from droplet_amd.abi import CX, CY, CZ, E, F, H, OPP, DRP_WALL_Y, DRP_WALL_Z, POP_MOM, POP_PHI, make_config
from oracle import pyoracle
from oracle.pyoracle import OracleLattice
from tests import golden_data as gd

C = np.array([CX, CY, CZ], dtype=float)  # [3][15]
Ew, Fw, Hw = np.array(E), np.array(F), np.array(H)


def test_lattice_identities():
    """D3Q15 weights (program/lbm_drop_xy.f90:70-84): isotropy up to fourth order."""
    d = np.eye(3)
    assert math.isclose(Ew.sum(), 1.0) and math.isclose(Hw.sum(), 1.0) and abs(Fw.sum()) < 1e-15
    np.testing.assert_allclose(np.einsum("i,ai->a", Ew, C), 0, atol=1e-16)
    np.testing.assert_allclose(np.einsum("i,ai,bi->ab", Ew, C, C), d / 3, atol=1e-16)
    np.testing.assert_allclose(np.einsum("i,ai,bi->ab", Fw, C, C), d, atol=1e-15)
    e4 = np.einsum("i,ai,bi,ci,di->abcd", Ew, C, C, C, C)
    iso = (np.einsum("ab,cd->abcd", d, d) + np.einsum("ac,bd->abcd", d, d) + np.einsum("ad,bc->abcd", d, d)) / 9
    np.testing.assert_allclose(e4, iso, atol=1e-16)
    for i in range(15):
        np.testing.assert_array_equal(C[:, OPP[i]], -C[:, i])


def _grid_fields(N, seed=0):
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    return rng, x.astype(float), y.astype(float), z.astype(float)


def test_stencils_exact_on_polynomials():
    """grad is exact on quadratics, lap on quadratics (various_cal :608-633)."""
    N = 12
    _, x, y, z = _grid_fields(N)
    phi = 3.0 + 0.25 * x - 0.5 * y + 0.125 * z + 0.0625 * x * x + 0.03125 * y * y - 0.25 * z * z + 0.5 * x * y
    u1, u2, u3 = 0.5 * x - 0.25 * z, 0.125 * y * y, 0.25 * x * z
    o = OracleLattice(make_config(N))
    o.init_fields(phi, np.full_like(phi, 1 / 3), u1, u2, u3)
    lap, gp, gu = o.stencils()
    inner = (slice(1, -1),) * 3
    np.testing.assert_allclose(lap[inner], 2 * (0.0625 + 0.03125 - 0.25), atol=1e-12)
    np.testing.assert_allclose(gp[0][inner], (0.25 + 0.125 * x + 0.5 * y)[inner], atol=1e-12)
    np.testing.assert_allclose(gp[1][inner], (-0.5 + 0.0625 * y + 0.5 * x)[inner], atol=1e-12)
    np.testing.assert_allclose(gp[2][inner], (0.125 - 0.5 * z)[inner], atol=1e-12)
    # gu[k][a] = d_a u_k
    np.testing.assert_allclose(gu[0][0][inner], 0.5, atol=1e-12)
    np.testing.assert_allclose(gu[0][2][inner], -0.25, atol=1e-12)
    np.testing.assert_allclose(gu[1][1][inner], (0.25 * y)[inner], atol=1e-12)
    np.testing.assert_allclose(gu[2][0][inner], (0.25 * z)[inner], atol=1e-12)
    np.testing.assert_allclose(gu[2][2][inner], (0.25 * x)[inner], atol=1e-12)
    np.testing.assert_allclose(gu[0][1][inner], 0.0, atol=1e-12)


def test_equilibrium_moments():
    """Moments of feq/geq (equilibrium_cal :452-469):
    sum feq = phi, sum c feq = phi u, sum cc feq = A delta + (2 kf/9) G,
    sum geq = 3p, sum c geq = u, sum cc geq = p delta + u u + (2 kg/9) G,
    with A = p0 - kf phi lap - kf/6 |grad phi|^2 and G = 4.5 gg - 1.5 |g|^2 delta."""
    N = 10
    rng, x, y, z = _grid_fields(N, 1)
    k = 2 * np.pi / N
    phi = 3.5 + 1.2 * np.sin(k * x) * np.cos(k * y) + 0.3 * np.sin(k * z)
    p = 1 / 3 + 0.01 * rng.standard_normal(phi.shape)
    u = [0.02 * rng.standard_normal(phi.shape) for _ in range(3)]
    cfg = make_config(N, kappa_g=0.013, nu2=0.2)
    o = OracleLattice(cfg)
    o.init_fields(phi, p, *u)
    fe, ge, tau = o.equilibria()
    lap, gp, _ = o.stencils()
    g2 = (gp ** 2).sum(0)
    G = 4.5 * np.einsum("a...,b...->ab...", gp, gp) - 1.5 * g2 * np.eye(3)[:, :, None, None, None]
    p0 = phi * cfg.T / (1 - cfg.b * phi) - cfg.a * phi ** 2
    A = p0 - cfg.kappa_f * phi * lap - cfg.kappa_f / 6 * g2
    U = np.array(u)
    np.testing.assert_allclose(fe.sum(0), phi, rtol=1e-14)
    np.testing.assert_allclose(np.einsum("ai,i...->a...", C, fe), phi * U, atol=1e-14)
    np.testing.assert_allclose(np.einsum("ai,bi,i...->ab...", C, C, fe),
                               A * np.eye(3)[:, :, None, None, None] + 2 * cfg.kappa_f / 9 * G, atol=1e-13)
    np.testing.assert_allclose(ge.sum(0), 3 * p, rtol=1e-14)
    np.testing.assert_allclose(np.einsum("ai,i...->a...", C, ge), U, atol=1e-15)
    uu = np.einsum("a...,b...->ab...", U, U)
    np.testing.assert_allclose(np.einsum("ai,bi,i...->ab...", C, C, ge),
                               p * np.eye(3)[:, :, None, None, None] + uu + 2 * cfg.kappa_g / 9 * G, atol=1e-14)
    # tau_g = 1/2 + 3 nu(phi), nu linear between (phi1, nu1) and (phi2, nu2) (ini_op :273-284)
    nu = (phi - cfg.phi1) / (cfg.phi2 - cfg.phi1) * (cfg.nu2 - cfg.nu1) + cfg.nu1
    np.testing.assert_allclose(tau, 0.5 + 3 * nu, rtol=1e-15)


def test_initial_populations_carry_fields():
    N = 8
    rng, *_ = _grid_fields(N, 2)
    phi = 2.211 + 2.6 * (rng.random((N, N, N)) < 0.3)
    p = 1 / 3 + 0.001 * rng.standard_normal((N, N, N))
    u = [0.01 * rng.standard_normal((N, N, N)) for _ in range(3)]
    o = OracleLattice(make_config(N, kappa_g=0.01))
    o.init_fields(phi, p, *u)
    f, g = o.populations(POP_PHI), o.populations(POP_MOM)
    np.testing.assert_allclose(f.sum(0), phi, rtol=1e-14)
    np.testing.assert_allclose(g.sum(0) / 3, p, rtol=1e-14)
    np.testing.assert_allclose(np.einsum("ai,i...->a...", C, g), np.array(u), atol=1e-15)


def test_phi_mass_conserved():
    """Streaming permutes populations and sum feq = phi: sum phi is invariant."""
    N = 16
    o = OracleLattice(make_config(N, kappa_g=0.01, nu2=0.3))
    o.init_droplets([(7.5, 7.5, 7.5, 8.0), (2.0, 12.0, 3.0, 5.0)])
    m0 = o.reduce()["phi_mass"]
    o.step(100)
    assert abs(o.reduce()["phi_mass"] - m0) / m0 < 1e-13


@pytest.mark.parametrize("wall", [DRP_WALL_Y, DRP_WALL_Z])
def test_couette_profile(wall):
    """Plane Couette flow between walls moving -U/+U: u = 2U z/zmax - U
    (prototype_program/single-phase/couette3d.f90:252-261)."""
    n, U = 16, 0.01
    dims = (4, n, 4) if wall == DRP_WALL_Y else (4, 4, n)
    cfg = make_config(*dims, nu1=0.1, nu2=0.1, kappa_g=0.0, wall_axis=wall, wall_u=U)
    o = OracleLattice(cfg, threads=1)
    o.init_fields(np.full(o.shape, cfg.phi1), np.full(o.shape, 1 / 3), 0, 0, 0)
    o.step(6000)
    u1 = o.macro(2)
    prof = u1[0, :, 0] if wall == DRP_WALL_Y else u1[:, 0, 0]
    ut = 2 * U * np.arange(n) / (n - 1) - U
    assert np.abs(prof - ut).sum() / np.abs(ut).sum() < 1e-10


def test_taylor_green_decay():
    """Decaying Taylor-Green vortex: amplitude exp(-nu (k1^2 + k2^2) t)
    (prototype_program/single-phase/taylor.f90:148-150)."""
    N, u0, nu, n = 32, 0.01, 0.1, 300
    cfg = make_config(N, 3, N, nu1=nu, nu2=nu, kappa_g=0.0)
    o = OracleLattice(cfg)
    z, y, x = np.meshgrid(np.arange(N), np.arange(3), np.arange(N), indexing="ij")
    k = 2 * np.pi / N
    u1 = -u0 * np.cos(k * x) * np.sin(k * z)
    u3 = u0 * np.sin(k * x) * np.cos(k * z)
    o.init_fields(np.full(o.shape, cfg.phi1), np.full(o.shape, 1 / 3), u1, 0, u3)
    o.step(n)
    amp = float(np.sum(o.macro(2) * u1) / np.sum(u1 * u1))
    assert abs(amp / math.exp(-nu * n * 2 * k * k) - 1) < 0.015


def _ellipse_phi(cfg, a, b, theta, plane_axis):
    z, y, x = np.meshgrid(np.arange(cfg.nz), np.arange(cfg.ny), np.arange(cfg.nx), indexing="ij")
    xc = 0.5 * (cfg.nx - 1)
    if plane_axis == DRP_WALL_Z:  # shear plane (x, z)
        X, Z = x - xc, z - 0.5 * (cfg.nz - 1)
    else:                          # shear plane (x, y)
        X, Z = x - xc, y - 0.5 * (cfg.ny - 1)
    c, s = math.cos(theta), math.sin(theta)
    r1 = (c * X + s * Z) / a
    r2 = (-s * X + c * Z) / b
    return np.where(r1 * r1 + r2 * r2 <= 1.0, cfg.phi2, cfg.phi1)


@pytest.mark.parametrize("wall", [DRP_WALL_Y, DRP_WALL_Z])
@pytest.mark.parametrize("theta", [0.35, 2.0])
def test_deformation_of_analytic_ellipse(wall, theta):
    """Taylor deformation (L-B)/(L+B) of an elliptic cylinder with semi-axes a, b
    (break_up3, prototype_program/two-phase/dropch.f90:857-948)."""
    N, a, b = 48, 14.0, 8.0
    cfg = make_config(N, wall_axis=wall)
    o = OracleLattice(cfg)
    o.init_fields(_ellipse_phi(cfg, a, b, theta, wall), 1 / 3, 0, 0, 0)
    L, B, D = o.deformation(0)
    assert abs(L - 2 * a) < 1.0 and abs(B - 2 * b) < 1.0
    assert abs(D - (a - b) / (a + b)) < 0.03
    L1, B1, D1 = o.deformation(1)
    assert abs(D1 - (a - b) / (a + b)) < 0.06
    L2, B2, D2 = o.deformation(2)
    assert abs(L2 - 2 * a) < 0.5 and abs(B2 - 2 * b) < 0.5 and abs(D2 - (a - b) / (a + b)) < 0.01


def test_diagnostics_closed_forms():
    """KE, dissipation (central-difference stencil) and bulk free energy of a shear wave."""
    N, U = 16, 0.02
    cfg = make_config(N, nu1=0.1, nu2=0.1)
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    k = 2 * np.pi / N
    o = OracleLattice(cfg)
    phi0 = 3.0
    o.init_fields(np.full(o.shape, phi0), 1 / 3, U * np.sin(k * y), 0, 0)
    st = o.reduce()
    assert math.isclose(st["kinetic"], U * U / 4, rel_tol=1e-12)
    # d_y u_x = U cos(k y) sin(k) for the D3Q15 gradient; strain^2 summed = 2 (d_y u_x)^2
    assert math.isclose(st["dissipation"], 0.5 * cfg.nu1 * U * U * math.sin(k) ** 2, rel_tol=1e-12)
    psi = phi0 * cfg.T * math.log(phi0 / (1 - cfg.b * phi0)) - cfg.a * phi0 * phi0
    assert math.isclose(st["free_energy"], N ** 3 * psi, rel_tol=1e-12)
    assert st["interface_area"] < 1e-12
    assert math.isclose(st["phi_mass"], N ** 3 * phi0, rel_tol=1e-14)
    assert abs(st["interaction"]) < 1e-15


def test_interaction_closed_form():
    """Interface-strain interaction sum_ab d_a phi d_b phi S_ab for phi = phi0 + A sin(kx),
    u1 = B sin(2kx): the D3Q15 gradient of a function of x is the central difference,
    so d_x phi = A sin(k) cos(kx), S_xx = B sin(2k) cos(2kx), and the sum over the grid
    is A^2 B N^3 sin(k)^2 sin(2k) / 4 (the mean of cos^2 t cos 2t is 1/4).  A missing
    1/2 in S, a missing square of grad phi or the gradient of another field would miss it."""
    N, A, B = 16, 0.3, 0.01
    cfg = make_config(N)
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    k = 2 * np.pi / N
    o = OracleLattice(cfg)
    o.init_fields(3.0 + A * np.sin(k * x), 1 / 3, B * np.sin(2 * k * x), 0, 0)
    want = A * A * B * N ** 3 * math.sin(k) ** 2 * math.sin(2 * k) / 4
    assert math.isclose(o.reduce()["interaction"], want, rel_tol=1e-12)
    # u1 = B sin(k y) has only S_xy = S_yx; with phi varying in x only (d_y phi = 0)
    # every product vanishes
    o.init_fields(3.0 + A * np.sin(k * x), 1 / 3, B * np.sin(k * y), 0, 0)
    assert abs(o.reduce()["interaction"]) < 1e-15


def _area_closed_form(libpath=None):
    """Interface area sum |grad phi| (At_cal, program/drop_analysis.f90:445-474) of
    phi = phi0 + A sin(kx) + B sin(ky).  Five D3Q15 directions have c_x = +1 and
    five c_x = -1, so the gradient of a function of x alone is the central
    difference (phi(x+1) - phi(x-1)) / 2: d_x phi = A sin(k) cos(kx), d_y phi =
    B sin(k) cos(ky), and the area is N sin(k) sum_xy sqrt(A^2 cos^2 kx + B^2 cos^2 ky).
    A missing square root, a wrong gradient scale or a component left out misses it."""
    N, A, B = 16, 0.3, 0.2
    k = 2 * np.pi / N
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    o = OracleLattice(make_config(N), libpath=libpath)
    o.init_fields(3.0 + A * np.sin(k * x) + B * np.sin(k * y), 1 / 3, 0, 0, 0)
    xs = np.arange(N)
    cx, cy = np.meshgrid(np.cos(k * xs), np.cos(k * xs), indexing="ij")
    want = N * math.sin(k) * float(np.sqrt(A * A * cx * cx + B * B * cy * cy).sum())
    return math.isclose(o.reduce()["interface_area"], want, rel_tol=1e-12)


def test_interface_area_closed_form():
    assert _area_closed_form()
    # one-dimensional profile: N^2 A sin(k) sum_x |cos kx|
    N, A = 16, 0.3
    k = 2 * np.pi / N
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    o = OracleLattice(make_config(N))
    o.init_fields(3.0 + A * np.sin(k * z), 1 / 3, 0, 0, 0)
    want = N * N * A * math.sin(k) * float(np.abs(np.cos(k * np.arange(N))).sum())
    assert math.isclose(o.reduce()["interface_area"], want, rel_tol=1e-12)


def test_mean_pressure_and_momentum_closed_forms():
    """<p> and sum u (reduction rows of drp_stats): p = p0 + P sin(kz) cos(kx) has
    mean p0; u1 = U sin(ky) sums to 0, a constant u2 = W to N^3 W and
    u3 = V cos(kx) + 2W to 2 N^3 W, so a mean taken as a sum, a component in the
    wrong slot or a sum over a partial grid misses them."""
    N, P, U, V, W, p0 = 16, 0.01, 0.02, 0.015, 0.003, 0.3
    k = 2 * np.pi / N
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    o = OracleLattice(make_config(N))
    o.init_fields(3.0, p0 + P * np.sin(k * z) * np.cos(k * x), U * np.sin(k * y), W, V * np.cos(k * x) + 2 * W)
    st = o.reduce()
    assert math.isclose(st["p_mean"], p0, rel_tol=1e-13)
    assert abs(st["momentum"][0]) < 1e-14
    assert math.isclose(st["momentum"][1], N ** 3 * W, rel_tol=1e-13)
    assert math.isclose(st["momentum"][2], 2 * N ** 3 * W, rel_tol=1e-13)


def test_threads_do_not_change_results():
    cfg = make_config(16, kappa_g=0.01, nu2=0.25, wall_axis=DRP_WALL_Y, wall_u=0.02)
    outs = []
    for t in (1, 4):
        o = OracleLattice(cfg, threads=t)
        o.init_droplets([(7.5, 7.5, 7.5, 9.0)])
        o.step(20)
        outs.append(o.fields())
    for k in outs[0]:
        np.testing.assert_array_equal(outs[0][k], outs[1][k])


def _golden_periodic_matches(libpath):
    z, meta = gd.load("periodic16")
    o = OracleLattice(gd.config_for(meta), libpath=libpath)
    o.init_droplets(gd.drop_for(meta), meta["phys"]["p_init"])
    o.step(10)
    return np.array_equal(o.fields()["phi"], z["fields_10"][0]) and np.array_equal(o.fields()["u1"], z["fields_10"][2])


def _moments_hold(libpath):
    N = 8
    rng = np.random.default_rng(3)
    phi = 3.5 + 0.5 * rng.standard_normal((N, N, N))
    o = OracleLattice(make_config(N, kappa_g=0.02), libpath=libpath)
    o.init_fields(phi, 1 / 3, 0.01, 0.0, 0.0)
    fe, ge, _ = o.equilibria()
    _, gp, _ = o.stencils()
    g2 = (gp ** 2).sum(0)
    Gxy = 4.5 * gp[0] * gp[1]
    return np.allclose(np.einsum("i,i...->...", C[0] * C[1], ge), 0.02 * 2 / 9 * Gxy, atol=1e-14)


def _stencil_exact(libpath):
    N = 10
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    o = OracleLattice(make_config(N), libpath=libpath)
    o.init_fields(3.0 + 0.01 * x * x, 1 / 3, 0, 0, 0)
    lap, _, _ = o.stencils()
    return np.allclose(lap[1:-1, 1:-1, 1:-1], 0.02, atol=1e-12)


def _couette_holds(libpath):
    n, U = 10, 0.01
    cfg = make_config(4, 4, n, kappa_g=0.0, wall_axis=DRP_WALL_Z, wall_u=U)
    o = OracleLattice(cfg, threads=1, libpath=libpath)
    o.init_fields(np.full(o.shape, cfg.phi1), 1 / 3, 0, 0, 0)
    o.step(3000)
    ut = 2 * U * np.arange(n) / (n - 1) - U
    return np.allclose(o.macro(2)[:, 0, 0], ut, atol=1e-9)


def _force_golden_matches(libpath):
    z, meta = gd.load("inject16")
    o = OracleLattice(gd.config_for(meta), libpath=libpath)
    o.init_droplets(gd.drop_for(meta), meta["phys"]["p_init"])
    o.step(meta["step_input"] - 1)
    o.inject_momentum(z["g_injected"])
    n = meta["steps"][1]
    o.step(n - o.step_count)
    return np.array_equal(o.fields()["u1"], z[f"fields_{n}"][2])


@pytest.mark.parametrize("plant,checks", [
    (1, [_golden_periodic_matches, _moments_hold]),   # Korteweg term of geq dropped
    (2, [_force_golden_matches]),                     # body-force sign flipped
    (3, [_golden_periodic_matches, _stencil_exact]),  # Laplacian scaled wrongly
    (4, [_couette_holds]),                            # bottom wall velocity sign
    (5, [_area_closed_form]),                         # interface area without the square root
])
def test_planted_mistakes_are_caught(tmp_path, plant, checks):
    """Every planted mistake makes its pins fail, while the clean build passes them."""
    for chk in checks:
        assert chk(None), f"clean restatement fails {chk.__name__}"
    lib = pyoracle.build_planted(plant, str(tmp_path))
    for chk in checks:
        assert not chk(lib), f"planted mistake {plant} not caught by {chk.__name__}"


def test_noneq_smagorinsky_closed_forms():
    """Non-equilibrium Smagorinsky (program/LES_2.f90:530-551, 713-733).
    (a) A uniform state stays at equilibrium, Pi = 0 up to rounding, so tau = 1/2 + 3 nu.
    (b) Steady Couette flow between walls moving -+U: BGK gives the
    non-equilibrium flux Pi_xy = Pi_yx = -tau gamma / 3 for the shear rate gamma,
    so sqrt(Q) = sqrt(2) tau gamma / 3 and tau solves
    tau = 1/2 + 3 (nu + C S),  S = (sqrt(nu^2 + 18 C sqrt(Q)) - nu) / (6 C),  C = Cs^2 Delta^2.
    With gamma measured from the profile the interior tau matches that root; the
    eddy viscosity is half the molecular one here, so a wrong constant (18, 6),
    a missing square root or the full instead of the non-equilibrium moment
    misses it by far more than the tolerance."""
    from droplet_amd.abi import DRP_LES_NONEQ
    nu, Cs, dl = 0.02, 0.17, 4.0
    cfg = make_config(6, kappa_g=0.01, nu1=nu, nu2=nu, les_cs=Cs, les_delta=dl, les_model=DRP_LES_NONEQ)
    o = OracleLattice(cfg, threads=1)
    o.init_fields(3.0, 1 / 3, 0.02, -0.01, 0.005)
    o.step(3)
    assert np.abs(o.tau() - (0.5 + 3 * nu)).max() <= 1e-12   # rounding-level non-equilibrium only

    ny, U = 32, 0.05
    cfg = make_config(4, ny, 4, kappa_g=0.0, nu1=nu, nu2=nu, wall_axis=DRP_WALL_Y, wall_u=U, les_cs=Cs,
                      les_delta=dl, les_model=DRP_LES_NONEQ)
    o = OracleLattice(cfg, threads=1)
    o.init_fields(np.full(o.shape, cfg.phi1), 1 / 3, 0, 0, 0)
    o.step(12000)
    u = o.macro(2)[0, :, 0]
    gamma = (u[20] - u[11]) / 9.0
    C = Cs * Cs * dl * dl
    tau = 0.5 + 3 * nu
    for _ in range(100):
        S = (math.sqrt(nu * nu + 18 * C * math.sqrt(2) * tau * gamma / 3) - nu) / (6 * C)
        tau = 0.5 + 3 * (nu + C * S)
    got = o.tau()[0, ny // 2, 0]
    assert got - (0.5 + 3 * nu) > 0.03            # the eddy viscosity matters
    assert abs(got - tau) <= 1e-6 * tau, (got, tau)
