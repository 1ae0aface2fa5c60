"""Derived fields, the phi-equation cross spectrum and the per-step energy budget
on the GPU, against the restatement (oracle/budget_oracle.c) and numpy.fft.

* drp_get_derived: chemical potential and free-energy density within 1e-14
  relative (the device logarithm may differ from the host's in the last bit),
  advection u . grad phi bit for bit; FP64 and FP32 lattices.
* drp_phi_contribution (program/drop_analysis.f90:476-578) against a literal
  numpy restatement: unnormalised transforms of u . grad phi and mu, half space
  k3 = 0 .. nz/2 with weight 2 off k3 = 0, per-shell sums over the weighted mode
  counts, the large / medium / small bands with the reference's minus sign.
* drp_step_energy_budget (program/energy_budget.f90:1251-1700): the step is the
  ordinary step (fields bit-identical to the restatement's) and the seven terms
  agree with the restatement's terms of the same step within 1e-11.
"""
import math

import numpy as np
import pytest

from droplet_amd.abi import (DRP_FORCE_FIELD, DRP_FORCE_TAYLOR_GREEN, DRP_FP32_SHIFTED, DRP_WALL_Y, DRP_WALL_Z,
                             make_config)
from oracle import pyoracle
from oracle.pyoracle import OracleLattice

pytestmark = pytest.mark.gpu


def _state(shape, seed):
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    r = np.sqrt((x - nx / 2) ** 2 + (y - ny / 2) ** 2 + (z - nz / 2) ** 2)
    phi = 2.211 + (4.895 - 2.211) * 0.5 * (1 + np.tanh((0.3 * min(shape) - r) / 2.0))
    phi = phi + 0.01 * rng.standard_normal(shape)
    u = [0.01 * np.sin(2 * np.pi * (x / nx + k * y / ny)) + 0.002 * rng.standard_normal(shape) for k in range(3)]
    return phi, 1 / 3 + 0.001 * rng.standard_normal(shape), u[0], u[1], u[2]


@pytest.mark.parametrize("precision", [0, DRP_FP32_SHIFTED])
def test_derived_fields_vs_restatement(precision):
    from droplet_amd.lattice import Lattice
    shape = (16, 18, 20)
    cfg = make_config(20, 18, 16, precision=precision, kappa_g=0.013)
    lat = Lattice(cfg)
    lat.init_fields(*_state(shape, 5))
    f = lat.fields()   # the FP64 view the derived fields are taken of
    fields = (f["phi"], f["p"], f["u1"], f["u2"], f["u3"])
    for kind in (0, 1, 2):
        got = lat.derived(kind)
        want = pyoracle.derived(cfg, fields, kind)
        if kind == 2:
            assert np.array_equal(got, want)
        else:
            np.testing.assert_allclose(got, want, rtol=1e-14, atol=0)
    with pytest.raises(RuntimeError):
        lat.derived(3)


def _shells(shape):
    def signed(n):
        i = np.arange(n)
        return np.where(i >= n // 2 + 1, i - n, i)
    nz, ny, nx = shape
    k3, k2, k1 = np.meshgrid(signed(nz), signed(ny), signed(nx), indexing="ij")
    return np.floor(np.sqrt((k1 * k1 + k2 * k2 + k3 * k3).astype(float))).astype(int)


def _contribution_ref(cfg, fields):
    """contribution restated in numpy, loop structure of drop_analysis.f90:495-578."""
    shape = fields[0].shape
    nz, nx = shape[0], shape[2]
    nb = nx // 2 + 1
    A = np.fft.fftn(pyoracle.derived(cfg, fields, 2))
    M = np.fft.fftn(pyoracle.derived(cfg, fields, 0))
    sh = _shells(shape)
    w = np.where(np.arange(nz)[:, None, None] == 0, 1.0, 2.0) * np.ones(shape)
    half = np.zeros(shape, bool)
    half[: nz // 2 + 1] = True
    keep = half & (sh < nb)
    num = np.bincount(sh[keep], (w * np.real(A * np.conj(M)))[keep], minlength=nb)
    den = np.bincount(sh[keep], w[keep], minlength=nb)
    res = np.where(den > 0, num / np.where(den > 0, den, 1), 0.0)
    large, medium, small = res[1:4].sum(), res[4:10].sum(), res[10:nb].sum()
    return res, -np.array([res[1:nb].sum(), large, medium, small])


@pytest.mark.parametrize("shape", [(24, 24, 24), (20, 30, 32)])
def test_phi_contribution_vs_numpy(shape):
    from droplet_amd.lattice import Lattice
    nz, ny, nx = shape
    cfg = make_config(nx, ny, nz, kappa_g=0.01)
    lat = Lattice(cfg)
    lat.init_fields(*_state(shape, 9))
    f = lat.fields()
    fields = (f["phi"], f["p"], f["u1"], f["u2"], f["u3"])
    C, bands = lat.phi_contribution()
    Cr, br = _contribution_ref(cfg, fields)
    scale = np.abs(Cr).max()
    assert np.abs(C - Cr).max() <= 1e-10 * scale
    np.testing.assert_allclose(bands, br, rtol=1e-9, atol=1e-10 * scale)
    C3, b3 = lat.phi_contribution(3)   # the shell sums are atomics: equal to rounding
    np.testing.assert_allclose(C3, C[:3], rtol=1e-12, atol=1e-12 * scale)
    np.testing.assert_allclose(b3, bands, rtol=1e-12, atol=1e-12 * scale)
    with pytest.raises(RuntimeError):
        lat.phi_contribution(nx // 2 + 2)


def _budget_case(case):
    # odd_bricks: the budget kernels work 64 x 4 x 16 bricks; nx = 70, ny = 7 and
    # nz = 5 leave partial bricks in every direction and halos that wrap by modulo
    shape = (5, 7, 70) if case == "odd_bricks" else (18, 20, 22)
    nz, ny, nx = shape
    if case == "odd_bricks":
        cfg = make_config(nx, ny, nz, kappa_g=0.01, nu1=0.05, nu2=0.12)
        force = None
    elif case == "walls_z":  # moving walls normal to z (dropch.f90), nu(phi) varying
        cfg = make_config(nx, ny, nz, kappa_g=0.01, nu1=0.05, nu2=0.12, wall_axis=DRP_WALL_Z, wall_u=0.02,
                          wall_start=1)
        force = None
    elif case == "periodic_tg":
        cfg = make_config(nx, ny, nz, kappa_g=0.01, nu1=0.05, nu2=0.12, force_kind=DRP_FORCE_TAYLOR_GREEN,
                          force_amp=2e-5, force_k=2 * math.pi / nx, force_start=3)
        force = None
    else:
        cfg = make_config(nx, ny, nz, kappa_g=0.01, nu1=0.1, nu2=0.1, wall_axis=DRP_WALL_Y, wall_u=0.02,
                          wall_start=2, force_kind=DRP_FORCE_FIELD)
        rng = np.random.default_rng(2)
        force = [1e-5 * rng.standard_normal(shape) for _ in range(3)]
    return cfg, shape, force


@pytest.mark.parametrize("case", ["periodic_tg", "walls_field", "odd_bricks", "walls_z"])
def test_energy_budget_vs_restatement(case):
    from droplet_amd.lattice import Lattice
    cfg, shape, force = _budget_case(case)
    init = _state(shape, 4)
    lat, ora = Lattice(cfg), OracleLattice(cfg)
    for o in (lat, ora):
        o.init_fields(*init)
        if force is not None:
            o.set_force(*force)
    for nsteps in (1, 4):   # budgets of steps 2 and 7 (the force switches on at step 3)
        lat.step(nsteps)
        ora.step(nsteps)
        got = lat.step_energy_budget()
        f = ora.fields()
        old = (f["phi"], f["p"], f["u1"], f["u2"], f["u3"])
        ora.step(1)
        f = ora.fields()
        new = (f["phi"], f["p"], f["u1"], f["u2"], f["u3"])
        fl = lat.fields()
        for k in f:
            assert np.array_equal(fl[k], f[k]), k
        want = pyoracle.budget_terms(cfg, old, new, ora.force(), ora.phi_range())
        scale = np.abs(want).max()
        np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-12 * scale, err_msg=f"{case} step {lat.step_count}")
        assert lat.step_count == ora.step_count


def test_energy_budget_errors():
    from droplet_amd.abi import DRP_MODEL_ALLEN_CAHN
    from droplet_amd.lattice import Lattice
    lat = Lattice(make_config(12, precision=DRP_FP32_SHIFTED))
    lat.init_droplets([(6, 6, 6, 6)])
    with pytest.raises(RuntimeError):
        lat.step_energy_budget()
    lat = Lattice(make_config(12, model=DRP_MODEL_ALLEN_CAHN, ac_a=0.4, ac_width=3.0, phi1=-0.5, phi2=0.5))
    lat.init_droplets([(6, 6, 6, 6)], 0.0)
    with pytest.raises(RuntimeError):
        lat.step_energy_budget()
