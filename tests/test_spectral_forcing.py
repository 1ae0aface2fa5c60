"""Low-shell linear forcing (program/lbm_drop_iht_xy.f90:722-841) in the CPU
restatement, and the turbulent-droplet initial state.

The reference computes the forcing with 2decomp + FFTW, which are not available,
so the restatement is pinned by numpy's FFT (an independent transform of the same
definition), by the injection invariant <F . u> = eps that the construction
guarantees, and by a single-mode closed form.
"""
import numpy as np
import pytest

from droplet_amd.abi import DRP_FORCE_SPECTRAL, make_config
from droplet_amd import hit
from oracle.pyoracle import OracleLattice


def numpy_forcing(u, kc, eps):
    """The reference's externalforce written with numpy.fft (u: three [z][y][x] arrays)."""
    nz, ny, nx = u[0].shape
    n = u[0].size
    uh = [np.fft.fftn(c) / n for c in u]
    kz, ky, kx = np.meshgrid(np.fft.fftfreq(nz) * nz, np.fft.fftfreq(ny) * ny, np.fft.fftfreq(nx) * nx,
                             indexing="ij")
    k2 = kx * kx + ky * ky + kz * kz
    shells = (k2 >= 1) & (k2 < kc * kc)
    elow = 0.5 * sum(float((np.abs(h[shells]) ** 2).sum()) for h in uh)
    alpha = 0.5 * eps / elow
    return [np.real(np.fft.ifftn(np.where(shells, alpha * h, 0.0))) * n for h in uh], elow


def _oracle(shape, kc, eps):
    nz, ny, nx = shape
    cfg = make_config(nx, ny, nz, force_kind=DRP_FORCE_SPECTRAL, force_amp=eps, force_k=float(kc))
    return OracleLattice(cfg, threads=4)


@pytest.mark.parametrize("shape,kc", [((8, 10, 12), 3), ((16, 16, 16), 2), ((12, 12, 12), 4)])
def test_restatement_matches_numpy_fft(shape, kc):
    rng = np.random.default_rng(7)
    u = [rng.standard_normal(shape) * 0.01 for _ in range(3)]
    eps = 1e-6
    o = _oracle(shape, kc, eps)
    fx, fy, fz, elow = o.spectral_force(*u)
    ref, elow_ref = numpy_forcing(u, kc, eps)
    assert abs(elow - elow_ref) / elow_ref < 1e-12
    scale = max(float(np.abs(r).max()) for r in ref)
    for got, r in zip((fx, fy, fz), ref):
        assert float(np.abs(got - r).max()) / scale < 1e-12
    # injection invariant: <F . u> = eps
    power = float((fx * u[0] + fy * u[1] + fz * u[2]).mean())
    assert abs(power - eps) / eps < 1e-12


def test_single_mode_closed_form():
    """u = A sin(2 pi x / N) (|k| = 1, forced) + B sin(2 pi 3 y / N) (|k| = 3, not forced):
    E_low = A^2 / 4 and F = (2 eps / A^2) A sin(2 pi x / N) along x, zero along y."""
    N, A, B, eps = 16, 0.02, 0.05, 3e-7
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    u1 = A * np.sin(2 * np.pi * x / N)
    u2 = B * np.sin(2 * np.pi * 3 * y / N)
    u3 = np.zeros_like(u1)
    o = _oracle((N, N, N), 3, eps)
    fx, fy, fz, elow = o.spectral_force(u1, u2, u3)
    assert abs(elow - A * A / 4) / (A * A / 4) < 1e-12
    assert np.abs(fx - 2 * eps / (A * A) * u1).max() < 1e-12 * 2 * eps / A
    assert np.abs(fy).max() < 1e-12 * 2 * eps / A and np.abs(fz).max() < 1e-12 * 2 * eps / A


def test_restatement_step_applies_the_force():
    """A forced step differs from an unforced one exactly by the collision term of F."""
    N = 12
    rng = np.random.default_rng(1)
    u = [rng.standard_normal((N, N, N)) * 0.005 for _ in range(3)]
    phi = np.full((N, N, N), 2.211)
    p = np.full((N, N, N), 1.0 / 3.0)
    forced = OracleLattice(make_config(N, force_kind=DRP_FORCE_SPECTRAL, force_amp=1e-7, force_k=3.0), threads=2)
    plain = OracleLattice(make_config(N), threads=2)
    for o in (forced, plain):
        o.init_fields(phi, p, *u)
        o.step(1)
    F = forced.force()
    ref, _ = numpy_forcing(u, 3, 1e-7)
    for a, b in zip(F, ref):
        assert np.abs(a - b).max() < 1e-12 * np.abs(b).max()
    # momentum gained in one step: the force enters each direction at its upstream node,
    # du(x) = sum_q 3 E_q c_q (c_q . F(x - c_q))  (fg_cal, program/lbm_drop_iht_xy.f90)
    from oracle.torch_step import CX, CY, CZ, E
    du_ref = [np.zeros((N, N, N)) for _ in range(3)]
    for q in range(1, 15):
        c = (CX[q], CY[q], CZ[q])
        up = [np.roll(f, shift=(c[2], c[1], c[0]), axis=(0, 1, 2)) for f in ref]   # F(x - c)
        dot = c[0] * up[0] + c[1] * up[1] + c[2] * up[2]
        for a in range(3):
            du_ref[a] += 3.0 * E[q] * c[a] * dot
    for a in range(3):
        du = forced.macro(2 + a) - plain.macro(2 + a)
        assert np.abs(du - du_ref[a]).max() < 1e-9 * np.abs(du_ref[a]).max()


def test_turbulent_initial_state():
    N = 32
    phi, p, u1, u2, u3, drops = hit.initial_state(N, seed=3, u_rms=0.03, count=6, D=8.0)
    rms = np.sqrt((u1 ** 2 + u2 ** 2 + u3 ** 2).mean() / 3)
    assert abs(rms - 0.03) < 1e-12
    # solenoidal in the spectral sense
    k1 = np.fft.fftfreq(N) * N
    kz, ky, kx = np.meshgrid(k1, k1, k1, indexing="ij")
    div = kx * np.fft.fftn(u1) + ky * np.fft.fftn(u2) + kz * np.fft.fftn(u3)
    assert np.abs(div).max() < 1e-9 * np.abs(np.fft.fftn(u1)).max()
    # energy only on shells 1..8, and a -5/3 slope between shells 2 and 6
    kmag = np.sqrt(kx * kx + ky * ky + kz * kz)
    e = 0.5 * (np.abs(np.fft.fftn(u1)) ** 2 + np.abs(np.fft.fftn(u2)) ** 2 + np.abs(np.fft.fftn(u3)) ** 2)
    shell = np.floor(kmag + 0.5)
    assert e[(shell > 8.5) | (shell < 0.5)].sum() < 1e-20 * e.sum()
    E = np.array([e[shell == s].sum() for s in range(1, 9)])
    slope = np.polyfit(np.log(np.arange(2, 7)), np.log(E[1:6]), 1)[0]
    assert abs(slope + 5.0 / 3.0) < 0.25
    # droplets: periodic separation >= D + gap, phi two-valued
    c = np.array([d[:3] for d in drops])
    for i in range(len(c)):
        for j in range(i):
            d = np.abs(c[i] - c[j])
            d = np.minimum(d, N - d)
            assert np.sqrt((d * d).sum()) >= 8.0 + 2.0
    assert set(np.unique(phi)) <= {hit.SET_B["phi1"], hit.SET_B["phi2"]}


def test_weber_number_sets_surface_tension():
    cfg = hit.hit_config(64, We=5.0, D=20.0, nu=1e-3)
    eps = 1e-9
    sigma = eps ** (2 / 3) * 20.0 ** (5 / 3) / 5.0
    assert abs(cfg.kappa_g - sigma / 2.95e-3) < 1e-15
    assert cfg.force_kind == DRP_FORCE_SPECTRAL and cfg.force_amp == eps and cfg.force_k == 3.0
