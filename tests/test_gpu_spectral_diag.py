"""FFT diagnostics on the GPU against numpy.fft restatements of the reference's
formulas (FFTW and 2decomp, which the reference calls, are absent; numpy.fft and
closed forms pin them instead).

* the 3-D transform itself against numpy.fft.fftn (radices 2, 3, 4, 5);
* energy spectrum = enesupe_cal (program/turbulence_hit_xy.f90:588-644):
  u_hat = FFT(u)/N^3, shells int|k| with k = i - n for i >= n/2 + 1, half space
  k3 = 0..n/2 weighted 2 off k3 = 0; closed forms of a single Fourier mode;
* low-pass / band-pass filters (program/LES_drop_xy.f90:851-971,
  program/drop_analysis.f90:722-797);
* per-scale interaction (program/drop_analysis.f90:1121-1265), restated in
  numpy literally as the reference computes it: one band-pass and inverse
  transform per shell, the mean of u^(k) . div(grad phi grad phi) times kappa_g.
"""
import math

import numpy as np
import pytest

from droplet_amd.abi import CX, CY, CZ, DRP_FP32_SHIFTED, make_config

pytestmark = pytest.mark.gpu


def _signed(n):
    i = np.arange(n)
    return np.where(i >= n // 2 + 1, i - n, i)


def _shells(shape):
    nz, ny, nx = shape
    k3, k2, k1 = np.meshgrid(_signed(nz), _signed(ny), _signed(nx), indexing="ij")
    return np.floor(np.sqrt((k1 * k1 + k2 * k2 + k3 * k3).astype(float))).astype(int)


def _spectrum_ref(u1, u2, u3, nshells):
    """enesupe_cal restated: half space in k3 (index 0..nz/2), weight 2 off k3 = 0."""
    n = u1.size
    e = 0.5 * sum(np.abs(np.fft.fftn(u) / n) ** 2 for u in (u1, u2, u3))
    nz = u1.shape[0]
    sh = _shells(u1.shape)
    w = np.ones_like(e)
    w[1:] = 2.0
    half = np.zeros(u1.shape, bool)
    half[: nz // 2 + 1] = True
    keep = half & (sh < nshells)
    return np.bincount(sh[keep], (e * w)[keep], minlength=nshells)


def _grad(f):
    """D3Q15 gradient: sum_i c_i f(x + c_i) / 10, arrays [z][y][x]."""
    g = np.zeros((3,) + f.shape)
    for q in range(1, 15):
        v = np.roll(f, (-CZ[q], -CY[q], -CX[q]), axis=(0, 1, 2))
        g[0] += CX[q] * v
        g[1] += CY[q] * v
        g[2] += CZ[q] * v
    return g / 10.0


def _lattice_with(fields, cfg):
    from droplet_amd.lattice import Lattice
    lat = Lattice(cfg)
    lat.init_fields(*fields)
    return lat


def _random_state(shape, seed, smooth=False):
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    if smooth:
        u1 = 0.01 * np.sin(2 * np.pi * x / nx) * np.cos(2 * np.pi * 2 * z / nz)
        u2 = 0.004 * np.cos(2 * np.pi * 3 * y / ny)
        u3 = 0.002 * np.sin(2 * np.pi * (x / nx + y / ny))
    else:
        u1, u2, u3 = (0.01 * rng.standard_normal(shape) for _ in range(3))
    r = np.sqrt((x - nx / 2) ** 2 + (y - ny / 2) ** 2 + (z - nz / 2) ** 2)
    phi = 2.211 + (4.895 - 2.211) * 0.5 * (1 + np.tanh((0.3 * min(shape) - r) / 2.0))
    phi = phi + 0.01 * rng.standard_normal(shape)
    return phi, np.full(shape, 1 / 3), u1, u2, u3


@pytest.mark.parametrize("shape", [(16, 16, 16), (30, 20, 24), (12, 36, 40), (64, 64, 64),
                                   (4, 6, 2048), (3, 1500, 4), (1024, 5, 3)])
def test_fft3d_matches_numpy(shape):
    """Including lines longer than one 1024-point LDS block (2048 along x, 1500
    along y) and of exactly one block (1024 along z), one line per workgroup."""
    from droplet_amd.lattice import fft3d
    rng = np.random.default_rng(sum(shape))
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    fwd = fft3d(a, -1)
    ref = np.fft.fftn(a)
    assert np.abs(fwd - ref).max() <= 1e-12 * np.abs(ref).max()
    back = fft3d(fwd, +1) / a.size
    assert np.abs(back - a).max() <= 1e-13 * np.abs(a).max()


@pytest.mark.parametrize("shape", [(32, 32, 32), (20, 24, 30)])
def test_energy_spectrum_matches_restatement(shape):
    f = _random_state(shape, 7)
    lat = _lattice_with(f, make_config(shape[2], shape[1], shape[0], kappa_g=0.01))
    ns = shape[2] // 2 + 1
    E = lat.energy_spectrum(ns)
    ref = _spectrum_ref(*f[2:], ns)
    assert np.abs(E - ref).max() <= 1e-12 * ref.max()


def test_energy_spectrum_closed_forms():
    """u1 = A cos(2 pi k.x / N) puts A^2/4 in shell int|k|; a smooth field's shells
    sum to the kinetic energy 0.5 <|u|^2> (Parseval, no Nyquist content)."""
    N, A = 32, 0.02
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    for k in [(3, 0, 0), (2, 2, 1), (0, 4, 3), (5, 1, 0)]:
        u1 = A * np.cos(2 * np.pi * (k[0] * x + k[1] * y + k[2] * z) / N)
        lat = _lattice_with((3.0, 1 / 3, u1, 0.0, 0.0), make_config(N))
        E = lat.energy_spectrum(N)
        s = int(math.sqrt(k[0] ** 2 + k[1] ** 2 + k[2] ** 2))
        assert math.isclose(E[s], A * A / 4, rel_tol=1e-12), (k, E[s])
        assert abs(E.sum() - E[s]) <= 1e-15
    f = _random_state((N, N, N), 3, smooth=True)
    lat = _lattice_with(f, make_config(N))
    ke = 0.5 * np.mean(f[2] ** 2 + f[3] ** 2 + f[4] ** 2)
    assert math.isclose(lat.energy_spectrum(N).sum(), ke, rel_tol=1e-12)
    assert math.isclose(lat.reduce()["kinetic"], ke, rel_tol=1e-12)


@pytest.mark.parametrize("kind,k", [(0, 3), (0, 6), (1, 2), (1, 5)])
def test_velocity_filters_match_numpy(kind, k):
    shape = (24, 20, 32)
    f = _random_state(shape, 11)
    lat = _lattice_with(f, make_config(shape[2], shape[1], shape[0]))
    got = lat.filter_velocity(kind, k)
    sh = _shells(shape)
    keep = (sh < k) if kind == 0 else (sh == k)
    for a, u in zip(got, f[2:]):
        ref = np.real(np.fft.ifftn(np.fft.fftn(u) * keep))
        assert np.abs(a - ref).max() <= 1e-14 + 1e-12 * np.abs(ref).max()


def test_scale_interaction_matches_per_shell_restatement():
    """contribution_each_scale restated: for every shell, band-pass u, transform
    back, take kappa_g * mean(u^(k) . w) with w_b = sum_a d_a(d_b phi d_a phi)."""
    shape, kg = (16, 16, 16), 0.004
    f = _random_state(shape, 5)
    lat = _lattice_with(f, make_config(16, kappa_g=kg))
    ns = 16
    C = lat.scale_interaction(ns)
    g = _grad(f[0])
    w = np.zeros((3,) + shape)
    for b in range(3):
        for a in range(3):
            w[b] += _grad(g[b] * g[a])[a]
    sh = _shells(shape)
    ref = np.zeros(ns)
    for s in range(ns):
        us = [np.real(np.fft.ifftn(np.fft.fftn(u) * (sh == s))) for u in f[2:]]
        ref[s] = kg * np.mean(us[0] * w[0] + us[1] * w[1] + us[2] * w[2])
    assert np.abs(C - ref).max() <= 1e-11 * np.abs(ref).max()
    # all shells together: kappa_g <u . w>
    assert math.isclose(C.sum(), kg * np.mean(np.sum(np.array(f[2:]) * w, axis=0)), rel_tol=1e-10)


def test_fp32_lattice_and_errors():
    from droplet_amd.lattice import DropletError
    shape = (16, 16, 16)
    f = _random_state(shape, 9, smooth=True)
    l64 = _lattice_with(f, make_config(16))
    l32 = _lattice_with(f, make_config(16, precision=DRP_FP32_SHIFTED))
    e64, e32 = l64.energy_spectrum(), l32.energy_spectrum()
    assert np.abs(e32 - e64).max() <= 1e-6 * e64.max()
    bad = _lattice_with(_random_state((14, 14, 14), 1), make_config(14))   # 7 is not a radix
    with pytest.raises(DropletError):
        bad.energy_spectrum()
