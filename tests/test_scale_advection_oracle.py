"""Scale-separated advection of phi (program/lbm_drop_scale_xy.f90), restated in
oracle/lbm_oracle.c.  The program needs FFTW and 2decomp, which are absent, so its
parity is pinned here instead of by a fixture:

* the band-pass (low_pass :846-903) against numpy.fft with the same shell mask
  (signed wavenumbers, int|k| in [k_low, k_high)) on odd and even grids;
* the coupling (equilibrium_cal :507-560): from advect_start on, the phi
  equilibria differ from the ordinary model's by exactly E_i phi 3 c_i.(u_band - u)
  and the momentum equilibria not at all (planted mistake 8, the band-passed
  velocity in the momentum equilibria too, is caught); before advect_start the
  steps are bit-identical to the ordinary model's;
* the physics it is for: a droplet in a uniform stream moves with the stream,
  but not when the band leaves out k = 0 (k_low = 1): phi is then carried by the
  fluctuations only.
"""
import math

import numpy as np
import pytest

from droplet_amd.abi import CX, CY, CZ, make_config
from oracle import pyoracle
from oracle.pyoracle import OracleLattice

E = np.array([2 / 9] + [1 / 9] * 6 + [1 / 72] * 8)


def _mask(shape, klo, khi):
    def signed(n):
        i = np.arange(n)
        return np.where(i >= n // 2 + 1, i - n, i)
    nz, ny, nx = shape
    k3, k2, k1 = np.meshgrid(signed(nz), signed(ny), signed(nx), indexing="ij")
    sh = np.floor(np.sqrt((k1 * k1 + k2 * k2 + k3 * k3).astype(float))).astype(int)
    return (sh >= klo) & (sh < khi)


@pytest.mark.parametrize("shape,klo,khi", [((12, 12, 12), 0, 2), ((10, 14, 12), 1, 3), ((13, 11, 12), 0, 3),
                                            ((12, 12, 16), 2, 4)])
def test_band_pass_matches_numpy(shape, klo, khi):
    rng = np.random.default_rng(sum(shape) + klo)
    nz, ny, nx = shape
    o = OracleLattice(make_config(nx, ny, nz), threads=4)
    u = [rng.standard_normal(shape) for _ in range(3)]
    got = o.band_pass(*u, klo, khi)
    m = _mask(shape, klo, khi)
    for g, v in zip(got, u):
        ref = np.real(np.fft.ifftn(np.fft.fftn(v) * m))
        assert np.abs(g - ref).max() <= 1e-13 * np.abs(v).max()


def _state(N, seed):
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    r = np.sqrt((x - N / 2) ** 2 + (y - N / 2) ** 2 + (z - N / 2) ** 2)
    phi = np.where(r <= N / 4, 4.895, 2.211) + 0.01 * rng.standard_normal((N, N, N))
    u = [0.01 * np.sin(2 * np.pi * (x / N + k * y / N)) + 0.003 * rng.standard_normal((N, N, N)) for k in range(3)]
    return phi, 1 / 3 + 0.001 * rng.standard_normal((N, N, N)), u[0], u[1], u[2]


def _coupling_holds(libpath):
    N = 12
    f = _state(N, 3)
    kw = dict(kappa_g=0.01, nu1=0.1, nu2=0.1)
    o_std = OracleLattice(make_config(N, **kw), threads=4, libpath=libpath)
    o_adv = OracleLattice(make_config(N, advect_k_high=3, advect_k_low=1, advect_start=1, **kw), threads=4,
                          libpath=libpath)
    for o in (o_std, o_adv):
        o.init_fields(*f)
    fe_s, ge_s, _ = o_std.equilibria()
    fe_a, ge_a, _ = o_adv.equilibria()
    ub = o_adv.band_pass(*f[2:], 1, 3)
    dcu = sum(np.asarray(c)[:, None, None, None] * (b - u)[None] for c, b, u in zip((CX, CY, CZ), ub, f[2:]))
    want = E[:, None, None, None] * f[0][None] * 3.0 * dcu
    return bool(np.array_equal(ge_a, ge_s) and np.abs((fe_a - fe_s) - want).max() <= 1e-14)


def test_coupling_is_the_phi_advection_term_only():
    assert _coupling_holds(None)


def test_planted_coupling_mistake_is_caught(tmp_path):
    assert not _coupling_holds(pyoracle.build_planted(8, str(tmp_path)))


def test_identical_before_advect_start():
    N = 12
    f = _state(N, 5)
    a = OracleLattice(make_config(N, kappa_g=0.01), threads=4)
    b = OracleLattice(make_config(N, kappa_g=0.01, advect_k_high=2, advect_start=6), threads=4)
    for o in (a, b):
        o.init_fields(*f)
        o.step(5)
    fa, fb = a.fields(), b.fields()
    assert all(np.array_equal(fa[k], fb[k]) for k in fa)
    a.step(1)
    b.step(1)
    assert not np.array_equal(a.fields()["phi"], b.fields()["phi"])


def _drift(cfg, steps):
    N = cfg.nx
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    r = np.sqrt((x - N / 2) ** 2 + (y - N / 2) ** 2 + (z - N / 2) ** 2)
    U = 0.02
    o = OracleLattice(cfg, threads=8)
    o.init_fields(2.211 + (4.895 - 2.211) * 0.5 * (1 + np.tanh((6.0 - r) / 1.5)), 1 / 3, U, 0.0, 0.0)

    def cx():
        w = np.clip(o.macro(0) - 2.211, 0, None)
        ang = 2 * np.pi * x / N   # periodic centroid along x
        return N * (np.arctan2((w * np.sin(ang)).sum(), (w * np.cos(ang)).sum()) / (2 * np.pi)) % N

    x0 = cx()
    o.step(steps)
    d = (cx() - x0 + N / 2) % N - N / 2
    return d, U * steps


def test_droplet_follows_only_the_kept_scales():
    N, steps = 24, 150
    kw = dict(kappa_g=0.002, nu1=0.1, nu2=0.1)
    d_std, want = _drift(make_config(N, **kw), steps)
    assert abs(d_std - want) <= 0.1 * want                     # carried by the uniform stream
    d_mean, _ = _drift(make_config(N, advect_k_high=2, advect_k_low=0, advect_start=1, **kw), steps)
    assert abs(d_mean - want) <= 0.1 * want                    # k = 0 kept: still carried
    d_fluct, _ = _drift(make_config(N, advect_k_high=3, advect_k_low=1, advect_start=1, **kw), steps)
    assert abs(d_fluct) <= 0.1 * want                          # k = 0 removed: not carried
