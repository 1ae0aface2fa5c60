"""Scale-separated advection of phi (program/lbm_drop_scale_xy.f90) on the GPU
against the restatement: the band-passed velocity comes from a truncated
separable DFT on the GPU and from a direct DFT in the restatement, so the fields
agree to rounding (1e-12 relative after 30 steps) rather than bit for bit; before
advect_start the step is the ordinary one and bit-identical.  Tile shapes,
walls, a Taylor-Green force, a viscosity ratio and the uniform-stream droplet
are covered.
"""
import math

import numpy as np
import pytest

from droplet_amd.abi import DRP_FORCE_TAYLOR_GREEN, DRP_WALL_Z, make_config
from oracle.pyoracle import OracleLattice

pytestmark = pytest.mark.gpu


def _state(shape, seed):
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    r = np.sqrt((x - nx / 2) ** 2 + (y - ny / 2) ** 2 + (z - nz / 2) ** 2)
    phi = np.where(r <= min(shape) / 4, 4.895, 2.211) + 0.01 * rng.standard_normal(shape)
    u = [0.01 * np.sin(2 * np.pi * (x / nx + k * y / ny)) + 0.003 * rng.standard_normal(shape) for k in range(3)]
    return phi, 1 / 3 + 0.001 * rng.standard_normal(shape), u[0], u[1], u[2]


def _max_rel(a, b):
    worst = 0.0
    for k in a:
        scale = max(float(np.abs(b[k]).max()), 1e-3)
        worst = max(worst, float(np.abs(a[k] - b[k]).max()) / scale)
    return worst


CASES = {
    "periodic_tg": dict(kappa_g=0.01, nu1=0.1, nu2=0.1, advect_k_high=2, advect_k_low=0, advect_start=4,
                        force_kind=DRP_FORCE_TAYLOR_GREEN, force_amp=1e-5, force_k=2 * math.pi / 22, force_start=2),
    "walls_ratio": dict(kappa_g=0.01, nu1=0.05, nu2=0.1, advect_k_high=3, advect_k_low=1, advect_start=4,
                        wall_axis=DRP_WALL_Z, wall_u=0.02, wall_start=1),
}


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("shape_id", [None, 0])
def test_scale_advection_vs_restatement(case, shape_id):
    from droplet_amd.lattice import Lattice
    shape = (20, 18, 22)
    nz, ny, nx = shape
    cfg = make_config(nx, ny, nz, **CASES[case])
    lat, ora = Lattice(cfg), OracleLattice(cfg, threads=8)
    if shape_id is not None:
        lat.tile_shape = shape_id
    f = _state(shape, 6)
    lat.init_fields(*f)
    ora.init_fields(*f)
    lat.step(3)
    ora.step(3)
    a, b = lat.fields(), ora.fields()
    assert all(np.array_equal(a[k], b[k]) for k in a), "before advect_start the step is the ordinary one"
    lat.step(27)
    ora.step(27)
    err = _max_rel(lat.fields(), ora.fields())
    print(f"scale-separated advection {case} shape {lat.tile_shape}: 30 steps, max rel err {err:.2e}")
    assert err <= 1e-12


def test_droplet_in_uniform_stream():
    """A droplet in a uniform stream U: carried U t with the ordinary model and
    with k = 0 in the band, not carried when the band leaves out k = 0."""
    from droplet_amd.lattice import Lattice
    N, steps, U = 32, 200, 0.02
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    r = np.sqrt((x - N / 2) ** 2 + (y - N / 2) ** 2 + (z - N / 2) ** 2)
    phi = 2.211 + (4.895 - 2.211) * 0.5 * (1 + np.tanh((8.0 - r) / 1.5))

    def centroid(lat):
        w = np.clip(lat.macro(0) - 2.211, 0, None)
        ang = 2 * np.pi * x / N
        return N * (np.arctan2((w * np.sin(ang)).sum(), (w * np.cos(ang)).sum()) / (2 * np.pi)) % N

    drift = {}
    for name, kw in (("ordinary", {}), ("band 0..1", dict(advect_k_high=2, advect_k_low=0, advect_start=1)),
                     ("band 1..2", dict(advect_k_high=3, advect_k_low=1, advect_start=1))):
        lat = Lattice(make_config(N, kappa_g=0.002, **kw))
        lat.init_fields(phi, 1 / 3, U, 0.0, 0.0)
        c0 = centroid(lat)
        lat.step(steps)
        drift[name] = (centroid(lat) - c0 + N / 2) % N - N / 2
    print(drift)
    want = U * steps
    assert abs(drift["ordinary"] - want) <= 0.1 * want
    assert abs(drift["band 0..1"] - want) <= 0.1 * want
    assert abs(drift["band 1..2"]) <= 0.1 * want
