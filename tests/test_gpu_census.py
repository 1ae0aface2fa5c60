"""GPU droplet census (drp_droplet_census) against the CPU restatement.

Cluster counts, volumes, first indices and per-node labels must be identical
(integer results).  Fields: random blob fields on cubic and non-cubic grids,
the special cases (no droplet, everything droplet, the checkerboard with the
largest possible number of clusters), the 71 sphere layouts of the dsd32
fixture (whose reference counts the census matches wherever no droplet crosses
a periodic seam), the FP32 lattice, and a running turbulent-droplet state at
256^3 (timed against the restatement).
"""
import os
import sys
import time

import numpy as np
import pytest
from scipy import ndimage

from droplet_amd.abi import DRP_FP32_SHIFTED, make_config
from oracle.pyoracle import census
from tests import golden_data as gd

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
from make_golden import sphere_field  # noqa: E402

pytestmark = pytest.mark.gpu
THR = 0.5 * (2.211 + 4.895)


def _lattice(phi, precision=0):
    from droplet_amd.lattice import Lattice
    nz, ny, nx = phi.shape
    lat = Lattice(make_config(nx, ny, nz, kappa_g=0.002, precision=precision))
    z0 = np.zeros(phi.shape)
    lat.init_fields(phi, np.full(phi.shape, 1.0 / 3.0), z0, z0, z0)
    return lat


def _check(lat, phi):
    f, v, lab = lat.droplet_census(THR, labels=True)
    of, ov, olab = census(phi, THR)
    np.testing.assert_array_equal(f, of)
    np.testing.assert_array_equal(v, ov)
    np.testing.assert_array_equal(lab, olab)
    return len(f)


@pytest.mark.parametrize("shape,seed,level", [((64, 64, 64), 1, 1.0), ((36, 40, 48), 2, 0.7), ((17, 9, 13), 3, 0.5),
                                              ((64, 64, 64), 4, 0.2)])
def test_census_random_fields(shape, seed, level):
    rng = np.random.default_rng(seed)
    noise = ndimage.gaussian_filter(rng.standard_normal(shape), 1.5, mode="wrap")
    phi = np.where(noise > level * noise.std(), 4.895, 2.211)
    n = _check(_lattice(phi), phi)
    print(f"{shape} level {level}: {n} droplets")


def test_census_special_fields():
    shape = (12, 10, 8)
    for phi in (np.full(shape, 2.211), np.full(shape, 4.895)):
        _check(_lattice(phi), phi)
    z, y, x = np.indices(shape)
    checker = np.where((x + y + z) % 2 == 0, 4.895, 2.211)
    assert _check(_lattice(checker), checker) == checker.size // 2
    # odd sides: the wrap joins some checkerboard nodes; still within the record buffer
    z, y, x = np.indices((9, 11, 13))
    odd = np.where((x + y + z) % 2 == 0, 4.895, 2.211)
    _check(_lattice(odd), odd)


def test_census_golden_layouts():
    z, meta = gd.load("dsd32")
    N = meta["N"]
    lat = None
    for k in range(len(meta["steps"])):
        d = z["drops"][k, :z["ndrops"][k]]
        phi = sphere_field(N, d)
        lat = _lattice(phi)
        n = _check(lat, phi)
        crosses = bool(((d[:, :3] - d[:, 3:4]) < 0).any() or ((d[:, :3] + d[:, 3:4]) > N - 1).any())
        if not crosses:
            assert n == z["counts"][k]


def test_census_fp32_lattice():
    rng = np.random.default_rng(7)
    noise = ndimage.gaussian_filter(rng.standard_normal((32, 32, 32)), 1.5, mode="wrap")
    phi = np.where(noise > 0.8 * noise.std(), 4.895, 2.211)
    lat = _lattice(phi, DRP_FP32_SHIFTED)
    lat.step(20)
    ph = lat.fields()["phi"]
    f, v, lab = lat.droplet_census(THR, labels=True)
    of, ov, olab = census(ph, THR)
    np.testing.assert_array_equal(f, of)
    np.testing.assert_array_equal(v, ov)
    np.testing.assert_array_equal(lab, olab)


def test_census_turbulent_droplets_256():
    from droplet_amd import hit
    from droplet_amd.lattice import Lattice
    N = 256
    phi, p, u1, u2, u3, drops = hit.initial_state(N, seed=3, u_rms=0.03, count=32, D=20.0)
    lat = Lattice(hit.hit_config(N, We=5.0, D=20.0))
    lat.init_fields(phi, p, u1, u2, u3)
    del p, u1, u2, u3
    lat.step(50)
    ph = lat.fields()["phi"]
    lat.droplet_census()
    t0 = time.perf_counter()
    reps = 10
    for _ in range(reps):
        f, v = lat.droplet_census()
    t_gpu = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    of, ov, _ = census(ph, 0.5 * (lat.cfg.phi1 + lat.cfg.phi2))   # parameter set B
    t_cpu = time.perf_counter() - t0
    np.testing.assert_array_equal(f, of)
    np.testing.assert_array_equal(v, ov)
    print(f"census 256^3 turbulent droplets: {len(f)} droplets, volumes {v.min()}..{v.max()}; "
          f"GPU {t_gpu * 1e3:.2f} ms per call (incl. host sync), restatement {t_cpu * 1e3:.0f} ms (1 core)")
    assert len(f) >= 20


def test_droplet_size_histogram():
    """Droplet size distribution (program/mkdsd.f90:68-100 over the radii
    program/percolation.f90:298 forms from volumes, r = (3 V / (4 * 3.14))^(1/3)):
    the GPU census volumes binned exactly as the restatement's volumes binned here,
    on a layout of spheres of several radii, with a d_max cut."""
    N = 48
    rng = np.random.default_rng(21)
    drops = [(float(x), float(y), float(z), float(r)) for (x, y, z), r in
             zip(rng.uniform(0, N, (14, 3)), rng.choice([1.5, 2.6, 3.2, 4.1, 5.0, 6.5], 14))]
    phi = sphere_field(N, drops)
    lat = _lattice(phi)
    scale, bin_, dmax, nbins = 1.135, 2.0, 11.0, 12
    counts, total = lat.droplet_size_histogram(scale, bin_, dmax, nbins, THR)
    _, vols, _ = census(phi, THR)
    want = np.zeros(nbins, np.int64)
    for v in vols:
        d = 2.0 * (3.0 * v / (4.0 * 3.14)) ** (1.0 / 3.0) / scale
        if 0 < d <= dmax and int(d / bin_) < nbins:
            want[int(d / bin_)] += 1
    np.testing.assert_array_equal(counts, want)
    assert total == want.sum() and 0 < total < len(vols)   # the cut drops the largest droplets
