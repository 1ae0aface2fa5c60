"""The Fortran binding (droplet_amd/fortran/droplet_amd.f90) and the shear driver."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "droplet_amd", "_lib", "shear_driver")

pytestmark = pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built")


def test_layout_check_and_error_path_without_gpu(tmp_path):
    """A grid of 0 nodes passes the module's struct-layout check and is then
    rejected by drp_create's argument validation (before any GPU call)."""
    r = subprocess.run([DRIVER, "0", "1"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "struct layout" not in r.stdout
    assert "drp_create failed (-1): grid must be at least 3^3" in r.stdout


@pytest.mark.gpu
def test_driver_matches_python_front_end(tmp_path):
    from droplet_amd.abi import DRP_WALL_Y, make_config
    from droplet_amd.lattice import Lattice
    from tests.golden.make_golden import read_records
    N, steps = 24, 40
    r = subprocess.run([DRIVER, str(N), str(steps), "10"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    rec = read_records(str(tmp_path / "fields.bin"), 5).reshape(N, N, N, 5)
    cfg = make_config(N, tau_f=0.7, nu1=0.1, nu2=0.1, kappa_f=0.01, kappa_g=(0.01 / 1.7039) ** (1 / 0.9991),
                      wall_axis=DRP_WALL_Y, wall_u=0.01, wall_start=1)
    lat = Lattice(cfg)
    lat.init_droplets([(0.5 * (N - 1),) * 3 + (0.4 * N,)])
    lat.step(steps)
    f = lat.fields()
    np.testing.assert_array_equal(rec[..., 0], f["phi"])
    np.testing.assert_array_equal(rec[..., 1], f["u1"])
    np.testing.assert_array_equal(rec[..., 4], f["p"])
    d = np.loadtxt(tmp_path / "deformation.d")
    assert d.shape == (steps // 10, 7)
    np.testing.assert_allclose(d[-1, 1:4], lat.deformation(0), rtol=1e-14)
