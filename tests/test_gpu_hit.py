"""Spectral low-shell forcing and the turbulent-droplet configuration on the GPU.

- the HIP forcing (truncated separable DFT) against numpy's FFT of the same
  definition and the injection invariant <F . u> = eps, FP64 and FP32 fields,
  several grid shapes and cutoffs;
- set-B droplets in forced turbulence with Smagorinsky LES: FP64 GPU against the
  CPU restatement, and FP32-shifted against FP64 with stated error.
"""
import numpy as np
import pytest

from droplet_amd import hit
from droplet_amd.abi import DRP_FORCE_SPECTRAL, DRP_FP32_SHIFTED, DRP_FP64, make_config
from oracle.pyoracle import OracleLattice
from tests.test_spectral_forcing import numpy_forcing

pytestmark = pytest.mark.gpu


def _dl():
    from droplet_amd import lattice as dl
    return dl


@pytest.mark.parametrize("precision", [DRP_FP64, DRP_FP32_SHIFTED], ids=["fp64", "fp32"])
@pytest.mark.parametrize("shape,kc", [((8, 10, 12), 3), ((16, 16, 16), 2), ((12, 14, 12), 4), ((20, 20, 20), 5),
                                      ((64, 48, 32), 3)])
def test_spectral_force_matches_numpy(shape, kc, precision):
    dl = _dl()
    nz, ny, nx = shape
    rng = np.random.default_rng(11)
    u = [rng.standard_normal(shape) * 0.01 for _ in range(3)]
    eps = 1e-6
    cfg = make_config(nx, ny, nz, precision=precision, force_kind=DRP_FORCE_SPECTRAL, force_amp=eps,
                      force_k=float(kc))
    lat = dl.Lattice(cfg)
    lat.init_fields(np.full(shape, cfg.phi1), np.full(shape, 1.0 / 3.0), *u)
    lat.step(1)
    F = lat.force()
    if precision == DRP_FP32_SHIFTED:   # the FP32 path forces from its FP32 velocity
        u = [a.astype(np.float32).astype(np.float64) for a in u]
    ref, _ = numpy_forcing(u, kc, eps)
    scale = max(float(np.abs(r).max()) for r in ref)
    err = max(float(np.abs(a - b).max()) for a, b in zip(F, ref)) / scale
    power = float(sum((f * c).mean() for f, c in zip(F, u)))
    print(f"{shape} kc={kc} {'fp32' if precision else 'fp64'}: max rel err {err:.2e}, <F.u>/eps - 1 = {power / eps - 1:.2e}")
    assert err < 1e-12
    assert abs(power - eps) / eps < 1e-12
    lat.close()


def _hit_case(N, precision=DRP_FP64, count=2, D=6.0, steps_kw=None):
    cfg = hit.hit_config(N, We=5.0, D=D, nu=0.01, precision=precision, force_amp=1e-6)
    phi, p, u1, u2, u3, drops = hit.initial_state(N, seed=3, u_rms=0.02, count=count, D=D)
    return cfg, (phi, p, u1, u2, u3)


def test_turbulent_droplets_fp64_vs_restatement():
    dl = _dl()
    N = 16
    cfg, fields = _hit_case(N)
    gpu = dl.Lattice(cfg)
    cpu = OracleLattice(cfg, threads=4)
    gpu.init_fields(*fields)
    cpu.init_fields(*fields)
    gpu.step(30)
    cpu.step(30)
    a, b = gpu.fields(), cpu.fields()
    worst = 0.0
    for k in a:
        scale = max(float(np.abs(b[k]).max()), 1e-3)
        worst = max(worst, float(np.abs(a[k] - b[k]).max()) / scale)
    print(f"turbulent droplets 16^3, LES + spectral forcing, 30 FP64 steps: max rel err vs restatement {worst:.2e}")
    assert worst <= 1e-10
    fa, fb = gpu.force(), cpu.force()
    assert max(float(np.abs(x - y).max()) for x, y in zip(fa, fb)) <= 1e-12 * max(float(np.abs(y).max()) for y in fb)


def test_turbulent_droplets_fp32_vs_fp64():
    dl = _dl()
    N = 32
    c64, fields = _hit_case(N, count=4, D=8.0)
    c32, _ = _hit_case(N, precision=DRP_FP32_SHIFTED, count=4, D=8.0)
    a, b = dl.Lattice(c64), dl.Lattice(c32)
    for o in (a, b):
        o.init_fields(*fields)
        o.step(300)
    fa, fb = a.fields(), b.fields()
    e_phi = float(np.abs(fa["phi"] - fb["phi"]).max()) / c64.phi2
    e_u = max(float(np.abs(fa[k] - fb[k]).max()) for k in ("u1", "u2", "u3"))
    m64, m32 = a.reduce()["phi_mass"], b.reduce()["phi_mass"]
    ke64, ke32 = a.reduce()["kinetic"], b.reduce()["kinetic"]
    print(f"turbulent droplets 32^3, 300 steps, FP32-shifted vs FP64: phi rel Linf {e_phi:.2e}, u Linf {e_u:.2e}, "
          f"mass rel {abs(m32 - m64) / m64:.2e}, KE rel {abs(ke32 - ke64) / ke64:.2e}")
    assert e_phi < 1e-3
    assert e_u < 1e-4
    assert abs(m32 - m64) / m64 < 1e-6
    assert abs(ke32 - ke64) / ke64 < 1e-3


def test_spectral_forcing_not_in_batches_or_tapes():
    dl = _dl()
    cfg = make_config(12, force_kind=DRP_FORCE_SPECTRAL, force_amp=1e-7, force_k=3.0)
    with pytest.raises(dl.DropletError):
        dl.Batch([cfg, cfg])
    lat = dl.Lattice(cfg)
    lat.init_droplets([(5.5, 5.5, 5.5, 5.0)])
    with pytest.raises(dl.DropletError):
        dl.Tape(lat, 4)
    lat.close()
