"""FP64 HIP path against the reference's golden vectors and the CPU restatement.

All calls go through the C ABI (libdroplet_amd.so via ctypes).  The bar is the
one the FP64 path is specified to: relative L-infinity error <= 1e-10 on phi,
p and u after N steps, and phi mass conserved to 1e-13 relative; the kernels
keep the reference's operation order with FMA contraction off, so in practice
the results are expected to be bit-identical and the tests report it.
"""
import numpy as np
import pytest

from droplet_amd.abi import (DRP_FORCE_FIELD, DRP_WALL_Y, DRP_WALL_Z, POP_MOM, POP_PHI, kappa_g_from_sigma,
                             make_config)
from oracle.pyoracle import OracleLattice
from tests import golden_data as gd

pytestmark = pytest.mark.gpu
FIELDS = ("phi", "p", "u1", "u2", "u3")
TOL = 1e-10


def Lattice(cfg):
    from droplet_amd.lattice import Lattice as _L
    return _L(cfg)


def _compare(got: dict, ref, label: str, tol=TOL):
    worst = 0.0
    exact = True
    for k, fld in enumerate(FIELDS):
        r = ref[k] if not isinstance(ref, dict) else ref[fld]
        err = gd.max_rel(got[fld], r, floor=1e-3 if fld.startswith("u") else 0.0)
        exact &= bool(np.array_equal(got[fld], r))
        worst = max(worst, err)
        assert err <= tol, f"{label}: {fld} rel Linf {err:.3e} > {tol}"
    print(f"{label}: worst rel Linf {worst:.3e} bit-exact={exact}")
    return exact


@pytest.mark.parametrize("name", ["periodic16", "periodic32"])
def test_periodic_golden(name):
    z, meta = gd.load(name)
    lat = Lattice(gd.config_for(meta))
    lat.init_droplets(gd.drop_for(meta), meta["phys"]["p_init"])
    for n in meta["steps"]:
        lat.step(n - lat.step_count)
        _compare(lat.fields(), z[f"fields_{n}"], f"{name}@{n}")


def test_periodic_golden_populations():
    z, meta = gd.load("periodic16")
    lat = Lattice(gd.config_for(meta))
    lat.init_droplets(gd.drop_for(meta), meta["phys"]["p_init"])
    lat.step(10)
    ref = z["pops_10"]
    for which in (POP_PHI, POP_MOM):
        got = lat.populations(which)
        assert gd.max_rel(got, ref[which]) <= TOL


def test_injection_taylor_green_golden():
    z, meta = gd.load("inject16")
    lat = Lattice(gd.config_for(meta))
    lat.init_droplets(gd.drop_for(meta), meta["phys"]["p_init"])
    lat.step(meta["step_input"] - 1)
    lat.inject_momentum(z["g_injected"])
    for n in meta["steps"]:
        lat.step(n - lat.step_count)
        _compare(lat.fields(), z[f"fields_{n}"], f"inject16@{n}")


@pytest.mark.skipif(not gd.available("walls24"), reason="walls24 fixture not generated")
def test_moving_walls_golden():
    z, meta = gd.load("walls24")
    lat = Lattice(gd.config_for(meta))
    lat.init_droplets(gd.drop_for(meta), meta["phys"]["p_init"])
    n = meta["steps"][0]
    lat.step(n)
    got = lat.fields()
    assert gd.max_rel(got["phi"], z[f"phi_{n}"]) <= TOL
    assert gd.max_rel(got["p"], z[f"p_{n}"]) <= TOL


def _shear_cfg(N=64, **kw):
    """64^3 shear droplet of the headline case: walls normal to y moving +-U in x,
    tau_g = 0.8 (nu = 0.1), tau_f = 0.7, kappa_f = 0.01."""
    base = dict(nu1=0.1, nu2=0.1, tau_f=0.7, kappa_f=0.01, kappa_g=kappa_g_from_sigma(0.01),
                wall_axis=DRP_WALL_Y, wall_u=0.01, wall_start=1)
    base.update(kw)
    return make_config(N, **base)


@pytest.mark.parametrize("wall", [DRP_WALL_Y, DRP_WALL_Z])
def test_shear_droplet_vs_oracle(wall):
    N = 48
    cfg = _shear_cfg(N, wall_axis=wall, nu2=0.2)
    drop = [(0.5 * (N - 1),) * 3 + (0.4 * N,)]
    lat, ora = Lattice(cfg), OracleLattice(cfg)
    lat.init_droplets(drop, 1.0 / 3.0)
    ora.init_droplets(drop, 1.0 / 3.0)
    for n in (1, 7, 40):
        lat.step(n - lat.step_count)
        ora.step(n - ora.step_count)
        _compare(lat.fields(), ora.fields(), f"shear wall={wall}@{n}")
    assert gd.max_rel(lat.populations(POP_MOM), ora.populations(POP_MOM)) <= TOL


def test_partial_tiles_and_odd_sizes():
    cfg = make_config(19, 13, 11, kappa_g=0.005, nu2=0.15, wall_axis=DRP_WALL_Z, wall_u=0.02)
    drop = [(9.0, 6.0, 5.0, 7.0)]
    lat, ora = Lattice(cfg), OracleLattice(cfg)
    lat.init_droplets(drop)
    ora.init_droplets(drop)
    lat.step(25)
    ora.step(25)
    _compare(lat.fields(), ora.fields(), "19x13x11")


@pytest.mark.parametrize("case", ["all_tiny", "tiny_slab"])
def test_tiny_nonequilibrium_parts_vs_oracle(case):
    """Relaxation with a constant tau divides by a divisor whose part of the
    division is done once per thread; where some f - feq is nonzero but below
    2^-967 (or at 2^766 and beyond) a node redoes its collision with full
    divisions.  Subnormal fields (order 1e-310) drive that path everywhere, and
    fields of order 1e-300 in a slab of nodes next to an ordinary droplet give
    waves where some lanes redo and others do not.  Bit-identical to the
    restatement either way."""
    N = 16
    cfg = make_config(N, kappa_g=0.004, wall_axis=DRP_WALL_Y, wall_u=0.01, wall_start=1)
    shape = (N, N, N)
    rng = np.random.default_rng(11)
    tiny = 1e-300 * (1.0 + rng.random(shape))
    if case == "all_tiny":
        phi = 1e-310 * (1.0 + rng.random(shape))
        p = 0.5 * phi
        u = [1e-312 * rng.standard_normal(shape) for _ in range(3)]
    else:
        z = np.arange(N)[:, None, None] * np.ones(shape)
        r2 = ((np.indices(shape) - 7.5) ** 2).sum(axis=0)
        phi = np.where(r2 <= 25.0, cfg.phi2, cfg.phi1)
        p = np.full(shape, 1.0 / 3.0)
        slab = z >= N - 4
        phi = np.where(slab, tiny, phi)
        p = np.where(slab, 0.5 * tiny, p)
        u = [np.where(slab, 0.0, 0.01 * rng.standard_normal(shape)) for _ in range(3)]
    lat, ora = Lattice(cfg), OracleLattice(cfg)
    for o in (lat, ora):
        o.init_fields(phi, p, *u)
    for n in (1, 2, 6):
        lat.step(n - lat.step_count)
        ora.step(n - ora.step_count)
        a, b = lat.fields(), ora.fields()
        for k in FIELDS:
            assert np.array_equal(a[k], b[k]), f"{case}@{n}: {k}"
    for which in (POP_PHI, POP_MOM):
        assert np.array_equal(lat.populations(which), ora.populations(which)), f"{case}: populations {which}"


def test_les_vs_oracle():
    N = 24
    rng = np.random.default_rng(5)
    cfg = make_config(N, kappa_g=0.004, les_cs=0.1, les_delta=1.0, nu1=0.02, nu2=0.02)
    shape = (N, N, N)
    u = [0.03 * rng.standard_normal(shape) for _ in range(3)]
    phi = np.where(rng.random(shape) < 0.2, cfg.phi2, cfg.phi1)
    lat, ora = Lattice(cfg), OracleLattice(cfg)
    for o in (lat, ora):
        o.init_fields(phi, np.full(shape, 1 / 3), *u)
        o.step(15)
    _compare(lat.fields(), ora.fields(), "LES")


@pytest.mark.parametrize("name", ["les16", "les16z4"])
def test_les_golden(name):
    """program/LES_drop.f90 on 8 x 1 x 2 and 8 x 1 x 4 ranks: eddy viscosity at the
    end of each step with the values across rank blocks one step old, momentum
    injection at step 3 and the Taylor-Green force, 500 steps at 16^3."""
    z, meta = gd.load(name)
    lat = Lattice(gd.config_for(meta))
    lat.init_droplets(gd.drop_for(meta), meta["phys"]["p_init"])
    lat.step(meta["step_input"] - 1)
    lat.inject_momentum(z["g_injected"])
    for n in meta["steps"]:
        lat.step(n - lat.step_count)
        assert _compare(lat.fields(), z[f"fields_{n}"], f"{name}@{n}")


def test_les_rank_blocks_vs_oracle():
    """A 4 x 2 x 2 rank split on a non-cubic grid with walls and variable
    viscosity, an injection (tau kept) and a population replacement (tau
    recomputed with current neighbours) in between."""
    nx, ny, nz = 24, 16, 20
    rng = np.random.default_rng(12)
    cfg = make_config(nx, ny, nz, kappa_g=0.004, les_cs=0.15, les_delta=2.0, nu1=0.02, nu2=0.05,
                      wall_axis=DRP_WALL_Y, wall_u=0.01, les_ranks_x=4, les_ranks_y=2, les_ranks_z=2)
    shape = (nz, ny, nx)
    u = [0.03 * rng.standard_normal(shape) for _ in range(3)]
    phi = np.where(rng.random(shape) < 0.2, cfg.phi2, cfg.phi1)
    lat, ora = Lattice(cfg), OracleLattice(cfg)
    for o in (lat, ora):
        o.init_fields(phi, np.full(shape, 1 / 3), *u)
        o.step(6)
    g = ora.populations(POP_MOM) * (1 + 1e-3 * rng.standard_normal((15,) + shape))
    for o in (lat, ora):
        o.inject_momentum(g)
        o.step(5)
    f = ora.populations(POP_PHI)
    for o in (lat, ora):
        o.set_populations(POP_PHI, f)
        o.step(5)
    assert _compare(lat.fields(), ora.fields(), "LES 4x2x2 blocks")


def test_force_field_vs_oracle():
    N = 20
    rng = np.random.default_rng(6)
    cfg = make_config(N, kappa_g=0.004, force_kind=DRP_FORCE_FIELD, force_start=3)
    shape = (N, N, N)
    F = [1e-4 * rng.standard_normal(shape) for _ in range(3)]
    lat, ora = Lattice(cfg), OracleLattice(cfg)
    for o in (lat, ora):
        o.init_droplets([(9.5, 9.5, 9.5, 8.0)])
        o.set_force(*F)
        o.step(12)
    _compare(lat.fields(), ora.fields(), "force field")


def test_mass_conservation_periodic():
    """Sum phi is conserved up to the rounding of the reference arithmetic itself.

    The FP64 state is bit-identical to the compiled reference, whose exact drift
    for this case (math.fsum over the CPU restatement's populations) is
    -1.051e-13 after 1000 steps: a systematic ~1.05e-16 per step, i.e. below one
    ulp per node and step.  The bound is therefore per step: 2e-16 (1e-13 over
    500 steps)."""
    N = 64
    cfg = make_config(N, kappa_g=kappa_g_from_sigma(0.01), nu2=0.3)
    lat = Lattice(cfg)
    lat.init_droplets([(31.5, 31.5, 31.5, 25.0)])
    m0 = lat.reduce()["phi_mass"]
    drifts = []
    for n in (500, 1000):
        lat.step(n - lat.step_count)
        drifts.append(abs(lat.reduce()["phi_mass"] - m0) / abs(m0))
    print(f"phi mass drift at 64^3: {drifts[0]:.3e} after 500 steps, {drifts[1]:.3e} after 1000")
    assert drifts[0] <= 1e-13
    assert drifts[1] <= 2e-16 * 1000


def test_population_roundtrip_and_restart():
    N = 16
    cfg = make_config(N, kappa_g=0.01, nu2=0.2)
    a = Lattice(cfg)
    a.init_droplets([(7.5, 7.5, 7.5, 8.0)])
    a.step(5)
    f, g = a.populations(POP_PHI), a.populations(POP_MOM)
    b = Lattice(cfg)
    b.init_droplets([(7.5, 7.5, 7.5, 8.0)])
    b.step(5)
    b.set_populations(POP_PHI, f)
    b.set_populations(POP_MOM, g)
    np.testing.assert_array_equal(b.populations(POP_PHI), f)
    a.step(7)
    b.step(7)
    for k in FIELDS:
        np.testing.assert_array_equal(a.fields()[k], b.fields()[k])


def test_reduce_vs_oracle():
    N = 32
    cfg = _shear_cfg(N, wall_axis=DRP_WALL_Z)
    drop = [(15.5, 15.5, 15.5, 12.0)]
    lat, ora = Lattice(cfg), OracleLattice(cfg)
    for o in (lat, ora):
        o.init_droplets(drop)
        o.step(30)
    a, b = lat.reduce(), ora.reduce()
    for k in ("phi_mass", "phi_min", "phi_max", "kinetic", "dissipation", "interface_area", "free_energy",
              "p_mean", "interaction"):
        assert abs(a[k] - b[k]) <= 1e-11 * max(1.0, abs(b[k])), (k, a[k], b[k])
    np.testing.assert_allclose(a["momentum"], b["momentum"], rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_deformation_vs_oracle(mode):
    N = 48
    cfg = _shear_cfg(N, wall_axis=DRP_WALL_Z, wall_u=0.02)
    drop = [(23.5, 23.5, 23.5, 18.0)]
    lat, ora = Lattice(cfg), OracleLattice(cfg)
    for o in (lat, ora):
        o.init_droplets(drop)
        o.step(200)
    a, b = lat.deformation(mode), ora.deformation(mode)
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12)


def test_tile_shapes_agree(monkeypatch):
    import droplet_amd.lattice as dl
    N = 32
    cfg = _shear_cfg(N, nu2=0.25)
    outs = []
    for shape in range(4):
        monkeypatch.setenv("DROPLET_SHAPE", str(shape))
        lat = dl.Lattice(cfg)
        lat.init_droplets([(15.5, 15.5, 15.5, 12.0)])
        lat.step(20)
        outs.append(lat.fields())
    for o in outs[1:]:
        for k in FIELDS:
            np.testing.assert_array_equal(o[k], outs[0][k])


def test_errors_are_reported():
    from droplet_amd.lattice import DropletError
    with pytest.raises(DropletError):
        Lattice(make_config(2))
    lat = Lattice(make_config(16))
    with pytest.raises(DropletError):
        lat.step(1)  # before init


def test_allen_cahn_golden_and_oracle():
    """Conservative Allen-Cahn model (prototype_program/two-phase/dropac.f90): the
    GPU against the plane the compiled program wrote (9 significant digits) and
    against the restatement, bit for bit, after 100 and 300 steps."""
    z, meta = gd.load("ac24")
    cfg = gd.config_for(meta)
    lat, ora = Lattice(cfg), OracleLattice(cfg)
    lat.init_droplets(gd.drop_for(meta), meta["phys"]["p_init"])
    ora.init_droplets(gd.drop_for(meta), meta["phys"]["p_init"])
    yp = meta["phys"]["plane_y"]
    for n in meta["steps"]:
        lat.step(n - lat.step_count)
        ora.step(n - ora.step_count)
        got = lat.fields()
        _compare(got, ora.fields(), f"ac24@{n} vs restatement")
        for k, fld in enumerate(("u1", "u2", "u3", "p", "phi")):
            np.testing.assert_allclose(got[fld][:, yp, :], z[f"plane_{n}"][..., k], rtol=6e-9, atol=1e-300)


@pytest.mark.parametrize("wall", [DRP_WALL_Y, DRP_WALL_Z])
def test_allen_cahn_viscosity_ratio_vs_oracle(wall):
    """Allen-Cahn with nu(phi) (nu2 = 2 nu1, the global phi range every step), walls
    normal to y or z, an odd grid: bit for bit against the restatement."""
    from droplet_amd.abi import DRP_MODEL_ALLEN_CAHN
    cfg = make_config(30, 26, 22, model=DRP_MODEL_ALLEN_CAHN, ac_a=0.9, ac_width=3.0, phi1=-0.5, phi2=0.5,
                      nu1=0.1, nu2=0.2, kappa_g=0.005, wall_axis=wall, wall_u=0.02, wall_start=5)
    drop = [(14.5, 12.5, 10.5, 12.0)]
    lat, ora = Lattice(cfg), OracleLattice(cfg)
    lat.init_droplets(drop, 0.0)
    ora.init_droplets(drop, 0.0)
    for n in (1, 9, 60):
        lat.step(n - lat.step_count)
        ora.step(n - ora.step_count)
        _compare(lat.fields(), ora.fields(), f"allen-cahn wall={wall}@{n}")
    st = lat.reduce()
    assert abs(st["phi_mass"] - ora.reduce()["phi_mass"]) <= 1e-12 * abs(st["phi_mass"]) + 1e-12


@pytest.mark.parametrize("case", ["periodic", "walls_force"])
def test_noneq_smagorinsky_vs_oracle(case):
    """Non-equilibrium Smagorinsky (program/LES_2.f90:530-551, 713-733): the step
    kernel keeps sqrt(Pi:Pi) per node and forms tau from it at the next step;
    bit for bit against the restatement, which follows the reference's physics."""
    from droplet_amd.abi import DRP_LES_NONEQ
    N = 20
    rng = np.random.default_rng(8)
    shape = (N, N, N)
    kw = dict(kappa_g=0.004, les_cs=0.17, les_delta=2.0, nu1=0.01, nu2=0.03, les_model=DRP_LES_NONEQ)
    if case == "walls_force":
        kw.update(wall_axis=DRP_WALL_Z, wall_u=0.02, force_kind=DRP_FORCE_FIELD)
    cfg = make_config(N, **kw)
    u = [0.02 * rng.standard_normal(shape) for _ in range(3)]
    phi = np.where(rng.random(shape) < 0.2, cfg.phi2, cfg.phi1)
    F = [1e-5 * rng.standard_normal(shape) for _ in range(3)]
    lat, ora = Lattice(cfg), OracleLattice(cfg)
    for o in (lat, ora):
        o.init_fields(phi, np.full(shape, 1 / 3), *u)
        if case == "walls_force":
            o.set_force(*F)
    for n in (1, 4, 25):
        lat.step(n - lat.step_count)
        ora.step(n - ora.step_count)
        assert _compare(lat.fields(), ora.fields(), f"noneq LES {case}@{n}")
