"""The CPU restatement against golden vectors written by the compiled reference.

The fixtures in tests/golden/ were produced by program/lbm_drop_xy.f90 and
prototype_program/two-phase/dropch.f90 (parameter-patched, MPI, one thread per
rank); tests/golden/make_golden.py is the script.  The restatement keeps the
reference's operation order, so agreement is required to be exact (bit for bit)
or, where stated, within a few ulps.
"""
import numpy as np
import pytest

from droplet_amd.abi import POP_MOM, POP_PHI
from oracle.pyoracle import OracleLattice
from tests import golden_data as gd

FIELDS = ("phi", "p", "u1", "u2", "u3")


def _run_to(o, n):
    o.step(n - o.step_count)


@pytest.mark.parametrize("name", ["periodic16", "periodic32"])
def test_periodic_fields_bit_exact(name):
    z, meta = gd.load(name)
    o = OracleLattice(gd.config_for(meta))
    o.init_droplets(gd.drop_for(meta), meta["phys"]["p_init"])
    for n in meta["steps"]:
        _run_to(o, n)
        ref = z[f"fields_{n}"]
        got = o.fields()
        for k, fld in enumerate(FIELDS):
            np.testing.assert_array_equal(got[fld], ref[k], err_msg=f"{name} step {n} {fld}")


def test_periodic_populations_bit_exact():
    z, meta = gd.load("periodic16")
    o = OracleLattice(gd.config_for(meta))
    o.init_droplets(gd.drop_for(meta), meta["phys"]["p_init"])
    o.step(10)
    ref = z["pops_10"]
    np.testing.assert_array_equal(o.populations(POP_PHI), ref[0])
    np.testing.assert_array_equal(o.populations(POP_MOM), ref[1])


def test_injection_and_taylor_green_force_bit_exact():
    z, meta = gd.load("inject16")
    o = OracleLattice(gd.config_for(meta))
    o.init_droplets(gd.drop_for(meta), meta["phys"]["p_init"])
    o.step(meta["step_input"] - 1)
    o.inject_momentum(z["g_injected"])
    for n in meta["steps"]:
        _run_to(o, n)
        ref = z[f"fields_{n}"]
        got = o.fields()
        for k, fld in enumerate(FIELDS):
            np.testing.assert_array_equal(got[fld], ref[k], err_msg=f"inject16 step {n} {fld}")


@pytest.mark.slow
@pytest.mark.skipif(not gd.available("walls24"), reason="walls24 fixture not generated")
def test_moving_walls_bit_exact():
    z, meta = gd.load("walls24")
    o = OracleLattice(gd.config_for(meta))
    o.init_droplets(gd.drop_for(meta), meta["phys"]["p_init"])
    n = meta["steps"][0]
    o.step(n)
    got = o.fields()
    np.testing.assert_array_equal(got["phi"], z[f"phi_{n}"])
    np.testing.assert_array_equal(got["p"], z[f"p_{n}"])


@pytest.mark.parametrize("name", ["les16", "les16z4"])
def test_les_injection_bit_exact(name):
    """program/LES_drop.f90: Smagorinsky eddy viscosity from grad u at the end of each
    step, with the values across its slab boundaries (8 x-slabs times 2 or 4 z-slabs)
    and periodic wraps one step old, after a momentum injection and with the
    Taylor-Green force; 500 steps at 16^3."""
    z, meta = gd.load(name)
    o = OracleLattice(gd.config_for(meta))
    o.init_droplets(gd.drop_for(meta), meta["phys"]["p_init"])
    o.step(meta["step_input"] - 1)
    o.inject_momentum(z["g_injected"])
    for n in meta["steps"]:
        _run_to(o, n)
        ref = z[f"fields_{n}"]
        got = o.fields()
        for k, fld in enumerate(FIELDS):
            np.testing.assert_array_equal(got[fld], ref[k], err_msg=f"{name} step {n} {fld}")


def test_les_fixture_pins_the_decomposition():
    """The same run with every neighbour current, or with another z split, misses
    the fixtures: they pin the one-step-old values across slab boundaries."""
    z, meta = gd.load("les16")
    for over in (dict(les_ranks_x=0, les_ranks_y=0, les_ranks_z=0), dict(les_ranks_x=8, les_ranks_y=1, les_ranks_z=1),
                 dict(les_ranks_x=8, les_ranks_y=1, les_ranks_z=4)):
        o = OracleLattice(gd.config_for(meta, **over))
        o.init_droplets(gd.drop_for(meta), meta["phys"]["p_init"])
        o.step(meta["step_input"] - 1)
        o.inject_momentum(z["g_injected"])
        o.step(500 - o.step_count)
        assert np.abs(o.fields()["phi"] - z["fields_500"][0]).max() > 1e-5, over


def test_allen_cahn_plane_and_deformation():
    """Conservative Allen-Cahn model (prototype_program/two-phase/dropac.f90) against
    the fields the compiled program wrote on the plane y = ymax/2 at steps 100, 200
    and 300 with 9 significant digits (its only field output), and against the
    deformation D it wrote every step (break.d): the restatement's break_up3 on the
    same state.  Walls normal to z move from step 50."""
    z, meta = gd.load("ac24")
    o = OracleLattice(gd.config_for(meta))
    o.init_droplets(gd.drop_for(meta), meta["phys"]["p_init"])
    yp = meta["phys"]["plane_y"]
    brk = z["break"]
    for n in meta["steps"]:
        _run_to(o, n)
        f = o.fields()
        ref = z[f"plane_{n}"]   # [z][x][u1, u2, u3, p, phi]
        for k, fld in enumerate(("u1", "u2", "u3", "p", "phi")):
            got = f[fld][:, yp, :]
            want = ref[..., k]
            # the text output rounds to 9 significant digits (es16.8)
            np.testing.assert_allclose(got, want, rtol=6e-9, atol=1e-300, err_msg=f"ac24 step {n} {fld}")
        row = brk[brk[:, 0] == n][0]
        L, B, D = o.deformation(0)
        assert abs(L - row[2]) <= 1e-7 * row[2] and abs(B - row[3]) <= 1e-7 * row[3], (n, L, B, row)
        assert abs(D - row[4]) <= 1e-8, (n, D, row[4])
