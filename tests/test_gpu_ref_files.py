"""The reference's per-rank files through the C ABI (drp_write/read_reference_files).

The layouts are checked with the readers and writers that tests/golden/make_golden.py
uses on the reference programs' own files (tests/golden/fortran_io.py):
- `output` files written by the library assemble to the lattice's fields exactly;
- `outputfg` files round-trip, and a restart from them continues bit-identically
  (step counter and phi range restored);
- the golden injection case run from g-input files written like the reference's
  input reproduces the reference's fields bit for bit;
- program/LES_drop.f90's x-z decomposition: output files assemble like the
  reference's, and the LES golden case run from (f, g) input files in that
  layout (g kept, as LES_drop.f90 reads them) reproduces its fields bit for bit.
"""
import os
import sys

import numpy as np
import pytest

from droplet_amd.abi import (DRP_FILE_FG, DRP_FILE_FG_G, DRP_FILE_FIELDS, DRP_FILE_G, DRP_FILE_LAYOUT_XZ, DRP_WALL_Y,
                             POP_MOM, POP_PHI, make_config)
from tests import golden_data as gd

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import fortran_io as fio  # noqa: E402

pytestmark = pytest.mark.gpu


def _dl():
    from droplet_amd import lattice as dl
    return dl


def _assemble_fields_2d(d, n, shape, rx, ry):
    nz, ny, nx = shape
    xp, yp = nx // rx, ny // ry
    out = np.zeros((5, nz, ny, nx))
    for r in range(rx * ry):
        x0, y0 = (r // ry) * xp, (r % ry) * yp
        a = fio.read_records(os.path.join(d, f"0_{r}_{n}.bin"), 5).reshape(nz, yp, xp, 5)
        out[:, :, y0:y0 + yp, x0:x0 + xp] = np.moveaxis(a, -1, 0)
    return out[[0, 4, 1, 2, 3]]


def test_output_files_layout(tmp_path):
    dl = _dl()
    N = 16
    lat = dl.Lattice(make_config(N, kappa_g=0.01, wall_axis=DRP_WALL_Y, wall_u=0.01))
    lat.init_droplets([(7.5, 7.5, 7.5, 8.0)])
    lat.step(12)
    f = lat.fields()
    lat.write_reference_files(str(tmp_path), 12, DRP_FILE_FIELDS, 8, 1)
    got = fio.assemble_periodic(str(tmp_path), 12, N, 8)
    for k, name in enumerate(("phi", "p", "u1", "u2", "u3")):
        assert np.array_equal(got[k], f[name])
    # a 2 x 4 split of a non-cubic grid
    lat2 = dl.Lattice(make_config(16, 12, 8, kappa_g=0.01))
    lat2.init_droplets([(7.5, 5.5, 3.5, 6.0)])
    lat2.step(3)
    f2 = lat2.fields()
    d2 = tmp_path / "split"
    d2.mkdir()
    lat2.write_reference_files(str(d2), 3, DRP_FILE_FIELDS, 2, 4)
    got2 = _assemble_fields_2d(str(d2), 3, (8, 12, 16), 2, 4)
    for k, name in enumerate(("phi", "p", "u1", "u2", "u3")):
        assert np.array_equal(got2[k], f2[name])


@pytest.mark.parametrize("les", [False, True], ids=["nu_phi", "les"])
def test_outputfg_restart_is_bit_identical(tmp_path, les):
    """A restart continues bit for bit: populations, step counter and phi range
    are restored; with LES (all neighbours current) the restored state's tau
    equals the one the continuing run computed at the end of its step."""
    dl = _dl()
    N = 16
    kw = dict(les_cs=0.15, les_delta=2.0) if les else {}
    cfg = make_config(N, kappa_g=0.01, nu2=0.2, wall_axis=DRP_WALL_Y, wall_u=0.01, **kw)   # nu(phi): the range matters
    a = dl.Lattice(cfg)
    a.init_droplets([(7.5, 7.5, 7.5, 8.0)])
    a.step(20)
    a.write_reference_files(str(tmp_path), 20, DRP_FILE_FG, 8, 1)
    pops = fio.assemble_pops(str(tmp_path), 20, N, 8)
    assert np.array_equal(pops[0], a.populations(POP_PHI)) and np.array_equal(pops[1], a.populations(POP_MOM))
    b = dl.Lattice(cfg)
    b.init_droplets([(3.0, 3.0, 3.0, 2.0)])          # anything: the files replace the state
    b.read_reference_files(str(tmp_path), 20, DRP_FILE_FG, 8, 1)
    assert b.step_count == 20
    a.step(15)
    b.step(15)
    fa, fb = a.fields(), b.fields()
    for k in fa:
        assert np.array_equal(fa[k], fb[k]), k


def test_restart_with_les_rank_blocks_matches_the_restatement(tmp_path):
    """With LES rank blocks the continuing run takes grad u across block faces
    from one-step-old halos, while a restart forms the restored state's tau with
    every neighbour current (as ini_op does): the restarted run differs from the
    continuing one, and equals the restatement restarted from the same files."""
    from oracle.pyoracle import OracleLattice
    dl = _dl()
    N = 16
    # equal viscosities: the restored phi range does not enter tau
    cfg = make_config(N, kappa_g=0.01, les_cs=0.15, les_delta=2.0, les_ranks_x=2, les_ranks_y=1, les_ranks_z=2)
    a = dl.Lattice(cfg)
    a.init_droplets([(7.5, 7.5, 7.5, 8.0)])
    a.step(20)
    a.write_reference_files(str(tmp_path), 20, DRP_FILE_FG, 8, 1)
    b = dl.Lattice(cfg)
    b.init_droplets([(3.0, 3.0, 3.0, 2.0)])
    b.read_reference_files(str(tmp_path), 20, DRP_FILE_FG, 8, 1)
    o = OracleLattice(cfg)
    o.set_populations(POP_PHI, a.populations(POP_PHI))
    o.set_populations(POP_MOM, a.populations(POP_MOM))
    a.step(5)
    b.step(5)
    o.step(5)
    fa, fb, fo = a.fields(), b.fields(), o.fields()
    assert not np.array_equal(fa["u1"], fb["u1"])        # the documented divergence
    for k in fb:
        assert np.array_equal(fb[k], fo[k]), k


def test_golden_injection_from_reference_input_files(tmp_path):
    dl = _dl()
    z, meta = gd.load("inject16")
    fio.write_g_input(str(tmp_path), 7, z["g_injected"], 8)
    lat = dl.Lattice(gd.config_for(meta))
    lat.init_droplets(gd.drop_for(meta), meta["phys"]["p_init"])
    lat.step(meta["step_input"] - 1)
    lat.read_reference_files(str(tmp_path), 7, DRP_FILE_G, 8, 1)
    for n in meta["steps"]:
        lat.step(n - lat.step_count)
        ref = z[f"fields_{n}"]
        got = lat.fields()
        for k, name in enumerate(("phi", "p", "u1", "u2", "u3")):
            assert np.array_equal(got[name], ref[k]), (n, name)


def test_file_errors(tmp_path):
    dl = _dl()
    lat = dl.Lattice(make_config(16, kappa_g=0.01))
    lat.init_droplets([(7.5, 7.5, 7.5, 8.0)])
    with pytest.raises(dl.DropletError):
        lat.read_reference_files(str(tmp_path), 5, DRP_FILE_FG, 8, 1)       # no files
    with pytest.raises(dl.DropletError):
        lat.write_reference_files(str(tmp_path), 5, DRP_FILE_FIELDS, 3, 1)  # 3 does not divide 16
    lat.write_reference_files(str(tmp_path), 5, DRP_FILE_FG, 8, 1)
    with pytest.raises(dl.DropletError):
        lat.read_reference_files(str(tmp_path), 5, DRP_FILE_FG, 4, 1)       # wrong split: sizes differ
    with pytest.raises(dl.DropletError):
        lat.read_reference_files(str(tmp_path), 5, DRP_FILE_G, 8, 1)        # (f, g) records read as g-only


def test_xz_layout_output_files(tmp_path):
    dl = _dl()
    N = 16
    lat = dl.Lattice(make_config(N, kappa_g=0.01))
    lat.init_droplets([(7.5, 7.5, 7.5, 8.0)])
    lat.step(5)
    f = lat.fields()
    for rx, rz in ((8, 2), (4, 4)):
        d = tmp_path / f"xz{rx}{rz}"
        d.mkdir()
        lat.write_reference_files(str(d), 5, DRP_FILE_FIELDS | DRP_FILE_LAYOUT_XZ, rx, rz)
        got = fio.assemble_xz(str(d), 5, N, rx, rz)
        for k, name in enumerate(("phi", "p", "u1", "u2", "u3")):
            assert np.array_equal(got[k], f[name]), (rx, rz, name)


def test_les_golden_from_reference_fg_input(tmp_path):
    dl = _dl()
    z, meta = gd.load("les16")
    g = z["g_injected"]
    fio.write_fg_input(str(tmp_path), 7, np.full_like(g, 12345.0), g, 8, 2)   # f values are read and dropped
    lat = dl.Lattice(gd.config_for(meta))
    lat.init_droplets(gd.drop_for(meta), meta["phys"]["p_init"])
    lat.step(meta["step_input"] - 1)
    lat.read_reference_files(str(tmp_path), 7, DRP_FILE_FG_G | DRP_FILE_LAYOUT_XZ, 8, 2)
    n = meta["steps"][0]
    lat.step(n - lat.step_count)
    got = lat.fields()
    for k, name in enumerate(("phi", "p", "u1", "u2", "u3")):
        assert np.array_equal(got[name], z[f"fields_{n}"][k]), name
    with pytest.raises(dl.DropletError):   # a 4 x 2 split needs 512 nodes per file, these hold 256
        lat.read_reference_files(str(tmp_path), 7, DRP_FILE_FG_G | DRP_FILE_LAYOUT_XZ, 4, 2)
