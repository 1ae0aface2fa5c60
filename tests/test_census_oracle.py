"""Droplet census: the CPU restatements against the compiled reference and an
independent labelling.

- ora_percolation_count (program/percolation.f90's own scan) reproduces the
  droplet counts the compiled reference wrote for the 71 fields of the dsd32
  fixture (tests/golden/make_golden.py case_dsd).
- ora_census (face neighbours, periodic in x, y, z: what the GPU census
  computes) is checked against scipy.ndimage.label on the open box with the
  labels joined across the three pairs of opposite faces afterwards.
- Where the two definitions meet: on fixture fields whose droplets stay clear
  of the periodic seams the counts agree; every disagreement involves a droplet
  that crosses a seam (the reference follows wraps through only some of its
  neighbour links).
"""
import os
import sys

import numpy as np
import pytest
from scipy import ndimage

from oracle.pyoracle import census, percolation_count
from tests import golden_data as gd

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
from make_golden import sphere_field  # noqa: E402


def _fixture():
    z, meta = gd.load("dsd32")
    thr = 0.5 * (meta["phi1"] + meta["phi2"])
    fields = [sphere_field(meta["N"], z["drops"][k, :z["ndrops"][k]]) for k in range(len(meta["steps"]))]
    return z, meta, thr, fields


def scipy_periodic(mask: np.ndarray):
    """Face-connected periodic clusters by scipy on the open box, then a
    union-find over the labels that meet across opposite faces: (first, volumes)
    with first the smallest flat index [z][y][x] of each cluster."""
    lab, n = ndimage.label(mask, ndimage.generate_binary_structure(3, 1))
    par = list(range(n + 1))

    def root(a):
        while par[a] != a:
            par[a] = par[par[a]]
            a = par[a]
        return a

    for ax in range(3):
        a = np.take(lab, 0, axis=ax).ravel()
        b = np.take(lab, -1, axis=ax).ravel()
        for p, q in zip(a, b):
            if p and q:
                rp, rq = root(p), root(q)
                if rp != rq:
                    par[max(rp, rq)] = min(rp, rq)
    roots = np.array([root(i) for i in range(n + 1)])
    flat = roots[lab].ravel()
    idx = np.flatnonzero(flat)
    out = {}
    for i in idx:
        r = flat[i]
        f, v = out.get(r, (i, 0))
        out[r] = (min(f, i), v + 1)
    items = sorted(out.values())
    return np.array([f for f, _ in items], np.int64), np.array([v for _, v in items], np.int64)


def test_percolation_restatement_matches_reference():
    z, meta, thr, fields = _fixture()
    got = [percolation_count(phi, thr) for phi in fields]
    np.testing.assert_array_equal(got, z["counts"])


@pytest.mark.parametrize("shape,seed,level", [((16, 16, 16), 1, 1.2), ((9, 13, 17), 2, 0.6), ((24, 20, 12), 3, 1.2),
                                              ((32, 32, 32), 4, 0.9), ((32, 32, 32), 4, 0.3)])
def test_census_matches_scipy_periodic(shape, seed, level):
    rng = np.random.default_rng(seed)
    noise = ndimage.gaussian_filter(rng.standard_normal(shape), 1.2, mode="wrap")
    phi = np.where(noise > level * noise.std(), 4.895, 2.211)
    f, v, lab = census(phi, 3.553)
    sf, sv = scipy_periodic(phi >= 3.553)
    assert len(f) >= 2
    np.testing.assert_array_equal(f, sf)
    np.testing.assert_array_equal(v, sv)
    # labels: each node carries its cluster's first index
    m = phi >= 3.553
    assert (lab[~m] == -1).all()
    assert set(np.unique(lab[m])) == set(f)
    assert v.sum() == m.sum()


def test_census_special_fields():
    shape = (6, 8, 10)
    empty = np.full(shape, 2.211)
    f, v, _ = census(empty, 3.553)
    assert len(f) == 0
    full = np.full(shape, 4.895)
    f, v, _ = census(full, 3.553)
    assert list(f) == [0] and list(v) == [full.size]
    z, y, x = np.indices(shape)
    checker = np.where((x + y + z) % 2 == 0, 4.895, 2.211)   # even sides: no two droplet nodes touch
    f, v, _ = census(checker, 3.553)
    assert len(f) == checker.size // 2 and (v == 1).all()


def test_census_is_periodic_in_every_direction():
    """One drop split across each pair of faces is one cluster; an open-box
    labelling (the planted mistake) counts it as two."""
    N = 12
    for ax in range(3):
        phi = np.full((N, N, N), 2.211)
        sl = [slice(4, 7)] * 3
        sl[ax] = slice(0, 2)
        phi[tuple(sl)] = 4.895
        sl[ax] = slice(N - 2, N)
        phi[tuple(sl)] = 4.895
        f, v, _ = census(phi, 3.553)
        assert len(f) == 1 and v[0] == 2 * 2 * 9
        assert ndimage.label(phi >= 3.553)[1] == 2


def test_census_vs_reference_counts():
    z, meta, thr, fields = _fixture()
    N = meta["N"]
    agree_off_seam = 0
    off_seam = 0
    for k, phi in enumerate(fields):
        d = z["drops"][k, :z["ndrops"][k]]
        crosses = bool(((d[:, :3] - d[:, 3:4]) < 0).any() or ((d[:, :3] + d[:, 3:4]) > N - 1).any())
        n = len(census(phi, thr)[0])
        if not crosses:
            off_seam += 1
            agree_off_seam += n == z["counts"][k]
    assert off_seam >= 5 and agree_off_seam == off_seam
