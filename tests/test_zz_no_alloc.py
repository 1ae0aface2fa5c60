"""drp_reduce_noalloc (the diagnostics reductions with no device allocation)
against drp_reduce, through the C ABI: two calls in a row, the second on a
smaller lattice, must give the same statistics; sizes beyond the stated
maximum and FP32 lattices are refused with an error."""
import ctypes as C

import numpy as np
import pytest

from droplet_amd import abi
from droplet_amd.abi import DRP_FP32_SHIFTED, DRP_WALL_Y, Stats, make_config

pytestmark = pytest.mark.gpu
MAX_NODES = 1 << 25


def _lib():
    from droplet_amd.lattice import lib
    L = lib()
    L.drp_reduce_noalloc.argtypes = [C.c_void_p, C.POINTER(Stats)]
    return L


def _stats(fn, h):
    st = Stats()
    rc = fn(h, C.byref(st))
    assert rc == 0, _lib().drp_last_error()
    return st.as_dict()


def _same(a, b, label):
    for k in a:
        va, vb = np.asarray(a[k], float), np.asarray(b[k], float)
        scale = max(float(np.abs(vb).max()), 1e-300)
        assert float(np.abs(va - vb).max()) <= 1e-10 * scale, f"{label}: {k} {va} vs {vb}"
    return all(np.array_equal(np.asarray(a[k]), np.asarray(b[k])) for k in a)


def test_reduce_noalloc_matches_reduce_twice():
    from droplet_amd.lattice import Lattice
    L = _lib()
    exact = []
    for N, steps in ((96, 30), (40, 12)):   # the second call on a smaller lattice
        cfg = make_config(N, kappa_g=0.005, nu2=0.2, wall_axis=DRP_WALL_Y, wall_u=0.02)
        lat = Lattice(cfg)
        lat.init_droplets([(0.5 * (N - 1), 0.5 * (N - 1) + 1.0, 0.5 * (N - 1), 0.45 * N)])
        lat.step(steps)
        got = _stats(L.drp_reduce_noalloc, lat._h)
        ref = _stats(L.drp_reduce, lat._h)
        exact.append(_same(got, ref, f"{N}^3"))
        assert got["interface_area"] > 0 and got["kinetic"] > 0
    print(f"bit-identical to drp_reduce: {exact}")


def test_reduce_noalloc_refuses_what_it_cannot_do():
    from droplet_amd.lattice import Lattice
    L = _lib()
    L.drp_last_error.restype = C.c_char_p
    lat = Lattice(make_config(16, precision=DRP_FP32_SHIFTED))
    lat.init_droplets([(7.5, 7.5, 7.5, 8.0)])
    assert L.drp_reduce_noalloc(lat._h, C.byref(Stats())) == abi.DRP_ERR_UNSUPPORTED
    del lat
    big = Lattice(make_config(338, 320, 320))   # 34.6 M nodes > 2^25, not initialised
    assert big.n > MAX_NODES
    assert L.drp_reduce_noalloc(big._h, C.byref(Stats())) == abi.DRP_ERR_ARG
    assert b"exceed its maximum" in L.drp_last_error()
    assert L.drp_reduce_noalloc(None, C.byref(Stats())) == abi.DRP_ERR_ARG
