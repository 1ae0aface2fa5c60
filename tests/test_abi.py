"""C ABI checks that need no GPU: every entry point declared in include/droplet.h
is exported by libdroplet_amd.so, the struct layouts of the ctypes mirror match
the compiled library, and host-only entry points behave."""
import ctypes as C
import os
import re

import pytest

from droplet_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "droplet.h")
LIB = os.path.join(ROOT, "droplet_amd", "_lib", "libdroplet_amd.so")

pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="libdroplet_amd.so not built")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(drp_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_entry_point_is_exported():
    lib = C.CDLL(LIB)
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts_match():
    lib = C.CDLL(LIB)
    sz = [C.c_size_t() for _ in range(3)]
    ver = lib.drp_abi_sizes(*[C.byref(s) for s in sz])
    assert ver == 3
    assert sz[0].value == C.sizeof(abi.Config)
    assert sz[1].value == C.sizeof(abi.Drop)
    assert sz[2].value == C.sizeof(abi.Stats)


def test_config_default_matches_reference_constants():
    lib = C.CDLL(LIB)
    c = abi.Config()
    assert lib.drp_config_default(C.byref(c)) == 0
    ref = abi.make_config(64)
    for k in ("phi1", "phi2", "a", "b", "T", "kappa_f", "tau_f"):
        assert getattr(c, k) == getattr(ref, k), k


def test_invalid_config_is_rejected_without_gpu_work():
    lib = C.CDLL(LIB)
    lib.drp_last_error.restype = C.c_char_p
    c = abi.make_config(2)
    h = C.c_void_p()
    rc = lib.drp_create(C.byref(c), C.byref(h))
    assert rc == abi.DRP_ERR_ARG and not h.value
    assert b"at least 3" in lib.drp_last_error()
    c = abi.make_config(16, anu=0.1)
    assert lib.drp_create(C.byref(c), C.byref(h)) == abi.DRP_ERR_UNSUPPORTED
    assert lib.drp_step(None, 1) == abi.DRP_ERR_ARG


@pytest.mark.parametrize("kw,msg", [
    (dict(les_ranks_x=8, les_ranks_y=0, les_ranks_z=2), b"all 0 or all positive"),
    (dict(les_ranks_x=3, les_ranks_y=1, les_ranks_z=1), b"do not divide"),
    (dict(les_ranks_x=-1, les_ranks_y=-1, les_ranks_z=-1), b"all 0 or all positive"),
    (dict(advect_k_high=6), b"advect_k_high"),
    (dict(advect_k_high=2, advect_k_low=2), b"advect_k_low"),
    (dict(tau_f=0.5), b"tau_f"),
    (dict(nx=65536, ny=32768, nz=3), b"x-y plane"),
])
def test_config_fields_are_validated_without_gpu_work(kw, msg):
    lib = C.CDLL(LIB)
    lib.drp_last_error.restype = C.c_char_p
    c = abi.make_config(16, les_cs=0.1, les_delta=1.0)
    for k, v in kw.items():
        setattr(c, k, v)
    h = C.c_void_p()
    assert lib.drp_create(C.byref(c), C.byref(h)) == abi.DRP_ERR_ARG and not h.value
    assert msg in lib.drp_last_error()
