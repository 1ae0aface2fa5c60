"""ctypes mirror of include/droplet.h (structs, constants) and config helpers.

The helpers restate how the reference derives its constants from the
dimensionless numbers in its ``parameter`` block, so a Python caller can build
the same configuration the Fortran program compiles in.
"""
from __future__ import annotations

import ctypes as C
import math

DRP_OK = 0
DRP_ERR_ARG = -1
DRP_ERR_HIP = -2
DRP_ERR_STATE = -3
DRP_ERR_UNSUPPORTED = -4
DRP_ERR_ALLOC = -5

DRP_FP64 = 0
DRP_FP32_SHIFTED = 1

DRP_WALL_NONE = 0
DRP_WALL_Y = 2
DRP_WALL_Z = 3

DRP_FORCE_NONE = 0
DRP_FORCE_TAYLOR_GREEN = 1
DRP_FORCE_FIELD = 2
DRP_FORCE_SPECTRAL = 3
DRP_MODEL_FREE_ENERGY = 0
DRP_MODEL_ALLEN_CAHN = 1
DRP_LES_STRAIN = 0
DRP_LES_NONEQ = 1
DRP_DERIVED_CHEMICAL_POTENTIAL = 0
DRP_DERIVED_FREE_ENERGY = 1
DRP_DERIVED_ADVECTION = 2
DRP_FILE_FIELDS, DRP_FILE_FG, DRP_FILE_G, DRP_FILE_FG_G = 0, 1, 2, 3
DRP_FILE_LAYOUT_XZ = 0x100

FIELD_PHI, FIELD_P, FIELD_U1, FIELD_U2, FIELD_U3 = range(5)
FIELDS = {"phi": FIELD_PHI, "p": FIELD_P, "u1": FIELD_U1, "u2": FIELD_U2, "u3": FIELD_U3}
POP_PHI = 0  # reference f
POP_MOM = 1  # reference g

# D3Q15 velocity set, program/lbm_drop_xy.f90:70-84 (q = 0..14 <-> reference i = 1..15)
CX = (0, 1, 0, 0, -1, 0, 0, 1, -1, 1, 1, -1, 1, -1, -1)
CY = (0, 0, 1, 0, 0, -1, 0, 1, 1, -1, 1, -1, -1, 1, -1)
CZ = (0, 0, 0, 1, 0, 0, -1, 1, 1, 1, -1, -1, -1, -1, 1)
E = (2.0 / 9.0,) + (1.0 / 9.0,) * 6 + (1.0 / 72.0,) * 8
H = (1.0,) + (0.0,) * 14
F = (-7.0 / 3.0,) + (1.0 / 3.0,) * 6 + (1.0 / 24.0,) * 8
OPP = (0, 4, 5, 6, 1, 2, 3, 11, 12, 13, 14, 7, 8, 9, 10)


class Config(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("precision", C.c_int32),
        ("tau_f", C.c_double), ("nu1", C.c_double), ("nu2", C.c_double),
        ("kappa_f", C.c_double), ("kappa_g", C.c_double),
        ("phi1", C.c_double), ("phi2", C.c_double),
        ("a", C.c_double), ("b", C.c_double), ("T", C.c_double), ("anu", C.c_double),
        ("wall_axis", C.c_int32), ("wall_start", C.c_int32), ("wall_u", C.c_double),
        ("force_kind", C.c_int32), ("force_start", C.c_int32),
        ("force_amp", C.c_double), ("force_k", C.c_double),
        ("les_cs", C.c_double), ("les_delta", C.c_double),
        ("les_ranks_x", C.c_int32), ("les_ranks_y", C.c_int32), ("les_ranks_z", C.c_int32),
        ("model", C.c_int32), ("les_model", C.c_int32),
        ("advect_k_high", C.c_int32), ("advect_k_low", C.c_int32), ("advect_start", C.c_int32),
        ("ac_a", C.c_double), ("ac_width", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class Drop(C.Structure):
    _fields_ = [("xc", C.c_double), ("yc", C.c_double), ("zc", C.c_double), ("diameter", C.c_double)]


class Stats(C.Structure):
    _fields_ = [
        ("phi_mass", C.c_double), ("phi_min", C.c_double), ("phi_max", C.c_double),
        ("kinetic", C.c_double), ("dissipation", C.c_double), ("interface_area", C.c_double),
        ("free_energy", C.c_double), ("p_mean", C.c_double), ("momentum", C.c_double * 3),
        ("interaction", C.c_double),
    ]

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "momentum"}
        d["momentum"] = tuple(self.momentum)
        return d


def kappa_g_from_sigma(sigma: float) -> float:
    """Surface-tension calibration sigma = 1.7039 kappa_g^0.9991 (program/lbm_drop_xy.f90:51)."""
    return (sigma / 1.7039) ** (1.0 / 0.9991)


def sigma_from_kappa_g(kappa_g: float) -> float:
    return 1.7039 * kappa_g ** 0.9991


def nu_from_tau(tau_g: float) -> float:
    """tau_g = 1/2 + 3 nu / ds (program/lbm_drop_xy.f90:582)."""
    return (tau_g - 0.5) / 3.0


def taylor_green_amplitude(umax: float, d_vortex: float, re: float) -> float:
    """2 pi^2 U^2 / (D_vortex Re), evaluated left to right as in program/lbm_drop_xy.f90:682-690."""
    pi = math.acos(-1.0)
    return 2.0 * pi * pi * umax * umax / (d_vortex * re)


def make_config(nx: int, ny: int | None = None, nz: int | None = None, **kw) -> Config:
    """Config with the reference defaults (program/lbm_drop_xy.f90:51-67), overridable by keyword."""
    c = Config()
    c.nx = nx
    c.ny = nx if ny is None else ny
    c.nz = nx if nz is None else nz
    c.precision = DRP_FP64
    c.tau_f = 0.7
    c.nu1 = c.nu2 = 0.1
    c.kappa_f = 0.01
    c.kappa_g = 0.0
    c.phi1, c.phi2 = 2.211, 4.895
    c.a, c.b, c.T = 9.0 / 49.0, 2.0 / 21.0, 0.55
    c.anu = 0.0
    c.wall_axis = DRP_WALL_NONE
    c.wall_start = 1
    c.wall_u = 0.0
    c.force_kind = DRP_FORCE_NONE
    c.force_start = 1
    c.force_amp = 0.0
    c.force_k = 0.0
    c.les_cs = 0.0
    c.les_delta = 0.0
    for k, v in kw.items():
        if not hasattr(c, k):
            raise KeyError(f"unknown config field {k!r}")
        setattr(c, k, v)
    return c
