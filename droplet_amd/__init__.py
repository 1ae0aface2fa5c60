"""droplet_amd — MI355X-native D3Q15 two-phase lattice Boltzmann step (HIP, gfx950).

The compute path is libdroplet_amd.so (droplet_amd/csrc), reached through the C ABI
of include/droplet.h; this package is its Python front-end.
"""
from .abi import (Config, Drop, Stats, make_config, kappa_g_from_sigma, sigma_from_kappa_g, nu_from_tau,
                  taylor_green_amplitude, DRP_FP64, DRP_FP32_SHIFTED, DRP_WALL_NONE, DRP_WALL_Y, DRP_WALL_Z,
                  DRP_FORCE_NONE, DRP_FORCE_TAYLOR_GREEN, DRP_FORCE_FIELD, POP_PHI, POP_MOM, FIELDS)

__all__ = ["Config", "Drop", "Stats", "make_config", "kappa_g_from_sigma", "sigma_from_kappa_g", "nu_from_tau",
           "taylor_green_amplitude", "DRP_FP64", "DRP_FP32_SHIFTED", "DRP_WALL_NONE", "DRP_WALL_Y", "DRP_WALL_Z",
           "DRP_FORCE_NONE", "DRP_FORCE_TAYLOR_GREEN", "DRP_FORCE_FIELD", "POP_PHI", "POP_MOM", "FIELDS",
           "Lattice"]


def __getattr__(name):
    if name == "Lattice":
        from .lattice import Lattice
        return Lattice
    raise AttributeError(name)
