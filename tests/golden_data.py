"""Loader for the golden-vector fixtures in tests/golden/ (numbers produced by the
compiled reference, see tests/golden/make_golden.py)."""
from __future__ import annotations

import json
import os

import numpy as np

from droplet_amd.abi import DRP_FORCE_TAYLOR_GREEN, DRP_MODEL_ALLEN_CAHN, DRP_WALL_Z, make_config

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name: str):
    path = os.path.join(GOLDEN, f"{name}.npz")
    z = np.load(path, allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    return z, meta


def available(name: str) -> bool:
    return os.path.exists(os.path.join(GOLDEN, f"{name}.npz"))


def config_for(meta: dict, **over):
    """drp_config reproducing the reference parameter set of a golden case."""
    ph = meta["phys"]
    kw = dict(nu1=ph["nu1"], nu2=ph["nu2"], kappa_g=ph["kappa_g"], tau_f=ph.get("tau_f", 0.7),
              kappa_f=ph.get("kappa_f", 0.01))
    if meta["case"].startswith(("inject", "les")):
        kw.update(force_kind=DRP_FORCE_TAYLOR_GREEN, force_start=meta["step_input"],
                  force_amp=ph["force_amp"], force_k=ph["force_k"])
    if "les_cs" in ph:
        kw.update(les_cs=ph["les_cs"], les_delta=ph["les_delta"], les_ranks_x=ph["les_ranks_x"],
                  les_ranks_y=ph["les_ranks_y"], les_ranks_z=ph["les_ranks_z"])
    if "wall_u" in ph:
        kw.update(wall_axis=DRP_WALL_Z, wall_u=ph["wall_u"], wall_start=ph["wall_start"])
    if "ac_a" in ph:
        kw = dict(nu1=ph["nu1"], nu2=ph["nu2"], kappa_g=ph["kappa_g"], model=DRP_MODEL_ALLEN_CAHN, ac_a=ph["ac_a"],
                  ac_width=ph["ac_width"], phi1=ph["phi1"], phi2=ph["phi2"], wall_axis=DRP_WALL_Z,
                  wall_u=ph["wall_u"], wall_start=ph["wall_start"])
    kw.update(over)
    return make_config(ph["nx"], ph["ny"], ph["nz"], **kw)


def drop_for(meta: dict):
    return [tuple(meta["phys"]["drop"])]


def max_rel(a: np.ndarray, b: np.ndarray, floor: float = 0.0) -> float:
    """max |a-b| / max(|b|_inf, floor): relative L-infinity error."""
    scale = max(float(np.abs(b).max()), floor)
    return float(np.abs(a - b).max() / scale) if scale > 0 else float(np.abs(a - b).max())
