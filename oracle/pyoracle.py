"""ctypes wrapper of the CPU restatement (oracle/lbm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from droplet_amd.abi import Config, Drop, Stats

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liblbm_oracle.so")
_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def build_planted(plant: int, out_dir: str) -> str:
    """Compile a copy of the restatement with one planted mistake (see ORACLE_PLANT)."""
    out = os.path.join(out_dir, f"liblbm_oracle_plant{plant}.so")
    subprocess.run(["gcc", "-O2", "-fPIC", "-fopenmp", "-ffp-contract=off", "-std=c11", f"-DORACLE_PLANT={plant}",
                    "-shared", "-o", out, os.path.join(HERE, "lbm_oracle.c"), os.path.join(HERE, "census_oracle.c"),
                    os.path.join(HERE, "budget_oracle.c"), "-lm"], check=True)
    return out


def lib(path: str | None = None):
    """ctypes binding of the restatement (the default build, or the library at `path`)."""
    global _lib
    if path is not None:
        return _bind(C.CDLL(path))
    if _lib is None:
        srcs = [os.path.join(HERE, f) for f in ("lbm_oracle.c", "census_oracle.c", "budget_oracle.c")]
        srcs.append(os.path.join(HERE, "..", "include", "droplet.h"))
        if not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < max(os.path.getmtime(f) for f in srcs):
            build()
        _lib = _bind(C.CDLL(LIB_PATH))
    return _lib


def _bind(L):
    vp, dp = C.c_void_p, C.POINTER(C.c_double)
    L.ora_create.restype = vp
    L.ora_create.argtypes = [C.POINTER(Config)]
    L.ora_destroy.argtypes = [vp]
    L.ora_set_threads.argtypes = [vp, C.c_int]
    L.ora_step_count.restype = C.c_int64
    L.ora_step_count.argtypes = [vp]
    L.ora_init_droplets.argtypes = [vp, C.POINTER(Drop), C.c_int, C.c_double]
    L.ora_init_fields.argtypes = [vp, dp, dp, dp, dp, dp]
    L.ora_step.argtypes = [vp, C.c_int]
    L.ora_get_macro.argtypes = [vp, C.c_int, dp]
    L.ora_get_populations.argtypes = [vp, C.c_int, dp]
    L.ora_set_populations.argtypes = [vp, C.c_int, dp]
    L.ora_inject_momentum.argtypes = [vp, dp]
    L.ora_set_force.argtypes = [vp, dp, dp, dp]
    L.ora_phi_range.argtypes = [vp, dp, dp]
    L.ora_reduce.argtypes = [vp, C.POINTER(Stats)]
    L.ora_deformation.argtypes = [vp, C.c_int, dp, dp, dp]
    L.ora_stencils.argtypes = [vp, dp, dp, dp]
    L.ora_equilibria.argtypes = [vp, dp, dp, dp]
    L.ora_get_tau.argtypes = [vp, dp]
    L.ora_get_force.argtypes = [vp, dp, dp, dp]
    L.ora_spectral_force_fields.argtypes = [vp, dp, dp, dp, dp, dp, dp]
    L.ora_spectral_force_fields.restype = C.c_double
    i64p, i32p = C.POINTER(C.c_int64), C.POINTER(C.c_int32)
    L.ora_census.argtypes = [dp, C.c_int, C.c_int, C.c_int, C.c_double, i64p, i64p, i64p, C.c_int64, i32p]
    L.ora_percolation_count.argtypes = [dp, C.c_int, C.c_int, C.c_int, C.c_double, i64p]
    L.ora_band_pass.argtypes = [vp, dp, dp, dp, C.c_int, C.c_int, dp, dp, dp]
    L.ora_derived.argtypes = [C.POINTER(Config), dp, C.c_int, dp]
    L.ora_budget_terms.argtypes = [C.POINTER(Config), dp, dp, dp, C.c_double, C.c_double, dp]
    return L


def census(phi: np.ndarray, threshold: float):
    """Face-connected periodic clusters of phi >= threshold (phi [z][y][x]):
    (first, volumes, labels) with first / volumes ascending in first."""
    phi = np.ascontiguousarray(phi, np.float64)
    nz, ny, nx = phi.shape
    n = phi.size
    cnt = C.c_int64(0)
    vol = np.zeros(n, np.int64)
    first = np.zeros(n, np.int64)
    lab = np.zeros(phi.shape, np.int32)
    i64 = lambda a: a.ctypes.data_as(C.POINTER(C.c_int64))  # noqa: E731
    rc = lib().ora_census(_dp(phi), nx, ny, nz, threshold, C.byref(cnt), i64(vol), i64(first), n,
                          lab.ctypes.data_as(C.POINTER(C.c_int32)))
    if rc:
        raise MemoryError("ora_census")
    k = cnt.value
    return first[:k].copy(), vol[:k].copy(), lab


def percolation_count(phi: np.ndarray, threshold: float) -> int:
    """program/percolation.f90's droplet count of one field (phi [z][y][x])."""
    phi = np.ascontiguousarray(phi, np.float64)
    nz, ny, nx = phi.shape
    out = C.c_int64(0)
    rc = lib().ora_percolation_count(_dp(phi), nx, ny, nz, threshold, C.byref(out))
    if rc:
        raise RuntimeError("percolation scan needs more than 1000 provisional labels" if rc == -2 else "ora_percolation_count")
    return out.value


def _mac(cfg: Config, fields) -> np.ndarray:
    """phi, p, u1, u2, u3 (each broadcast to [z][y][x]) as one contiguous [5][n] array."""
    shape = (cfg.nz, cfg.ny, cfg.nx)
    return np.ascontiguousarray(np.stack([np.broadcast_to(np.asarray(f, np.float64), shape) for f in fields]))


def derived(cfg: Config, fields, kind: int, libpath: str | None = None) -> np.ndarray:
    """Chemical potential (0), free-energy density (1) or u . grad phi (2) of the fields
    (phi, p, u1, u2, u3), restated (budget_oracle.c)."""
    L = lib(libpath) if libpath else lib()
    m = _mac(cfg, fields)
    out = np.empty(m.shape[1:])
    if L.ora_derived(C.byref(cfg), _dp(m), int(kind), _dp(out)):
        raise ValueError("ora_derived")
    return out


def budget_terms(cfg: Config, old, new, force=None, phi_range=None, libpath: str | None = None) -> np.ndarray:
    """The seven energy-budget terms of one step from the fields it started from (old)
    and produced (new), each (phi, p, u1, u2, u3); force (fx, fy, fz) or None;
    phi_range (min, max) of the new phi, default the clamped extremes (budget_oracle.c)."""
    L = lib(libpath) if libpath else lib()
    mo, mn = _mac(cfg, old), _mac(cfg, new)
    f = None if force is None else _mac(cfg, tuple(force) + (0.0, 0.0))[:3].copy()
    if phi_range is None:
        phi_range = (min(cfg.phi2, float(mn[0].min())), max(cfg.phi1, float(mn[0].max())))
    out = np.zeros(7)
    rc = L.ora_budget_terms(C.byref(cfg), _dp(mo), _dp(mn), None if f is None else _dp(f), float(phi_range[0]),
                            float(phi_range[1]), _dp(out))
    if rc:
        raise ValueError("ora_budget_terms")
    return out


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


class OracleLattice:
    """Same call surface as droplet_amd.Lattice, computed by the CPU restatement."""

    def __init__(self, cfg: Config, threads: int | None = None, libpath: str | None = None):
        self._L = lib(libpath) if libpath else lib()
        self.cfg = cfg
        self.shape = (cfg.nz, cfg.ny, cfg.nx)
        self.n = cfg.nx * cfg.ny * cfg.nz
        self._h = self._L.ora_create(C.byref(cfg))
        if not self._h:
            raise ValueError("ora_create failed")
        if threads is None:
            threads = int(os.environ.get("DROPLET_ORACLE_THREADS", min(8, os.cpu_count() or 1)))
        self._L.ora_set_threads(self._h, threads)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._L.ora_destroy(h)
            self._h = None

    def init_droplets(self, drops, p_init: float = 1.0 / 3.0):
        arr = (Drop * len(drops))(*[Drop(*d) for d in drops])
        self._L.ora_init_droplets(self._h, arr, len(drops), p_init)

    def init_fields(self, phi, p, u1, u2, u3):
        fs = [np.ascontiguousarray(np.broadcast_to(np.asarray(a, np.float64), self.shape)) for a in (phi, p, u1, u2, u3)]
        self._L.ora_init_fields(self._h, *[_dp(a) for a in fs])

    def step(self, n: int = 1):
        self._L.ora_step(self._h, int(n))

    @property
    def step_count(self) -> int:
        return int(self._L.ora_step_count(self._h))

    def macro(self, field: int) -> np.ndarray:
        out = np.empty(self.shape, np.float64)
        if self._L.ora_get_macro(self._h, field, _dp(out)) != 0:
            raise ValueError(field)
        return out

    def fields(self) -> dict:
        return {k: self.macro(i) for i, k in enumerate(("phi", "p", "u1", "u2", "u3"))}

    def populations(self, which: int) -> np.ndarray:
        out = np.empty((15,) + self.shape, np.float64)
        self._L.ora_get_populations(self._h, which, _dp(out))
        return out

    def set_populations(self, which: int, pops: np.ndarray):
        a = np.ascontiguousarray(pops, np.float64)
        assert a.shape == (15,) + self.shape
        self._L.ora_set_populations(self._h, which, _dp(a))

    def inject_momentum(self, g: np.ndarray):
        a = np.ascontiguousarray(g, np.float64)
        assert a.shape == (15,) + self.shape
        self._L.ora_inject_momentum(self._h, _dp(a))

    def set_force(self, fx, fy, fz):
        fs = [np.ascontiguousarray(np.broadcast_to(np.asarray(a, np.float64), self.shape)) for a in (fx, fy, fz)]
        self._L.ora_set_force(self._h, *[_dp(a) for a in fs])

    def force(self):
        out = [np.empty(self.shape, np.float64) for _ in range(3)]
        self._L.ora_get_force(self._h, *[_dp(a) for a in out])
        return out

    def spectral_force(self, u1, u2, u3):
        """(fx, fy, fz, E_low) of the low-shell forcing for the given velocity (config's force_k, force_amp)."""
        us = [np.ascontiguousarray(np.broadcast_to(np.asarray(a, np.float64), self.shape)) for a in (u1, u2, u3)]
        out = [np.empty(self.shape, np.float64) for _ in range(3)]
        elow = self._L.ora_spectral_force_fields(self._h, *[_dp(a) for a in us], *[_dp(a) for a in out])
        return out[0], out[1], out[2], elow

    def band_pass(self, u1, u2, u3, klo: int, khi: int):
        """The restated low_pass (program/lbm_drop_scale_xy.f90:846-903): the shells
        klo <= int|k| < khi of (u1, u2, u3)."""
        us = [np.ascontiguousarray(np.broadcast_to(np.asarray(a, np.float64), self.shape)) for a in (u1, u2, u3)]
        out = [np.empty(self.shape, np.float64) for _ in range(3)]
        if self._L.ora_band_pass(self._h, *[_dp(a) for a in us], int(klo), int(khi), *[_dp(a) for a in out]):
            raise ValueError("ora_band_pass")
        return out

    def phi_range(self):
        mn, mx = C.c_double(), C.c_double()
        self._L.ora_phi_range(self._h, C.byref(mn), C.byref(mx))
        return mn.value, mx.value

    def reduce(self) -> dict:
        st = Stats()
        self._L.ora_reduce(self._h, C.byref(st))
        return st.as_dict()

    def deformation(self, mode: int = 0):
        L, B, D = C.c_double(), C.c_double(), C.c_double()
        rc = self._L.ora_deformation(self._h, mode, C.byref(L), C.byref(B), C.byref(D))
        if rc != 0:
            raise ValueError(f"deformation failed ({rc})")
        return L.value, B.value, D.value

    def stencils(self):
        """(lap phi, grad phi [3], grad u [3][3] with [k][a] = d_a u_k) of the current fields."""
        lap = np.empty(self.shape)
        gp = np.empty((3,) + self.shape)
        gu = np.empty((9,) + self.shape)
        self._L.ora_stencils(self._h, _dp(lap), _dp(gp), _dp(gu))
        return lap, gp, gu.reshape((3, 3) + self.shape)

    def tau(self):
        """tau_g the next step will use (with LES, the one the last step left)."""
        out = np.empty(self.shape)
        self._L.ora_get_tau(self._h, _dp(out))
        return out

    def equilibria(self):
        """(feq [15], geq [15], tau_g) of the current fields."""
        fe = np.empty((15,) + self.shape)
        ge = np.empty((15,) + self.shape)
        tau = np.empty(self.shape)
        self._L.ora_equilibria(self._h, _dp(fe), _dp(ge), _dp(tau))
        return fe, ge, tau
