"""Python front-end of the HIP library (ctypes over include/droplet.h).

Mirrors the reference driver's call surface: construct (par + allocation),
init_droplets / init_fields (ini + ini_op), step (loop body), fields /
populations (output / outputfg), inject_momentum (step_input injection),
reduce and deformation (analysis hooks).

The library must be built first (``python -m droplet_amd.build``); importing
this module on a machine without it raises instead of falling back to any CPU
path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .abi import Config, Drop, Stats, POP_MOM, POP_PHI, FIELDS  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libdroplet_amd.so")
_lib = None


class DropletError(RuntimeError):
    def __init__(self, code: int, where: str):
        msg = lib().drp_last_error().decode(errors="replace")
        super().__init__(f"{where} failed ({code}): {msg}")
        self.code = code


def lib():
    """Load libdroplet_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not built; run `python -m droplet_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp, dp, sz = C.c_void_p, C.POINTER(C.c_double), C.c_size_t
        L.drp_last_error.restype = C.c_char_p
        L.drp_config_default.argtypes = [C.POINTER(Config)]
        L.drp_create.argtypes = [C.POINTER(Config), C.POINTER(vp)]
        L.drp_destroy.argtypes = [vp]
        L.drp_destroy.restype = None
        L.drp_init_droplets.argtypes = [vp, C.POINTER(Drop), C.c_int, C.c_double]
        L.drp_init_fields.argtypes = [vp, dp, dp, dp, dp, dp]
        L.drp_init_droplets_tanh.argtypes = [vp, C.POINTER(Drop), C.c_int, C.c_double, C.c_double]
        L.drp_step.argtypes = [vp, C.c_int]
        L.drp_sync.argtypes = [vp]
        L.drp_step_count.argtypes = [vp]
        L.drp_step_count.restype = C.c_int64
        L.drp_get_macro.argtypes = [vp, C.c_int, dp, sz]
        L.drp_get_populations.argtypes = [vp, C.c_int, dp, sz]
        L.drp_set_populations.argtypes = [vp, C.c_int, dp, sz]
        L.drp_inject_momentum.argtypes = [vp, dp, sz]
        L.drp_set_force.argtypes = [vp, dp, dp, dp, sz]
        L.drp_reduce.argtypes = [vp, C.POINTER(Stats)]
        L.drp_get_force.argtypes = [vp, dp, dp, dp, sz]
        L.drp_write_reference_files.argtypes = [vp, C.c_char_p, C.c_int64, C.c_int, C.c_int, C.c_int]
        L.drp_read_reference_files.argtypes = [vp, C.c_char_p, C.c_int64, C.c_int, C.c_int, C.c_int]
        L.drp_deformation.argtypes = [vp, C.c_int, dp, dp, dp]
        L.drp_time_steps.argtypes = [vp, C.c_int, dp]
        L.drp_tile_shape_count.argtypes = []
        L.drp_set_tile_shape.argtypes = [vp, C.c_int]
        L.drp_get_tile_shape.argtypes = [vp]
        L.drp_autotune.argtypes = [C.POINTER(Config), C.c_int, C.POINTER(C.c_int), dp]
        L.drp_copy_bandwidth.argtypes = [C.c_size_t, C.c_int, dp]
        L.drp_copy_bandwidth_soa.argtypes = [C.c_size_t, C.c_int, C.c_int, dp]
        i64p = C.POINTER(C.c_int64)
        L.drp_droplet_census.argtypes = [vp, C.c_double, i64p, i64p, i64p, C.c_int64, C.POINTER(C.c_int32),
                                         C.c_size_t]
        L.drp_batch_create.argtypes = [C.POINTER(Config), C.c_int, C.POINTER(vp)]
        L.drp_batch_size.argtypes = [vp]
        L.drp_batch_member.argtypes = [vp, C.c_int]
        L.drp_batch_member.restype = vp
        L.drp_batch_step.argtypes = [vp, C.c_int]
        L.drp_batch_use_graph.argtypes = [vp, C.c_int]
        L.drp_batch_set_tile_shape.argtypes = [vp, C.c_int]
        L.drp_batch_sync.argtypes = [vp]
        L.drp_batch_time_steps.argtypes = [vp, C.c_int, dp]
        L.drp_batch_autotune.argtypes = [C.POINTER(Config), C.c_int, C.c_int, C.POINTER(C.c_int), dp]
        L.drp_copy_bandwidth_kind.argtypes = [C.c_size_t, C.c_int, C.c_int, dp]
        L.drp_energy_spectrum.argtypes = [vp, dp, C.c_int]
        L.drp_scale_interaction.argtypes = [vp, dp, C.c_int]
        L.drp_filter_velocity.argtypes = [vp, C.c_int, C.c_int, dp, dp, dp, sz]
        L.drp_fft3d.argtypes = [C.c_int, C.c_int, C.c_int, dp, C.c_int]
        L.drp_get_derived.argtypes = [vp, C.c_int, dp, sz]
        L.drp_phi_contribution.argtypes = [vp, dp, C.c_int, dp]
        L.drp_step_energy_budget.argtypes = [vp, dp]
        L.drp_droplet_size_histogram.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_double,
                                                 C.POINTER(C.c_int64), C.c_int64, C.POINTER(C.c_int64)]
        L.drp_batch_destroy.argtypes = [vp]
        L.drp_batch_destroy.restype = None
        L.drp_tape_create.argtypes = [vp, C.c_int, C.POINTER(vp)]
        L.drp_tape_step.argtypes = [vp, C.c_int]
        L.drp_tape_loss.argtypes = [vp, dp, dp, dp, dp]
        L.drp_tape_backward.argtypes = [vp, dp, dp, dp]
        L.drp_tape_bytes.argtypes = [vp]
        L.drp_tape_bytes.restype = C.c_int64
        L.drp_tape_destroy.argtypes = [vp]
        L.drp_tape_destroy.restype = None
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _check(rc: int, where: str):
    if rc != 0:
        raise DropletError(rc, where)


class Lattice:
    """One D3Q15 two-phase lattice resident on the GPU."""

    def __init__(self, cfg: Config):
        self.cfg = cfg
        self.shape = (cfg.nz, cfg.ny, cfg.nx)
        self.n = cfg.nx * cfg.ny * cfg.nz
        h = C.c_void_p()
        _check(lib().drp_create(C.byref(cfg), C.byref(h)), "drp_create")
        self._h = h

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and not getattr(self, "_borrowed", False):
            lib().drp_destroy(h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- init (ini + ini_op)
    def init_droplets(self, drops, p_init: float = 1.0 / 3.0):
        arr = (Drop * max(1, len(drops)))(*[Drop(*d) for d in drops])
        _check(lib().drp_init_droplets(self._h, arr, len(drops), p_init), "drp_init_droplets")

    def init_droplets_tanh(self, drops, width: float, p_init: float = 1.0 / 3.0):
        """Diffuse-interface droplets, phi = mid + half tanh((D/2 - r)/(sqrt2 width))."""
        arr = (Drop * len(drops))(*[Drop(*d) for d in drops])
        _check(lib().drp_init_droplets_tanh(self._h, arr, len(drops), p_init, width), "drp_init_droplets_tanh")

    def init_fields(self, phi, p, u1, u2, u3):
        fs = [np.ascontiguousarray(np.broadcast_to(np.asarray(a, np.float64), self.shape)) for a in (phi, p, u1, u2, u3)]
        _check(lib().drp_init_fields(self._h, *[_dp(a) for a in fs]), "drp_init_fields")

    # -- time stepping
    def step(self, n: int = 1):
        _check(lib().drp_step(self._h, int(n)), "drp_step")

    def sync(self):
        _check(lib().drp_sync(self._h), "drp_sync")

    def time_steps(self, n: int) -> float:
        """Run n steps between HIP events on the handle's stream; returns milliseconds."""
        ms = C.c_double()
        _check(lib().drp_time_steps(self._h, int(n), C.byref(ms)), "drp_time_steps")
        return ms.value

    @property
    def tile_shape(self) -> int:
        return int(lib().drp_get_tile_shape(self._h))

    @tile_shape.setter
    def tile_shape(self, shape: int):
        _check(lib().drp_set_tile_shape(self._h, int(shape)), "drp_set_tile_shape")

    @property
    def step_count(self) -> int:
        return int(lib().drp_step_count(self._h))

    # -- output hooks
    def macro(self, field: int) -> np.ndarray:
        out = np.empty(self.shape, np.float64)
        _check(lib().drp_get_macro(self._h, int(field), _dp(out), out.size), "drp_get_macro")
        return out

    def fields(self) -> dict:
        return {k: self.macro(i) for k, i in FIELDS.items()}

    def populations(self, which: int) -> np.ndarray:
        out = np.empty((15,) + self.shape, np.float64)
        _check(lib().drp_get_populations(self._h, int(which), _dp(out), out.size), "drp_get_populations")
        return out

    def set_populations(self, which: int, pops: np.ndarray):
        a = np.ascontiguousarray(pops, np.float64)
        if a.shape != (15,) + self.shape:
            raise ValueError(f"populations must have shape {(15,) + self.shape}")
        _check(lib().drp_set_populations(self._h, int(which), _dp(a), a.size), "drp_set_populations")

    def inject_momentum(self, g: np.ndarray):
        a = np.ascontiguousarray(g, np.float64)
        if a.shape != (15,) + self.shape:
            raise ValueError(f"populations must have shape {(15,) + self.shape}")
        _check(lib().drp_inject_momentum(self._h, _dp(a), a.size), "drp_inject_momentum")

    def set_force(self, fx, fy, fz):
        fs = [np.ascontiguousarray(np.broadcast_to(np.asarray(a, np.float64), self.shape)) for a in (fx, fy, fz)]
        _check(lib().drp_set_force(self._h, *[_dp(a) for a in fs], self.n), "drp_set_force")

    def force(self):
        """(fx, fy, fz) of the body-force field the last step used."""
        out = [np.empty(self.shape, np.float64) for _ in range(3)]
        _check(lib().drp_get_force(self._h, *[_dp(a) for a in out], self.n), "drp_get_force")
        return tuple(out)

    # -- the reference's per-rank files (output / outputfg / step_input g files)
    def write_reference_files(self, directory: str, n: int, kind: int, ranks_x: int = 8, ranks_y: int = 1):
        _check(lib().drp_write_reference_files(self._h, os.fsencode(directory), int(n), int(ranks_x), int(ranks_y),
                                               int(kind)), "drp_write_reference_files")

    def read_reference_files(self, directory: str, n: int, kind: int, ranks_x: int = 8, ranks_y: int = 1):
        _check(lib().drp_read_reference_files(self._h, os.fsencode(directory), int(n), int(ranks_x), int(ranks_y),
                                              int(kind)), "drp_read_reference_files")

    # -- analysis hooks
    def reduce(self) -> dict:
        st = Stats()
        _check(lib().drp_reduce(self._h, C.byref(st)), "drp_reduce")
        return st.as_dict()

    def deformation(self, mode: int = 0):
        L, B, D = C.c_double(), C.c_double(), C.c_double()
        _check(lib().drp_deformation(self._h, int(mode), C.byref(L), C.byref(B), C.byref(D)), "drp_deformation")
        return L.value, B.value, D.value

    def energy_spectrum(self, nshells: int | None = None) -> np.ndarray:
        """E[s] of the shells int|k| = s (drp_energy_spectrum; default nx/2 + 1 shells)."""
        ns = self.cfg.nx // 2 + 1 if nshells is None else int(nshells)
        out = np.zeros(ns)
        _check(lib().drp_energy_spectrum(self._h, _dp(out), ns), "drp_energy_spectrum")
        return out

    def scale_interaction(self, nshells: int | None = None) -> np.ndarray:
        """kappa_g <u^(s) . div(grad phi grad phi)> per shell s (drp_scale_interaction)."""
        ns = self.cfg.nx // 2 + 2 if nshells is None else int(nshells)
        out = np.zeros(ns)
        _check(lib().drp_scale_interaction(self._h, _dp(out), ns), "drp_scale_interaction")
        return out

    def derived(self, kind: int) -> np.ndarray:
        """Chemical potential (kind 0), free-energy density (1) or u . grad phi (2) of the
        current state, [z][y][x] (drp_get_derived)."""
        out = np.empty(self.shape, np.float64)
        _check(lib().drp_get_derived(self._h, int(kind), _dp(out), self.n), "drp_get_derived")
        return out

    def phi_contribution(self, nshells: int | None = None):
        """(C[s], bands) of drp_phi_contribution; bands = (all, large, medium, small)."""
        ns = self.cfg.nx // 2 + 1 if nshells is None else int(nshells)
        out = np.zeros(ns)
        bands = np.zeros(4)
        _check(lib().drp_phi_contribution(self._h, _dp(out), ns, _dp(bands)), "drp_phi_contribution")
        return out, bands

    def step_energy_budget(self) -> np.ndarray:
        """One step, then its energy-budget terms (left1, left2, right1 .. right5)."""
        out = np.zeros(7)
        _check(lib().drp_step_energy_budget(self._h, _dp(out)), "drp_step_energy_budget")
        return out

    def filter_velocity(self, kind: int, k: int):
        """(u1, u2, u3) low-passed (kind 0: shells < k) or band-passed (kind 1: shell k)."""
        out = [np.empty(self.shape, np.float64) for _ in range(3)]
        _check(lib().drp_filter_velocity(self._h, int(kind), int(k), *[_dp(a) for a in out], self.n),
               "drp_filter_velocity")
        return tuple(out)

    def droplet_size_histogram(self, scale: float, bin: float, d_max: float, nbins: int,
                               threshold: float | None = None):
        """(counts, total) of diameters 2 r / scale (drp_droplet_size_histogram)."""
        thr = 0.5 * (self.cfg.phi1 + self.cfg.phi2) if threshold is None else float(threshold)
        counts = np.zeros(nbins, np.int64)
        total = C.c_int64(0)
        _check(lib().drp_droplet_size_histogram(self._h, thr, float(scale), float(bin), float(d_max),
                                                counts.ctypes.data_as(C.POINTER(C.c_int64)), int(nbins),
                                                C.byref(total)), "drp_droplet_size_histogram")
        return counts, total.value

    def droplet_census(self, threshold: float | None = None, cap: int = 1 << 16, labels: bool = False):
        """Droplets of the current state (drp_droplet_census): (first, volumes[, labels]),
        ascending in first; threshold defaults to (phi1 + phi2) / 2 as in
        program/percolation.f90."""
        thr = 0.5 * (self.cfg.phi1 + self.cfg.phi2) if threshold is None else float(threshold)
        cnt = C.c_int64(0)
        vol = np.zeros(cap, np.int64)
        first = np.zeros(cap, np.int64)
        lab = np.zeros(self.shape, np.int32) if labels else None
        i64 = lambda a: a.ctypes.data_as(C.POINTER(C.c_int64))  # noqa: E731
        _check(lib().drp_droplet_census(self._h, thr, C.byref(cnt), i64(vol), i64(first), cap,
                                        lab.ctypes.data_as(C.POINTER(C.c_int32)) if labels else None,
                                        self.n if labels else 0), "drp_droplet_census")
        k = cnt.value
        if k > cap:
            raise ValueError(f"{k} droplets exceed cap = {cap}")
        out = (first[:k].copy(), vol[:k].copy())
        return out + (lab,) if labels else out


def tile_shape_count() -> int:
    return int(lib().drp_tile_shape_count())


def autotune(cfg: Config, nsteps: int = 10):
    """Fastest step-kernel tile shape for cfg; returns (best, [ms per step per shape])."""
    n = tile_shape_count()
    best = C.c_int(-1)
    ms = (C.c_double * n)()
    _check(lib().drp_autotune(C.byref(cfg), int(nsteps), C.byref(best), ms), "drp_autotune")
    return best.value, list(ms)


def batch_autotune(cfgs, nsteps: int = 10):
    """Fastest tile shape of the batched step for these members; returns (best, [ms per step per shape])."""
    n = tile_shape_count()
    best = C.c_int(-1)
    ms = (C.c_double * n)()
    arr = (Config * len(cfgs))(*cfgs)
    _check(lib().drp_batch_autotune(arr, len(cfgs), int(nsteps), C.byref(best), ms), "drp_batch_autotune")
    return best.value, list(ms)


def copy_bandwidth_kind(nbytes: int, kind: int, reps: int = 20) -> float:
    """GB/s of yardstick `kind` (0 grid-stride copy, 1 hipMemcpyDtoD, 2 unrolled non-temporal
    copy, 3 unrolled plain copy, 4 read only, 5 write only, 6 twenty SoA streams with
    non-temporal loads) over nbytes."""
    gbs = C.c_double()
    _check(lib().drp_copy_bandwidth_kind(int(nbytes), int(kind), int(reps), C.byref(gbs)), "drp_copy_bandwidth_kind")
    return gbs.value


def fft3d(a: np.ndarray, sign: int = -1) -> np.ndarray:
    """The library's 3-D FFT of a complex [z][y][x] array (unnormalised)."""
    buf = np.ascontiguousarray(a, np.complex128).copy()
    nz, ny, nx = buf.shape
    _check(lib().drp_fft3d(nx, ny, nz, buf.ctypes.data_as(C.POINTER(C.c_double)), int(sign)), "drp_fft3d")
    return buf


def copy_bandwidth_soa(bytes_per_stream: int, streams: int = 40, reps: int = 10) -> float:
    """GB/s of `streams` SoA arrays copied at once (read + write counted)."""
    gbs = C.c_double()
    _check(lib().drp_copy_bandwidth_soa(int(bytes_per_stream), int(streams), int(reps), C.byref(gbs)),
           "drp_copy_bandwidth_soa")
    return gbs.value


def copy_bandwidth(nbytes: int = 1 << 30, reps: int = 20) -> float:
    """GB/s of a plain device stream copy (read + write counted)."""
    gbs = C.c_double()
    _check(lib().drp_copy_bandwidth(int(nbytes), int(reps), C.byref(gbs)), "drp_copy_bandwidth")
    return gbs.value


class Batch:
    """Independent lattices (any sizes and parameters) advanced by one launch per step."""

    def __init__(self, cfgs, tile_shape=None, tune_steps: int = 10):
        """tile_shape: None keeps the library default, an int selects a shape, "auto"
        runs drp_batch_autotune on these members first (self.tuned holds its timings)."""
        self.tuned = None
        if tile_shape == "auto":
            tile_shape, ms = batch_autotune(cfgs, tune_steps)
            self.tuned = ms
        arr = (Config * len(cfgs))(*cfgs)
        h = C.c_void_p()
        _check(lib().drp_batch_create(arr, len(cfgs), C.byref(h)), "drp_batch_create")
        self._h = h
        self.members = []
        for i, cfg in enumerate(cfgs):
            m = Lattice.__new__(Lattice)
            m.cfg = cfg
            m.shape = (cfg.nz, cfg.ny, cfg.nx)
            m.n = cfg.nx * cfg.ny * cfg.nz
            m._h = C.c_void_p(lib().drp_batch_member(h, i))
            m._borrowed = True
            self.members.append(m)
        self.tile_shape_used = tile_shape
        if tile_shape is not None:
            self.set_tile_shape(int(tile_shape))

    def __len__(self):
        return len(self.members)

    def __getitem__(self, i) -> "Lattice":
        return self.members[i]

    def step(self, n: int = 1):
        _check(lib().drp_batch_step(self._h, int(n)), "drp_batch_step")

    def sync(self):
        _check(lib().drp_batch_sync(self._h), "drp_batch_sync")

    def time_steps(self, n: int) -> float:
        ms = C.c_double()
        _check(lib().drp_batch_time_steps(self._h, int(n), C.byref(ms)), "drp_batch_time_steps")
        return ms.value

    def use_graph(self, enable: bool):
        _check(lib().drp_batch_use_graph(self._h, int(bool(enable))), "drp_batch_use_graph")

    def set_tile_shape(self, shape: int):
        _check(lib().drp_batch_set_tile_shape(self._h, int(shape)), "drp_batch_set_tile_shape")

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            for m in self.members:
                m._h = None
            lib().drp_batch_destroy(h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Tape:
    """Records the FP64 run of a lattice with a checkpoint every K steps; backward()
    returns dJ/d(mobility, kappa_g, wall speed) of J = w0 D + w1 A (drp_tape_*)."""

    def __init__(self, lat: Lattice, checkpoint_every: int):
        h = C.c_void_p()
        _check(lib().drp_tape_create(lat._h, int(checkpoint_every), C.byref(h)), "drp_tape_create")
        self._h = h
        self.lattice = lat   # keeps the lattice alive while the tape exists

    def step(self, n: int = 1):
        _check(lib().drp_tape_step(self._h, int(n)), "drp_tape_step")

    def loss(self, w=(1.0, 1.0)):
        """(J, D, A) of the current state."""
        wv = (C.c_double * 2)(*w)
        J, D, A = C.c_double(), C.c_double(), C.c_double()
        _check(lib().drp_tape_loss(self._h, wv, C.byref(J), C.byref(D), C.byref(A)), "drp_tape_loss")
        return J.value, D.value, A.value

    def backward(self, w=(1.0, 1.0)):
        """(J, [dJ/dM, dJ/dkappa_g, dJ/dU]) by the checkpointed adjoint sweep."""
        wv = (C.c_double * 2)(*w)
        J = C.c_double()
        g = (C.c_double * 3)()
        _check(lib().drp_tape_backward(self._h, wv, C.byref(J), g), "drp_tape_backward")
        return J.value, list(g)

    @property
    def nbytes(self) -> int:
        return int(lib().drp_tape_bytes(self._h))

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().drp_tape_destroy(h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
