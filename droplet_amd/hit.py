"""Forced isotropic turbulence with droplets: configuration and initial state.

Parameter set B of program/lbm_drop_iht_xy.f90:37-70 (phi1 = 0.2638, phi2 = 0.4031,
a = b = 1, T = 0.293, kappa_f = 0.06, tau_f = 0.7, kappa_g = sigma / 2.95e-3,
sigma = eps^(2/3) D^(5/3) / We, eps = nu^3 / eta^4 with the Kolmogorov scale
eta = 1), the linear forcing of the shells 1 <= |k| < kc (kc = 3, externalforce
:722-841) and Smagorinsky LES (program/LES_drop_xy.f90:661-693).

The reference starts its droplet runs from momentum populations of a developed
turbulence run read from disk (:789-828).  Here the initial velocity is a
synthetic solenoidal field with E(k) ~ k^-5/3 on the shells 1..8 and a given
u_rms, drawn from numpy.random.default_rng(seed); the same generator then places
non-overlapping droplets.  These helpers build inputs on the host; the time
step itself runs in the HIP library.
"""
from __future__ import annotations

import math

import numpy as np

from .abi import DRP_FORCE_SPECTRAL, DRP_FP64, Config, make_config

SET_B = dict(phi1=0.2638, phi2=0.4031, a=1.0, b=1.0, T=0.293, kappa_f=0.06, tau_f=0.7)


def sigma_from_weber(We: float, eps: float, D: float) -> float:
    """sigma = eps^(2/3) D^(5/3) / We (program/lbm_drop_iht_xy.f90:49)."""
    return eps ** (2.0 / 3.0) * D ** (5.0 / 3.0) / We


def hit_config(N: int, We: float, D: float = 20.0, nu: float = 1e-3, eta: float = 1.0, kc: int = 3,
               les_cs: float = 0.1, les_delta: float = 1.0, precision: int = DRP_FP64, force_start: int = 1,
               **over) -> Config:
    """N^3 periodic box, set-B droplets at Weber number We, spectral forcing with
    eps = nu^3 / eta^4 (:70), Smagorinsky Cs, Delta."""
    eps = nu ** 3 / eta ** 4
    sigma = sigma_from_weber(We, eps, D)
    kw = dict(SET_B, nu1=nu, nu2=nu, kappa_g=sigma / 2.95e-3, precision=precision, les_cs=les_cs,
              les_delta=les_delta, force_kind=DRP_FORCE_SPECTRAL, force_amp=eps, force_k=float(kc),
              force_start=force_start)
    kw.update(over)
    return make_config(N, **kw)


def initial_velocity(N: int, rng: np.random.Generator, u_rms: float = 0.03, kmin: int = 1, kmax: int = 8,
                     slope: float = -5.0 / 3.0):
    """Solenoidal periodic field (u1, u2, u3), each [z][y][x], with shell energy
    E(k) ~ k^slope for kmin <= |k| < kmax + 1 and zero elsewhere, scaled to u_rms
    (= sqrt(<|u|^2> / 3))."""
    k1 = np.fft.fftfreq(N) * N
    kz, ky, kx = np.meshgrid(k1, k1, k1, indexing="ij")
    k2 = kx * kx + ky * ky + kz * kz
    kmag = np.sqrt(k2)
    shell = np.floor(kmag + 0.5)
    band = (shell >= kmin) & (shell <= kmax)
    # random complex amplitudes, projected onto the plane normal to k
    a = [rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N)) for _ in range(3)]
    safe = np.where(k2 > 0, k2, 1.0)
    kdota = (kx * a[0] + ky * a[1] + kz * a[2]) / safe
    a = [a[0] - kx * kdota, a[1] - ky * kdota, a[2] - kz * kdota]
    # scale every shell to E(k) ~ k^slope
    e_mode = 0.5 * (np.abs(a[0]) ** 2 + np.abs(a[1]) ** 2 + np.abs(a[2]) ** 2)
    target = np.zeros_like(kmag)
    for s in range(kmin, kmax + 1):
        m = band & (shell == s)
        tot = e_mode[m].sum()
        if tot > 0:
            target[m] = s ** slope / tot
    scale = np.sqrt(np.where(band, target, 0.0))
    u = [np.real(np.fft.ifftn(ai * scale)) for ai in a]   # real part: the field is symmetrised
    rms = math.sqrt(sum(float((ui * ui).mean()) for ui in u) / 3.0)
    return [np.ascontiguousarray(ui * (u_rms / rms)) for ui in u]


def place_droplets(N: int, count: int, D: float, rng: np.random.Generator, gap: float = 2.0, tries: int = 100000):
    """count centres in [0, N)^3 whose periodic distances are at least D + gap."""
    centres = []
    for _ in range(tries):
        if len(centres) == count:
            break
        c = rng.uniform(0.0, N, 3)
        ok = True
        for o in centres:
            d = np.abs(c - o)
            d = np.minimum(d, N - d)
            if float(np.sqrt((d * d).sum())) < D + gap:
                ok = False
                break
        if ok:
            centres.append(c)
    if len(centres) < count:
        raise ValueError(f"could not place {count} droplets of diameter {D} in {N}^3")
    return [(float(c[0]), float(c[1]), float(c[2]), float(D)) for c in centres]


def droplet_field(N: int, drops, phi1: float, phi2: float):
    """Sharp-step phi of periodic droplets (ini :224-235 with periodic images)."""
    z, y, x = np.meshgrid(np.arange(N, dtype=np.float64), np.arange(N, dtype=np.float64),
                          np.arange(N, dtype=np.float64), indexing="ij")
    phi = np.full((N, N, N), phi1)
    for (xc, yc, zc, D) in drops:
        dx = np.abs(x - xc)
        dy = np.abs(y - yc)
        dz = np.abs(z - zc)
        dx, dy, dz = np.minimum(dx, N - dx), np.minimum(dy, N - dy), np.minimum(dz, N - dz)
        phi[dx * dx + dy * dy + dz * dz <= (0.5 * D) ** 2] = phi2
    return phi


def initial_state(N: int, seed: int = 3, u_rms: float = 0.03, count: int = 32, D: float = 20.0,
                  phi1: float = SET_B["phi1"], phi2: float = SET_B["phi2"]):
    """(phi, p, u1, u2, u3, drops) of the turbulent droplet case, numpy.random.default_rng(seed)."""
    rng = np.random.default_rng(seed)
    u1, u2, u3 = initial_velocity(N, rng, u_rms)
    drops = place_droplets(N, count, D, rng)
    phi = droplet_field(N, drops, phi1, phi2)
    p = np.full((N, N, N), 1.0 / 3.0)
    return phi, p, u1, u2, u3, drops
