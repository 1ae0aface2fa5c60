"""Fortran unformatted sequential record files as the reference programs write
and read them (4-byte record markers around every record): the readers and
writers the golden-vector generator uses on the reference's own output files,
shared with the tests of the library's file entry points."""
from __future__ import annotations

import os

import numpy as np


def read_records(path: str, nval: int) -> np.ndarray:
    """Fortran unformatted sequential file of equal records of nval doubles."""
    raw = np.fromfile(path, dtype=np.uint8)
    rec = 8 * nval + 8
    if raw.size % rec:
        raise ValueError(f"{path}: size {raw.size} is not a multiple of {rec}")
    r = raw.reshape(-1, rec)
    head = r[:, :4].copy().view(np.int32).ravel()
    tail = r[:, -4:].copy().view(np.int32).ravel()
    if not ((head == 8 * nval).all() and (tail == 8 * nval).all()):
        raise ValueError(f"{path}: bad record markers")
    return r[:, 4:-4].copy().view(np.float64).reshape(-1, nval)


def write_records(path: str, values: np.ndarray) -> None:
    """One double per record (the g-only input format of program/lbm_drop_xy.f90:799)."""
    v = np.ascontiguousarray(values, np.float64).ravel()
    out = np.empty((v.size, 16), np.uint8)
    marker = np.array([8], np.int32).view(np.uint8)
    out[:, :4] = marker
    out[:, 4:12] = v.view(np.uint8).reshape(-1, 8)
    out[:, 12:] = marker
    out.tofile(path)


def assemble_periodic(outdir: str, step: int, N: int, nranks: int = 8) -> np.ndarray:
    """output() files 0_<rank>_<n>.bin, records (phi,u1,u2,u3,p), z-y-x loops, x split over ranks;
    returns [5][z][y][x] in the order (phi, p, u1, u2, u3)."""
    xp = N // nranks
    out = np.zeros((5, N, N, N))
    for r in range(nranks):
        a = read_records(os.path.join(outdir, f"0_{r}_{step}.bin"), 5).reshape(N, N, xp, 5)
        out[:, :, :, r * xp:(r + 1) * xp] = np.moveaxis(a, -1, 0)
    return out[[0, 4, 1, 2, 3]]


def assemble_pops(outdir: str, step: int, N: int, nranks: int = 8) -> np.ndarray:
    """outputfg() files, records (f_i, g_i), i fastest then x, y, z; returns [2][15][z][y][x]."""
    xp = N // nranks
    f = np.zeros((15, N, N, N))
    g = np.zeros((15, N, N, N))
    for r in range(nranks):
        a = read_records(os.path.join(outdir, f"0_{r}_{step}_fg.bin"), 2).reshape(N, N, xp, 15, 2)
        f[:, :, :, r * xp:(r + 1) * xp] = np.moveaxis(a[..., 0], -1, 0)
        g[:, :, :, r * xp:(r + 1) * xp] = np.moveaxis(a[..., 1], -1, 0)
    return np.stack([f, g])


def xz_blocks(N: int, ranks_x: int, ranks_z: int):
    """(rank, x slice, z slice) of program/LES_drop.f90's decomposition (:153-176):
    ranks_x slabs in x, ranks_z in z, y whole; rank = z index * ranks_x + x index."""
    xp, zp = N // ranks_x, N // ranks_z
    for r in range(ranks_x * ranks_z):
        ix, iz = r % ranks_x, r // ranks_x
        yield r, slice(ix * xp, (ix + 1) * xp), slice(iz * zp, (iz + 1) * zp)


def assemble_xz(outdir: str, step: int, N: int, ranks_x: int, ranks_z: int) -> np.ndarray:
    """output() files of the x-z decomposition (program/LES_drop.f90:800-822), records
    (phi,u1,u2,u3,p); returns [5][z][y][x] in the order (phi, p, u1, u2, u3)."""
    out = np.zeros((5, N, N, N))
    for r, xs, zs in xz_blocks(N, ranks_x, ranks_z):
        a = read_records(os.path.join(outdir, f"0_{r}_{step}.bin"), 5)
        a = a.reshape(zs.stop - zs.start, N, xs.stop - xs.start, 5)
        out[:, zs, :, xs] = np.moveaxis(a, -1, 0)
    return out[[0, 4, 1, 2, 3]]


def write_fg_input(indir: str, num: int, f: np.ndarray, g: np.ndarray, ranks_x: int = 8,
                   ranks_z: int = 1) -> None:
    """(f_i, g_i) record pairs per direction and node, the layout outputfg writes and
    program/LES_drop.f90:921-935 reads back (keeping g), over the x-z decomposition:
    f, g [15][z][y][x]."""
    marker = np.array([16], np.int32).view(np.uint8)
    for r, xs, zs in xz_blocks(g.shape[-1], ranks_x, ranks_z):
        fs = np.moveaxis(f[:, zs, :, xs], 0, -1)
        gs = np.moveaxis(g[:, zs, :, xs], 0, -1)
        v = np.ascontiguousarray(np.stack([fs, gs], axis=-1), np.float64).reshape(-1, 2)
        out = np.empty((v.shape[0], 24), np.uint8)
        out[:, :4] = marker
        out[:, 4:20] = v.view(np.uint8).reshape(-1, 16)
        out[:, 20:] = marker
        out.tofile(os.path.join(indir, f"0_{r}_{num}_fg.bin"))


def write_g_input(indir: str, num: int, g: np.ndarray, nranks: int = 8) -> None:
    """The momentum-population input files read at step_input (:789-828): g [15][z][y][x]."""
    N = g.shape[-1]
    xp = N // nranks
    for r in range(nranks):
        sub = g[:, :, :, r * xp:(r + 1) * xp]  # [i][z][y][x] -> records i fastest, then x, y, z
        write_records(os.path.join(indir, f"0_{r}_{num}_fg.bin"), np.moveaxis(sub, 0, -1))
