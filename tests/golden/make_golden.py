"""Generate the golden-vector fixtures of this directory from the compiled reference.

Run in the development container only (needs /root/reference, amdflang and
MPICH):  python tests/golden/make_golden.py

Each case builds a parameter-patched copy of a reference program with
oracle/ref_build.py, runs it under mpiexec with OMP_NUM_THREADS=1 (the reference
has OpenMP races on module variables, so threads would make it nondeterministic),
reads the Fortran unformatted output files it writes and stores inputs, outputs
and the parameter set as a compressed .npz.  Nothing of the reference itself is
stored: only numbers it produced.
"""
from __future__ import annotations

import json
import math
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)
from oracle import ref_build  # noqa: E402

MPIEXEC = "/opt/conda/bin/mpiexec"
CX = np.array([0, 1, 0, 0, -1, 0, 0, 1, -1, 1, 1, -1, 1, -1, -1])
CY = np.array([0, 0, 1, 0, 0, -1, 0, 1, 1, -1, 1, -1, -1, 1, -1])
CZ = np.array([0, 0, 0, 1, 0, 0, -1, 1, 1, 1, -1, -1, -1, -1, 1])
E = np.array([2 / 9] + [1 / 9] * 6 + [1 / 72] * 8)


# ------------------------------------------------------------------ Fortran I/O
from fortran_io import (assemble_periodic, assemble_pops, assemble_xz, read_records, write_fg_input,  # noqa: E402
                        write_g_input, write_records)


def fnum(v: float) -> str:
    """Fortran double literal of a Python float (shortest round-trip digits)."""
    s = repr(float(v))
    return s.replace("e", "d") if "e" in s else s + "d0"


def run(exe: str, nranks: int, cwd: str) -> None:
    """Run under mpiexec; stdout (the reference prints its step counter) goes to
    $DROPLET_GOLDEN_LOG when set, so long runs can be followed."""
    env = dict(os.environ, OMP_NUM_THREADS="1")
    log = os.environ.get("DROPLET_GOLDEN_LOG")
    with open(log, "w") if log else open(os.devnull, "w") as out:
        subprocess.run([MPIEXEC, "-n", str(nranks), exe], cwd=cwd, env=env, check=True,
                       stdout=out, stderr=subprocess.PIPE)


# ------------------------------------------------------------------ periodic program
def periodic_params(N, step, eta, Re, Ca, Dv, Dr, umax, step_input=10 ** 8):
    return dict(
        xmax=str(N - 1), ymax=str(N - 1), zmax=str(N - 1), step=str(step),
        step_input=str(step_input), step_input_file_num="0",
        datadir_input='"./in/"', datadir_output='"./out/"', datadir_output_fg='"./out/fg/"',
        step_output="1", step_putput_fg="10",
        Re=fnum(Re), eta=fnum(eta), Ca=fnum(Ca), D_vortex=fnum(Dv),
        D_ratio=fnum(Dr), umax=fnum(umax),
        xc="0.5d0*dble(xmax)", yc="0.5d0*dble(ymax)", zc="0.5d0*dble(zmax)")


def periodic_phys(N, eta, Re, Ca, Dv, Dr, umax):
    """Constants exactly as program/lbm_drop_xy.f90:40-51 derives them."""
    D = Dr * Dv
    nu1 = umax * Dv * Dr / Re
    nu2 = eta * nu1
    sigma = nu1 * umax / Ca
    pi = math.acos(-1.0)
    return dict(nx=N, ny=N, nz=N, nu1=nu1, nu2=nu2, kappa_g=(sigma / 1.7039) ** (1.0 / 0.9991),
                tau_f=0.7, kappa_f=0.01, drop=[0.5 * (N - 1)] * 3 + [D], p_init=1.0 / 3.0,
                force_amp=2.0 * pi * pi * umax * umax / (Dv * Re), force_k=pi / Dv)


def case_periodic(name, N, steps, keep, eta, Re, Ca, Dv, Dr, umax, pops_at=None):
    params = periodic_params(N, steps, eta, Re, Ca, Dv, Dr, umax)
    exe = ref_build.build("program/lbm_drop_xy.f90", f"lbm_drop_xy_{name}", params)
    with tempfile.TemporaryDirectory() as tmp:
        run(exe, 8, tmp)
        data = {f"fields_{n}": assemble_periodic(os.path.join(tmp, "out"), n, N) for n in keep}
        if pops_at is not None:
            data[f"pops_{pops_at}"] = assemble_pops(os.path.join(tmp, "out", "fg"), pops_at, N)
    meta = dict(case=name, program="program/lbm_drop_xy.f90", nranks=8, params=params,
                phys=periodic_phys(N, eta, Re, Ca, Dv, Dr, umax), steps=list(keep),
                sha256=ref_build.sha256_of(os.path.join(ref_build.REF_ROOT, "program/lbm_drop_xy.f90")))
    return data, meta


def case_injection(name, N, steps, keep, step_input, seed, eta, Re, Ca, Dv, Dr, umax):
    """HIT-style injection at step_input (program/lbm_drop_xy.f90:789-828) plus the
    Taylor-Green body force computed at that step (externalforce :668-693)."""
    params = periodic_params(N, steps, eta, Re, Ca, Dv, Dr, umax, step_input=step_input)
    params["step_input_file_num"] = "7"
    exe = ref_build.build("program/lbm_drop_xy.f90", f"lbm_drop_xy_{name}", params)
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    k = 2 * np.pi / N
    u1 = 0.02 * np.sin(k * x) * np.cos(k * z) + 0.002 * rng.standard_normal((N, N, N))
    u2 = 0.002 * rng.standard_normal((N, N, N))
    u3 = -0.02 * np.cos(k * x) * np.sin(k * z) + 0.002 * rng.standard_normal((N, N, N))
    p = 1.0 / 3.0 + 1e-3 * rng.standard_normal((N, N, N))
    g = np.empty((15, N, N, N))
    for i in range(15):
        cu = CX[i] * u1 + CY[i] * u2 + CZ[i] * u3
        g[i] = E[i] * (3 * p + 3 * cu + 4.5 * cu * cu - 1.5 * (u1 * u1 + u2 * u2 + u3 * u3))
    with tempfile.TemporaryDirectory() as tmp:
        os.makedirs(os.path.join(tmp, "in"))
        write_g_input(os.path.join(tmp, "in"), 7, g)
        run(exe, 8, tmp)
        data = {f"fields_{n}": assemble_periodic(os.path.join(tmp, "out"), n, N) for n in keep}
    data["g_injected"] = g
    meta = dict(case=name, program="program/lbm_drop_xy.f90", nranks=8, params=params,
                phys=periodic_phys(N, eta, Re, Ca, Dv, Dr, umax), steps=list(keep), step_input=step_input,
                sha256=ref_build.sha256_of(os.path.join(ref_build.REF_ROOT, "program/lbm_drop_xy.f90")))
    return data, meta


def case_walls(name, N, steps, D, nu, uw, kappag):
    """Moving-wall shear droplet, prototype_program/two-phase/dropch.f90 (walls normal to z).
    Nx = 3 is hard-coded (:146) and the field gather to rank 0 names ranks 0..8
    explicitly (:1103-1880), so the program runs on 9 ranks (Ny = 3); with Ny >= 2
    both walls are applied (with Ny = 1 the `else if (group_y == Ny-1)` branch of
    :684 never runs)."""
    assert steps % 30000 == 0, "dropch writes fields only when mod(n,30000) == 0 (:1905)"
    params = dict(xmax=str(N - 1), ymax=str(N - 1), zmax=str(N - 1), step=str(steps), start="1",
                  datadir='"./out/"', D=fnum(D) + " * ds", nu1=fnum(nu), nu2=fnum(nu),
                  uw=fnum(uw), kappag=fnum(kappag))
    exe = ref_build.build("prototype_program/two-phase/dropch.f90", f"dropch_{name}", params)
    with tempfile.TemporaryDirectory() as tmp:
        run(exe, 9, tmp)
        rec = read_records(os.path.join(tmp, "out", f"{steps}.bin"), 8).reshape(N, N, N, 8)
    xyz = rec[..., :3]
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    assert (xyz[..., 0] == x).all() and (xyz[..., 1] == y).all() and (xyz[..., 2] == z).all()
    data = {f"phi_{steps}": rec[..., 3], f"p_{steps}": rec[..., 7]}
    meta = dict(case=name, program="prototype_program/two-phase/dropch.f90", nranks=9, params=params,
                phys=dict(nx=N, ny=N, nz=N, nu1=nu, nu2=nu, tau_f=0.7, kappa_f=0.01, kappa_g=kappag,
                          wall_u=uw, wall_start=1, drop=[0.5 * (N - 1)] * 3 + [D], p_init=0.0),
                steps=[steps],
                sha256=ref_build.sha256_of(os.path.join(ref_build.REF_ROOT, "prototype_program/two-phase/dropch.f90")))
    return data, meta


def case_allen_cahn(name, N, steps, D, A, W, start, nu=0.2, uw=0.01, kappag=8.132e-3):
    """Conservative Allen-Cahn shear droplet, prototype_program/two-phase/dropac.f90
    (walls normal to z moving from step `start`).  Nx = 3 is hard-coded and the
    gather to rank 0 names ranks 0..8, so it runs on 9 ranks.  The program writes
    text only: every 100 steps the plane y = ymax/2 as (x, z, u1, u2, u3, p, phi)
    with 9 significant digits (output :932-952), and every step
    (n, shear strain, L, B, D) of its deformation measure to break.d (:929)."""
    assert steps % 100 == 0
    params = dict(xmax=str(N - 1), ymax=str(N - 1), zmax=str(N - 1), step=str(steps), start=str(start),
                  datadir='"./out/"', D=fnum(D) + " * ds", nu1=fnum(nu), nu2=fnum(nu), uw=fnum(uw),
                  kappag=fnum(kappag), A=fnum(A), W=fnum(W) + " * ds")
    exe = ref_build.build("prototype_program/two-phase/dropac.f90", f"dropac_{name}", params)
    keep = list(range(100, steps + 1, 100))
    data = {}
    with tempfile.TemporaryDirectory() as tmp:
        run(exe, 9, tmp)
        for n in keep:
            rows = np.loadtxt(os.path.join(tmp, "out", f"{n}.d"))
            assert rows.shape == (N * N, 7)
            assert (rows[:, 0].reshape(N, N) == np.arange(N)[None, :]).all()
            assert (rows[:, 1].reshape(N, N) == np.arange(N)[:, None]).all()
            data[f"plane_{n}"] = rows[:, 2:].reshape(N, N, 5)   # [z][x][u1, u2, u3, p, phi]
        data["break"] = np.loadtxt(os.path.join(tmp, "break.d"))  # n, strain, L, B, D per step
    meta = dict(case=name, program="prototype_program/two-phase/dropac.f90", nranks=9, params=params,
                phys=dict(nx=N, ny=N, nz=N, nu1=nu, nu2=nu, kappa_g=kappag, ac_a=A, ac_width=W, phi1=-0.5,
                          phi2=0.5, wall_u=uw, wall_start=start, drop=[0.5 * (N - 1)] * 3 + [D], p_init=0.0,
                          plane_y=(N - 1) // 2),
                steps=keep,
                sha256=ref_build.sha256_of(os.path.join(ref_build.REF_ROOT, "prototype_program/two-phase/dropac.f90")))
    return data, meta


def injected_momentum(N, seed, amp=0.02, noise=0.002):
    """A momentum-population field to inject: equilibria of a Taylor-Green-like
    velocity plus seeded noise (numpy.random.default_rng(seed))."""
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    k = 2 * np.pi / N
    u1 = amp * np.sin(k * x) * np.cos(k * z) + noise * rng.standard_normal((N, N, N))
    u2 = noise * rng.standard_normal((N, N, N))
    u3 = -amp * np.cos(k * x) * np.sin(k * z) + noise * rng.standard_normal((N, N, N))
    p = 1.0 / 3.0 + 1e-3 * rng.standard_normal((N, N, N))
    g = np.empty((15, N, N, N))
    for i in range(15):
        cu = CX[i] * u1 + CY[i] * u2 + CZ[i] * u3
        g[i] = E[i] * (3 * p + 3 * cu + 4.5 * cu * cu - 1.5 * (u1 * u1 + u2 * u2 + u3 * u3))
    return g


def case_les(name, N, steps, step_input, seed, delta, Cs, Re, Dv, D, sigma, umax=0.1, ranks_z=2):
    """Smagorinsky LES with a droplet, program/LES_drop.f90 on 8 * ranks_z ranks (8 slabs
    in x times ranks_z in z; its z-neighbour set-up needs at least 2 z slabs, :184-192): momentum populations injected at step_input from (f, g) pair files (the
    f values are ignored, :921-945), Taylor-Green force from then on (:774-798),
    eddy viscosity from grad u at the end of each step with halos exchanged at its
    start (physics :598-693).  Fields written when n is a multiple of 500."""
    assert steps % 500 == 0, "LES_drop.f90 writes fields only when mod(n,500) == 0"
    params = dict(xmax=str(N - 1), ymax=str(N - 1), zmax=str(N - 1), step=str(steps), step_input=str(step_input),
                  step_input_file_num="7", datadir_input='"./in/"', datadir='"./out/"', datadir2='"./out/fg/"',
                  delta=fnum(delta) + "*ds", Cs=fnum(Cs), Re=fnum(Re), D_vortex=fnum(Dv), D=fnum(D),
                  sigma=fnum(sigma), umax=fnum(umax))
    exe = ref_build.build("program/LES_drop.f90", f"les_drop_{name}", params)
    g = injected_momentum(N, seed)
    f_dummy = np.full_like(g, 12345.0)
    with tempfile.TemporaryDirectory() as tmp:
        os.makedirs(os.path.join(tmp, "in"))
        write_fg_input(os.path.join(tmp, "in"), 7, f_dummy, g, 8, ranks_z)
        run(exe, 8 * ranks_z, tmp)
        fields = assemble_xz(os.path.join(tmp, "out"), steps, N, 8, ranks_z)
    pi = math.acos(-1.0)
    nu1 = umax * Dv / Re
    data = {f"fields_{steps}": fields, "g_injected": g}
    meta = dict(case=name, program="program/LES_drop.f90", nranks=8 * ranks_z, params=params, step_input=step_input,
                phys=dict(nx=N, ny=N, nz=N, nu1=nu1, nu2=1.0 * nu1, tau_f=0.7, kappa_f=0.01,
                          kappa_g=(sigma / 1.7039) ** (1.0 / 0.9991), drop=[0.5 * (N - 1)] * 3 + [D], p_init=1.0 / 3.0,
                          force_amp=2.0 * pi * pi * umax * umax / (Dv * Re), force_k=pi / Dv,
                          les_cs=Cs, les_delta=delta * 1.0, les_ranks_x=8, les_ranks_y=1, les_ranks_z=ranks_z),
                steps=[steps],
                sha256=ref_build.sha256_of(os.path.join(ref_build.REF_ROOT, "program/LES_drop.f90")))
    return data, meta


def sphere_field(N, drops, phi1=2.211, phi2=4.895):
    """Sharp periodic spheres: phi2 where the minimum-image distance to a centre
    (xc, yc, zc) is <= r, phi1 elsewhere; [z][y][x].  Pure float64 arithmetic,
    so the tests rebuild the same fields from the stored lists."""
    z, y, x = np.meshgrid(np.arange(N, dtype=np.float64), np.arange(N, dtype=np.float64),
                          np.arange(N, dtype=np.float64), indexing="ij")
    phi = np.full((N, N, N), phi1)
    for xc, yc, zc, r in drops:
        dx = np.abs(x - xc)
        dy = np.abs(y - yc)
        dz = np.abs(z - zc)
        dx = np.minimum(dx, N - dx)
        dy = np.minimum(dy, N - dy)
        dz = np.minimum(dz, N - dz)
        phi[dx * dx + dy * dy + dz * dz <= r * r] = phi2
    return phi


def case_dsd(name, N, seed):
    """program/percolation.f90 on 71 fields (its loop reads 1_<step>.bin for step =
    5000, 6000, ..., 75000 and appends (step, droplet count) to drop_num_1.d):
    1-12 sharp spheres of radius 1.5-5 at seeded positions, periodic images
    included, so some cross the periodic seams and some overlap."""
    params = dict(xmax=str(N - 1), ymax=str(N - 1), zmax=str(N - 1), phi1="2.211d0", phi2="4.895d0",
                  datadir='"./"', datadir2='"./radius/"')
    exe = ref_build.build("program/percolation.f90", f"percolation_{name}", params)
    rng = np.random.default_rng(seed)
    steps = list(range(5000, 75001, 1000))
    drops = np.zeros((len(steps), 12, 4))
    ndrops = np.zeros(len(steps), np.int64)
    with tempfile.TemporaryDirectory() as tmp:
        os.makedirs(os.path.join(tmp, "radius"))
        for k, st in enumerate(steps):
            nd = 1 + k % 12
            ndrops[k] = nd
            drops[k, :nd, :3] = rng.uniform(0.0, N, (nd, 3))
            drops[k, :nd, 3] = rng.choice([1.5, 2.0, 2.6, 3.2, 4.1, 5.0], nd)
            phi = sphere_field(N, drops[k, :nd])
            # one record per node, z outermost, x fastest (:50-58)
            write_records(os.path.join(tmp, f"1_{st}.bin"), phi.ravel())
        subprocess.run([exe], cwd=tmp, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE)
        rows = np.loadtxt(os.path.join(tmp, "drop_num_1.d"))
    assert rows.shape == (len(steps), 2) and np.array_equal(rows[:, 0], steps)
    data = {"drops": drops, "ndrops": ndrops, "counts": rows[:, 1].astype(np.int64)}
    meta = dict(case=name, program="program/percolation.f90", params=params, N=N, steps=steps,
                phi1=2.211, phi2=4.895,
                sha256=ref_build.sha256_of(os.path.join(ref_build.REF_ROOT, "program/percolation.f90")))
    return data, meta


def save(name, data, meta):
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, meta=np.array(json.dumps(meta)), **data)
    print(f"wrote {path} ({os.path.getsize(path) / 1e6:.2f} MB)")


CASES = {
    # 16^3, viscosity ratio 2 (nu(phi) depends on the global phi range every step)
    "periodic16": lambda: case_periodic("periodic16", 16, 200, (1, 10, 200), eta=2.0, Re=0.8, Ca=0.1,
                                        Dv=16.0, Dr=0.5, umax=0.01, pops_at=10),
    # 32^3, equal viscosities, 2000 steps (the step count of the 64^3 shear case)
    "periodic32": lambda: case_periodic("periodic32", 32, 2000, (2000,), eta=1.0, Re=1.2, Ca=0.05,
                                        Dv=32.0, Dr=0.375, umax=0.01),
    # injected momentum populations + Taylor-Green body force from step 3 on
    "inject16": lambda: case_injection("inject16", 16, 60, (3, 4, 60), step_input=3, seed=16, eta=1.5,
                                       Re=0.8, Ca=2.0, Dv=16.0, Dr=0.5, umax=0.02),
    # moving walls normal to z, 24^3, 30000 steps; D = 12 because a D = 10 drop dissolves
    # into the bulk near step 9450, after which the unbounded axis walk of the program's
    # deformation measure (:905-944) never exits
    "walls24": lambda: case_walls("walls24", 24, 30000, D=12.0, nu=0.1, uw=0.01, kappag=0.01),
    # Smagorinsky LES (Cs 0.173, Delta 4) with a droplet, injection + Taylor-Green force, 500 steps
    # conservative Allen-Cahn model (dropac.f90), 24^3, mobility M = (1 - A)/6 = 0.1, W = 4,
    # walls moving from step 50, 300 steps
    "ac24": lambda: case_allen_cahn("ac24", 24, 300, D=12.0, A=0.4, W=4.0, start=50),
    # droplet counts of program/percolation.f90 for 71 seeded sphere layouts at 32^3
    "dsd32": lambda: case_dsd("dsd32", 32, seed=31),
    "les16": lambda: case_les("les16", 16, 500, step_input=3, seed=21, delta=4.0, Cs=0.173, Re=80.0, Dv=8.0,
                              D=8.0, sigma=1e-4),
    # the same on 32 ranks: 8 x-slabs times 4 z-slabs of 4 nodes
    "les16z4": lambda: case_les("les16z4", 16, 500, step_input=3, seed=22, delta=4.0, Cs=0.173, Re=80.0, Dv=8.0,
                                D=8.0, sigma=1e-4, ranks_z=4),
}


def main():
    if not ref_build.available():
        sys.exit("reference sources / amdflang / MPICH not available")
    for name in (sys.argv[1:] or list(CASES)):
        save(name, *CASES[name]())


if __name__ == "__main__":
    main()
