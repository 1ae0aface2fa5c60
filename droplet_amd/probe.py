"""Small GPU drivers for profiling runs (rocprofv3) of single kernels.

    python -m droplet_amd.probe copy [--mib 1024] [--reps 20]
        stream copy of a known size: the calibration kernel for FETCH_SIZE/WRITE_SIZE
    python -m droplet_amd.probe copysoa [--grid 256] [--reps 20]
        10, 20 and 40 structure-of-arrays streams of grid^3 doubles (40: floats'
        worth of bytes) copied at once: the achievable rate of the step's pattern
    python -m droplet_amd.probe yardsticks [--reps 20]
        independent bandwidth ceilings at 1, 2 and 4 GiB: grid-stride copy,
        hipMemcpyDtoD, 4x-unrolled non-temporal and plain copies, read only, write
        only, 20 SoA streams with non-temporal loads; then the copysoa streams at 256^3
    python -m droplet_amd.probe step [--grid 256] [--steps 20] [--fp32] [--shape S]
        the 256^3 shear-droplet step kernel of bench.py, a few launches
    python -m droplet_amd.probe hit [--grid 256] [--steps 20] [--fp32] [--we 5]
        forced isotropic turbulence with 32 droplets of radius 10 (set-B parameters,
        Smagorinsky Cs = 0.1, spectral forcing of the shells 1 <= |k| < 3, solenoidal
        k^-5/3 initial velocity with u_rms = 0.03, numpy.random.default_rng(3)):
        MLUPS of the whole step (forcing + LES + step kernels)
    python -m droplet_amd.probe allen-cahn [--grid 256] [--steps 20]
        the 256^3 shear droplet of bench.py with the conservative Allen-Cahn phase
        field (A = 0.4, i.e. mobility 0.1, W = 4, phi in [-0.5, 0.5]): MLUPS
    python -m droplet_amd.probe spectra [--grid 256] [--reps 3]
        FFT diagnostics of the 256^3 turbulent-droplet state: energy spectrum,
        per-scale interaction and a low-pass filter, milliseconds per call
    python -m droplet_amd.probe adjoint [--grid 256] [--steps 20] [--checkpoint 10]
        taped FP64 run of the shear droplet and the checkpointed adjoint sweep of
        J = D + A: ms per forward and per reverse step
    python -m droplet_amd.probe noneq|scale|budget [--grid 256] [--steps 20]
        bench.py's shear droplet with the non-equilibrium Smagorinsky model, with
        scale-separated phi advection, or stepped through drp_step_energy_budget:
        ms per step (and per budget call)
    python -m droplet_amd.probe diag [--grid 256] [--reps 20]
        ms per drp_reduce, drp_deformation and copy of the five fields to the host,
        per write of the reference's field files at grid^3 and per write and
        read-back of its population (restart) files at (grid/2)^3
    python -m droplet_amd.probe ensemble [--members 64] [--steps 60] [--fp64]
        64 shear-droplet members, N uniform in {48, 64, 96, 128}, radius N/5,
        Ca uniform in [0.05, 0.5], centre offsets within +-2 cells
        (numpy.random.default_rng(4)), FP32 shifted storage: aggregate MLUPS of
        one batched launch per step (hipGraph) against stepping the members one
        after another
"""
from __future__ import annotations

import argparse
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def ensemble(a, dl):
    import numpy as np
    from droplet_amd.abi import DRP_FP32_SHIFTED, DRP_FP64, DRP_WALL_Y, make_config
    rng = np.random.default_rng(4)
    kappa_g = 0.002
    sigma = 1.7039 * kappa_g ** 0.9991
    cfgs, drops = [], []
    for _ in range(a.members):
        N = int(rng.choice(a.sizes))
        Ca = float(rng.uniform(0.05, 0.5))
        D = 2.0 * N / 5.0
        U = Ca * (N - 1) * sigma / (0.1 * D)   # Ca = U nu D / (H sigma)
        cfgs.append(make_config(N, precision=DRP_FP64 if a.fp64 else DRP_FP32_SHIFTED, nu1=0.1, nu2=0.1,
                                kappa_g=kappa_g, wall_axis=DRP_WALL_Y, wall_u=U, wall_start=1))
        off = rng.uniform(-2.0, 2.0, 3)
        drops.append([tuple(0.5 * (N - 1) + off) + (D,)])
    nodes = sum(c.nx * c.ny * c.nz for c in cfgs)
    b = dl.Batch(cfgs)
    if a.shape >= 0:
        b.set_tile_shape(a.shape)
    for m, d in zip(b.members, drops):
        m.init_droplets(d)
    b.step(12)
    ms_batch = b.time_steps(a.steps)
    b.close()
    solos = []
    for c, d in zip(cfgs, drops):
        s = dl.Lattice(c)
        s.init_droplets(d)
        s.step(2)
        solos.append(s)
    for s in solos:
        s.sync()
    import time
    t0 = time.perf_counter()
    for _ in range(a.steps):
        for s in solos:
            s.step(1)
    for s in solos:
        s.sync()
    ms_loop = (time.perf_counter() - t0) * 1e3
    print(json.dumps({"members": a.members, "nodes": nodes, "steps": a.steps,
                      "batch_ms_per_step": ms_batch / a.steps, "batch_MLUPS": nodes * a.steps / ms_batch / 1e3,
                      "loop_ms_per_step": ms_loop / a.steps, "loop_MLUPS": nodes * a.steps / ms_loop / 1e3,
                      "speedup": ms_loop / ms_batch}))


def turbulence(a, dl):
    import time
    from droplet_amd import hit
    from droplet_amd.abi import DRP_FP32_SHIFTED, DRP_FP64
    N = a.grid
    t0 = time.perf_counter()
    phi, p, u1, u2, u3, drops = hit.initial_state(N, seed=3, u_rms=0.03, count=32, D=20.0)
    t_init = time.perf_counter() - t0
    cfg = hit.hit_config(N, We=a.we, D=20.0, precision=DRP_FP32_SHIFTED if a.fp32 else DRP_FP64)
    lat = dl.Lattice(cfg)
    if a.shape >= 0:
        lat.tile_shape = a.shape
    lat.init_fields(phi, p, u1, u2, u3)
    del phi, p, u1, u2, u3
    lat.step(3)
    ms = lat.time_steps(a.steps)
    st = lat.reduce()
    print(json.dumps({"workload": f"turbulent droplets {N}^3, 32 x R10, We {a.we}, LES Cs 0.1, spectral forcing kc 3",
                      "precision": "fp32_shifted" if a.fp32 else "fp64", "steps": a.steps, "shape": lat.tile_shape,
                      "ms_per_step": ms / a.steps, "MLUPS": N ** 3 * a.steps / (ms * 1e-3) / 1e6,
                      "kinetic": st["kinetic"], "phi_mass": st["phi_mass"], "host_init_s": round(t_init, 1)}))


def allen_cahn(a, dl):
    from droplet_amd.abi import DRP_MODEL_ALLEN_CAHN, DRP_WALL_Y, make_config
    from bench import shear_config
    base, drops, phys = shear_config(a.grid)
    cfg = make_config(a.grid, model=DRP_MODEL_ALLEN_CAHN, ac_a=0.4, ac_width=4.0, phi1=-0.5, phi2=0.5,
                      nu1=0.1, nu2=0.1, kappa_g=base.kappa_g, wall_axis=DRP_WALL_Y, wall_u=base.wall_u)
    lat = dl.Lattice(cfg)
    if a.shape >= 0:
        lat.tile_shape = a.shape
    lat.init_droplets(drops, 0.0)
    lat.step(3)
    ms = lat.time_steps(a.steps)
    n = a.grid ** 3
    print(json.dumps({"workload": f"Allen-Cahn shear droplet {a.grid}^3", "shape": lat.tile_shape, "steps": a.steps,
                      "ms_per_step": ms / a.steps, "MLUPS": n * a.steps / (ms * 1e-3) / 1e6,
                      "algorithmic_GBs": 320 * n / (ms / a.steps * 1e-3) / 1e9}))


def spectra(a, dl):
    """FFT diagnostics of the turbulent-droplet state, ms per call (host sync included)."""
    import time
    from droplet_amd import hit
    N = a.grid
    phi, p, u1, u2, u3, _ = hit.initial_state(N, seed=3, u_rms=0.03, count=32, D=20.0)
    lat = dl.Lattice(hit.hit_config(N, We=a.we, D=20.0))
    lat.init_fields(phi, p, u1, u2, u3)
    del phi, p, u1, u2, u3
    lat.step(2)
    out = {"grid": N}
    for name, fn in (("energy_spectrum", lambda: lat.energy_spectrum()),
                     ("scale_interaction", lambda: lat.scale_interaction()),
                     ("low_pass_k8", lambda: lat.filter_velocity(0, 8))):
        fn()
        t0 = time.perf_counter()
        for _ in range(a.reps):
            r = fn()
        out[name + "_ms"] = (time.perf_counter() - t0) * 1e3 / a.reps
    E = lat.energy_spectrum()
    out["E_shells_1_8"] = [float(v) for v in E[1:9]]
    out["kinetic"] = lat.reduce()["kinetic"]
    out["sum_E"] = float(E.sum())
    print(json.dumps(out))


def adjoint(a, dl):
    """Taped FP64 run of bench.py's shear droplet (checkpoint every K steps) and
    the checkpointed adjoint sweep of J = D + A: ms per taped forward step and per
    reverse step (segment recomputation, adjoint collision and stencil kernels)."""
    import time
    from bench import shear_config
    cfg, drops, _ = shear_config(a.grid)
    lat = dl.Lattice(cfg)
    lat.init_droplets(drops, 1.0 / 3.0)
    tape = dl.Tape(lat, a.checkpoint)
    t0 = time.perf_counter()
    tape.step(a.steps)
    J, D, A = tape.loss()
    t1 = time.perf_counter()
    J2, g = tape.backward((1.0, 1.0))
    t2 = time.perf_counter()
    n = a.grid ** 3
    fwd, rev = (t1 - t0) * 1e3 / a.steps, (t2 - t1) * 1e3 / a.steps
    print(json.dumps({"workload": f"adjoint of the shear droplet {a.grid}^3", "steps": a.steps,
                      "checkpoint_every": a.checkpoint, "tape_GB": round(tape.nbytes / 1e9, 2),
                      "forward_ms_per_step": round(fwd, 3), "reverse_ms_per_step": round(rev, 3),
                      "reverse_MLUPS": round(n / (rev * 1e-3) / 1e6, 1), "J": J, "D": D, "A": A,
                      "dJ": g}))
    tape.close()
    lat.close()


def variants(a, dl):
    """bench.py's 256^3 shear droplet (FP64) in one of the model variants: the
    non-equilibrium Smagorinsky model (Cs = 0.17, Delta = 1, nu2 = 2 nu1), the
    scale-separated advection of the shells int|k| < 3 from step 1, or the step
    with the energy budget of energy_budget.f90 (ms per call of
    drp_step_energy_budget against a plain step)."""
    from bench import shear_config
    from droplet_amd.abi import DRP_LES_NONEQ
    cfg, drops, _ = shear_config(a.grid)
    if a.what == "noneq":
        cfg.les_model, cfg.les_cs, cfg.les_delta, cfg.nu2 = DRP_LES_NONEQ, 0.17, 1.0, 2.0 * cfg.nu1
    elif a.what == "scale":
        cfg.advect_k_high, cfg.advect_k_low, cfg.advect_start = 3, 0, 1
    lat = dl.Lattice(cfg)
    lat.init_droplets(drops, 1.0 / 3.0)
    lat.step(3)
    n = a.grid ** 3
    out = {"workload": f"{a.what}: shear droplet {a.grid}^3", "steps": a.steps}
    if a.what == "budget":
        import time
        lat.step_energy_budget()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            terms = lat.step_energy_budget()
        out["budget_ms_per_call"] = (time.perf_counter() - t0) * 1e3 / a.steps
        out["terms"] = [float(v) for v in terms]
    ms = lat.time_steps(a.steps)
    out.update({"ms_per_step": ms / a.steps, "MLUPS": n * a.steps / (ms * 1e-3) / 1e6})
    print(json.dumps(out))
    lat.close()


def diagnostics(a, dl):
    """Analysis hooks and file formats on bench.py's shear droplet (FP64): ms per
    drp_reduce, drp_deformation and drp_get_macro of the five fields; the
    reference-format field files (output) at grid^3 and a population restart
    (outputfg written, then read back) at (grid/2)^3, in a temporary directory."""
    import shutil
    import tempfile
    import time
    from bench import shear_config
    from droplet_amd.abi import DRP_FILE_FG, DRP_FILE_FIELDS

    def per_call(fn, reps):
        fn()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        return (time.perf_counter() - t0) * 1e3 / reps

    out = {"grid": a.grid}
    cfg, drops, _ = shear_config(a.grid)
    lat = dl.Lattice(cfg)
    lat.init_droplets(drops, 1.0 / 3.0)
    lat.step(3)
    out["reduce_ms"] = per_call(lat.reduce, a.reps)
    out["deformation_ms"] = per_call(lambda: lat.deformation(0), a.reps)
    out["fields_to_host_ms"] = per_call(lat.fields, a.reps)
    d = tempfile.mkdtemp(prefix="drp_files_")
    try:
        out["write_fields_ms"] = per_call(lambda: lat.write_reference_files(d, 3, DRP_FILE_FIELDS), 2)
        out["fields_file_bytes"] = sum(os.path.getsize(os.path.join(d, f)) for f in os.listdir(d))
        lat.close()
        shutil.rmtree(d)
        os.makedirs(d)
        g = a.grid // 2
        cfg, drops, _ = shear_config(g)
        lat = dl.Lattice(cfg)
        lat.init_droplets(drops, 1.0 / 3.0)
        lat.step(3)
        out["restart_grid"] = g
        out["write_fg_ms"] = per_call(lambda: lat.write_reference_files(d, 3, DRP_FILE_FG), 1)
        out["fg_file_bytes"] = sum(os.path.getsize(os.path.join(d, f)) for f in os.listdir(d))
        out["read_fg_ms"] = per_call(lambda: lat.read_reference_files(d, 3, DRP_FILE_FG), 1)
        lat.close()
    finally:
        shutil.rmtree(d, ignore_errors=True)
    print(json.dumps(out))


def droplet_census(a, dl):
    """Census of the turbulent-droplet state (32 droplets of radius 10) after a
    few steps: droplet count, volume range and time per call (HIP sync included)."""
    import time
    from droplet_amd import hit
    from droplet_amd.abi import DRP_FP32_SHIFTED, DRP_FP64
    N = a.grid
    phi, p, u1, u2, u3, _ = hit.initial_state(N, seed=3, u_rms=0.03, count=32, D=20.0)
    cfg = hit.hit_config(N, We=a.we, D=20.0, precision=DRP_FP32_SHIFTED if a.fp32 else DRP_FP64)
    lat = dl.Lattice(cfg)
    lat.init_fields(phi, p, u1, u2, u3)
    del phi, p, u1, u2, u3
    lat.step(10)
    lat.droplet_census()
    t0 = time.perf_counter()
    for _ in range(a.reps):
        first, vol = lat.droplet_census()
    ms = (time.perf_counter() - t0) * 1e3 / a.reps
    print(json.dumps({"grid": N, "droplets": int(len(first)), "volume_min": int(vol.min()), "volume_max": int(vol.max()),
                      "ms_per_call": ms, "reps": a.reps}))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["copy", "copysoa", "yardsticks", "step", "autotune", "ensemble", "hit", "census", "allen-cahn",
                                     "spectra", "adjoint", "noneq", "scale", "budget", "diag"])
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--fp32", action="store_true")
    ap.add_argument("--shape", type=int, default=-1)
    ap.add_argument("--members", type=int, default=64)
    ap.add_argument("--sizes", type=lambda s: [int(v) for v in s.split(",")], default=[48, 64, 96, 128],
                    help="ensemble member grid sizes to draw from (comma separated)")
    ap.add_argument("--fp64", action="store_true")
    ap.add_argument("--we", type=float, default=5.0)
    ap.add_argument("--checkpoint", type=int, default=10, help="adjoint: checkpoint every K steps")
    a = ap.parse_args(argv)
    from droplet_amd import lattice as dl
    if a.what == "ensemble":
        ensemble(a, dl)
        return
    if a.what == "hit":
        turbulence(a, dl)
        return
    if a.what == "census":
        droplet_census(a, dl)
        return
    if a.what == "allen-cahn":
        allen_cahn(a, dl)
        return
    if a.what == "spectra":
        spectra(a, dl)
        return
    if a.what == "adjoint":
        adjoint(a, dl)
        return
    if a.what in ("noneq", "scale", "budget"):
        variants(a, dl)
        return
    if a.what == "diag":
        diagnostics(a, dl)
        return
    if a.what == "copy":
        gbs = dl.copy_bandwidth(a.mib << 20, a.reps)
        print(json.dumps({"copy_bytes": a.mib << 20, "reps": a.reps, "GB/s": round(gbs, 1)}))
        return
    if a.what == "yardsticks":
        names = ["gridstride", "memcpyDtoD", "unroll4_nt", "unroll4_plain", "read_only", "write_only", "soa20_nt"]
        for mib in (1024, 2048, 4096):
            out = {"MiB": mib}
            for k, name in enumerate(names):
                out[name + " GB/s"] = round(dl.copy_bandwidth_kind(mib << 20, k, a.reps), 1)
            print(json.dumps(out), flush=True)
        a.what, a.grid = "copysoa", 256
    if a.what == "copysoa":
        n = a.grid ** 3
        out = {"grid": a.grid}
        for k in (10, 20, 40):
            out[f"{k} streams GB/s"] = round(dl.copy_bandwidth_soa(8 * n if k <= 20 else 4 * n, k, a.reps), 1)
        print(json.dumps(out))
        return
    from bench import shear_config
    cfg, drops, _ = shear_config(a.grid, precision=1 if a.fp32 else 0)
    if a.what == "autotune":
        best, ms = dl.autotune(cfg, a.steps)
        print(json.dumps({"best_shape": best, "ms_per_step": ms}))
        return
    lat = dl.Lattice(cfg)
    if a.shape >= 0:
        lat.tile_shape = a.shape
    lat.init_droplets(drops)
    ms = lat.time_steps(a.steps)
    n = a.grid ** 3
    print(json.dumps({"grid": a.grid, "steps": a.steps, "shape": lat.tile_shape, "ms_per_step": ms / a.steps,
                      "MLUPS": n * a.steps / (ms * 1e-3) / 1e6}))


if __name__ == "__main__":
    main()
