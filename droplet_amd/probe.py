"""Small GPU drivers for profiling runs (rocprofv3) of single kernels.

    python -m droplet_amd.probe copy [--mib 1024] [--reps 20]
        stream copy of a known size: the calibration kernel for FETCH_SIZE/WRITE_SIZE
    python -m droplet_amd.probe step [--grid 256] [--steps 20] [--fp32] [--shape S]
        the 256^3 shear-droplet step kernel of bench.py, a few launches
"""
from __future__ import annotations

import argparse
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["copy", "step", "autotune"])
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--fp32", action="store_true")
    ap.add_argument("--shape", type=int, default=-1)
    a = ap.parse_args(argv)
    from droplet_amd import lattice as dl
    if a.what == "copy":
        gbs = dl.copy_bandwidth(a.mib << 20, a.reps)
        print(json.dumps({"copy_bytes": a.mib << 20, "reps": a.reps, "GB/s": round(gbs, 1)}))
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
