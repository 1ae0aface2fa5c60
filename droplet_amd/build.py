"""Build the HIP library libdroplet_amd.so in-tree (droplet_amd/_lib/) with hipcc for gfx950.

    python -m droplet_amd.build          # or __graft_entry__.build()

Compiled with FMA contraction off (-ffp-contract=off): the FP64 kernels round
every product and sum on its own, in the reference's order.  Packed FP32
instructions are disabled for all device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(OUT_DIR, "libdroplet_amd.so")
ARCH = os.environ.get("DROPLET_ARCH", "gfx950")

SOURCES = ["step_kernels.hip", "step_fp32.hip", "aux_kernels.hip", "aux_fp32.hip", "adjoint.hip", "spectral.hip", "probe.hip",
           "census.hip", "noalloc.hip", "spectral_diag.hip", "budget.hip",
           "lattice.cpp", "deformation.cpp", "fp32_shifted.cpp", "batch.cpp", "adjoint.cpp", "ref_files.cpp"]
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function", "-Wno-pass-failed",
          "-Wno-unused-command-line-argument", "-Wno-unknown-cuda-version"]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise FileNotFoundError("hipcc not found")


# the step kernels are compiled once per tile shape (-DDRP_STEP_SHAPE=k) plus once
# for their dispatcher, so that the template instances build in parallel
PER_SHAPE = ("step_kernels.hip", "step_fp32.hip")
SHAPES = 7


def units():
    """(source, extra flags, object name) of every translation unit."""
    out = []
    for src in SOURCES:
        out.append((src, [], src + ".o"))
        if src in PER_SHAPE:
            out += [(src, [f"-DDRP_STEP_SHAPE={k}"], f"{src}.{k}.o") for k in range(SHAPES)]
    return out


def build(verbose: bool = False, jobs: int = 8) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    cc = hipcc()
    with tempfile.TemporaryDirectory(prefix="droplet_build_") as tmp:
        objs, procs = [], []
        for src, extra, oname in units():
            path = os.path.join(CSRC, src)
            obj = os.path.join(tmp, oname)
            cmd = [cc, f"--offload-arch={ARCH}", "-x", "hip", *COMMON, *DEVICE_FLAGS, *extra,
                   f"-I{CSRC}", "-c", path, "-o", obj]
            if verbose:
                print(" ".join(cmd), file=sys.stderr)
            procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
            objs.append(obj)
            if len(procs) >= jobs:
                _wait(procs.pop(0))
        while procs:
            _wait(procs.pop(0))
        tmp_lib = LIB + ".tmp"
        link = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp_lib]
        r = subprocess.run(link, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if r.returncode:
            raise RuntimeError("link failed:\n" + r.stdout)
        os.replace(tmp_lib, LIB)
    return LIB


FORTRAN = os.path.join(HERE, "fortran")
FC = "/opt/rocm/llvm/bin/amdflang"


def build_fortran(verbose: bool = False) -> str | None:
    """Compile the iso_c_binding module and the shear driver against the library
    (droplet_amd/_lib/shear_driver); skipped when amdflang is absent."""
    if not os.path.exists(FC):
        return None
    exe = os.path.join(OUT_DIR, "shear_driver")
    with tempfile.TemporaryDirectory(prefix="droplet_fortran_") as tmp:
        cmd = [FC, "-O2", f"-J{tmp}", os.path.join(FORTRAN, "droplet_amd.f90"), os.path.join(FORTRAN, "shear_driver.f90"),
               f"-L{OUT_DIR}", "-ldroplet_amd", f"-Wl,-rpath,$ORIGIN", "-o", exe + ".tmp"]
        if verbose:
            print(" ".join(cmd), file=sys.stderr)
        r = subprocess.run(cmd, cwd=tmp, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if r.returncode:
            raise RuntimeError("Fortran driver build failed:\n" + r.stdout)
    os.replace(exe + ".tmp", exe)
    return exe


def _wait(item):
    src, p = item
    out, _ = p.communicate()
    if p.returncode:
        raise RuntimeError(f"compile of {src} failed:\n{out.decode(errors='replace')}")
    if out and out.strip():
        text = out.decode(errors="replace")
        # the host pass does not know the device-only target feature; that note is harmless
        lines = [ln for ln in text.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
        if lines:
            print(f"[{src}]\n" + "\n".join(lines), file=sys.stderr)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_fortran(verbose="-v" in sys.argv))
