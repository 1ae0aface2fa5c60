"""Build recipe for the reference Fortran programs used as the golden-vector source.

Test infrastructure only: nothing on the product path imports or runs this.

The reference (nakano-gendai/droplet) has no runtime configuration: grid size,
step count, output cadence, output directories and physical parameters are all
compile-time ``parameter`` constants in ``module globals`` (e.g.
program/lbm_drop_xy.f90:9-67, prototype_program/two-phase/dropch.f90:8-80).
This recipe therefore

1. reads the reference source where it lies under /root/reference,
2. rewrites ONLY named ``parameter`` lines (value to the right of ``=``) in a
   temporary copy created outside the repository,
3. compiles that copy with amdflang (ROCm) against MPICH from /opt/conda,
4. places the executable in ``oracle/_ref/`` (git-ignored) and deletes the copy.

No statement other than the named parameter lines is touched; a parameter that
does not match exactly one line aborts the build.  The executables need
``mpiexec`` (MPICH) to run; ``tests/golden/make_golden.py`` drives them.
"""
from __future__ import annotations

import hashlib
import os
import re
import shutil
import subprocess
import sys
import tempfile

REF_ROOT = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
OUT_DIR = os.path.join(HERE, "_ref")
FC = "/opt/rocm/llvm/bin/amdflang"
MPI_PREFIX = "/opt/conda"

_PARAM_RE = r"^(\s*(?:integer|real\(8\)|character\(\*\))\s*,\s*parameter\s*::\s*{name}\s*=\s*)([^!]*?)(\s*(?:!.*)?)$"


def sha256_of(path: str) -> str:
    with open(path, "rb") as fh:
        return hashlib.sha256(fh.read()).hexdigest()


def patch_parameters(text: str, params: dict[str, str]) -> str:
    """Replace the value of each named ``parameter`` constant; nothing else."""
    lines = text.split("\n")
    for name, value in params.items():
        rx = re.compile(_PARAM_RE.format(name=re.escape(name)), re.IGNORECASE)
        hits = [k for k, ln in enumerate(lines) if rx.match(ln)]
        if len(hits) != 1:
            raise RuntimeError(f"parameter {name!r}: expected exactly one declaration, found {len(hits)}")
        k = hits[0]
        lines[k] = rx.sub(lambda m: m.group(1) + value + m.group(3), lines[k])
    return "\n".join(lines)


def build(rel_source: str, exe_name: str, params: dict[str, str]) -> str:
    """Compile a parameter-patched copy of a reference program into oracle/_ref/."""
    src = os.path.join(REF_ROOT, rel_source)
    if not os.path.isfile(src):
        raise FileNotFoundError(src)
    if not os.path.exists(FC):
        raise FileNotFoundError(FC)
    os.makedirs(OUT_DIR, exist_ok=True)
    with open(src, "r", encoding="utf-8", errors="surrogateescape") as fh:
        text = fh.read()
    patched = patch_parameters(text, params)
    exe = os.path.join(OUT_DIR, exe_name)
    with tempfile.TemporaryDirectory(prefix="droplet_ref_") as tmp:
        copy = os.path.join(tmp, os.path.basename(rel_source))
        with open(copy, "w", encoding="utf-8", errors="surrogateescape") as fh:
            fh.write(patched)
        # -O2 without -ffast-math: no reassociation; the x86-64 baseline ISA has no FMA,
        # so every product and sum is rounded separately, as written.
        cmd = [FC, "-O2", "-fopenmp", f"-I{MPI_PREFIX}/include", copy,
               f"-L{MPI_PREFIX}/lib", f"-Wl,-rpath,{MPI_PREFIX}/lib", "-lmpifort", "-lmpi",
               "-o", exe]
        subprocess.run(cmd, cwd=tmp, check=True, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    with open(exe + ".params", "w") as fh:
        fh.write(f"source={rel_source}\nsha256={sha256_of(src)}\n")
        for k, v in params.items():
            fh.write(f"{k}={v}\n")
    return exe


def available() -> bool:
    return os.path.isdir(REF_ROOT) and os.path.exists(FC) and os.path.exists(os.path.join(MPI_PREFIX, "bin", "mpiexec"))



STANDARD = {
    # periodic droplet of program/lbm_drop_xy.f90 at 64^3 for CPU timing (cpu_baseline kind
    # "reference"): 8 ranks (Nx = 8 is hard-coded, :143), 300 steps, no field output
    # inside the timed steps.  nu = 0.1, sigma = 0.02 (kappa_g = 0.0117), D = 24: the
    # parameters of the periodic32 golden case, which stay finite (Re = 1.2 here gives
    # nu = 0.2, sigma = 0.04, and the run turns NaN before step 100).
    "lbm_drop_xy_bench64": ("program/lbm_drop_xy.f90", dict(
        xmax="63", ymax="63", zmax="63", step="300", step_input="100000000", step_input_file_num="0",
        datadir_input='"./in/"', datadir_output='"./out/"', datadir_output_fg='"./out/fg/"',
        step_output="100000000", step_putput_fg="100000000", Re="2.4d0", eta="1.0d0", Ca="0.05d0",
        D_vortex="64.0d0", D_ratio="0.375d0", umax="0.01d0",
        xc="0.5d0*dble(xmax)", yc="0.5d0*dble(ymax)", zc="0.5d0*dble(zmax)")),
}


def build_standard() -> list[str]:
    """Reference executables used outside the golden-vector generator."""
    return [build(src, name, params) for name, (src, params) in STANDARD.items()]


if __name__ == "__main__":  # pragma: no cover - manual use
    print("reference available:", available(), file=sys.stderr)
