"""Build the Fortran drop-in (ISO_C_BINDING modules + driver) with amdflang.

Produces spectral2d_amd/fortran/spectral2d linked against
spectral2d_amd/libs2d.so (rpath $ORIGIN/..). Module files and objects go
to spectral2d_amd/fortran/_build.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
FDIR = os.path.join(HERE, "fortran")
BUILD = os.path.join(FDIR, "_build")
EXE = os.path.join(FDIR, "spectral2d")
# dependency order of the modules
SOURCES = ["Parameters.F90", "s2d_c.F90", "Spectral.F90", "TimeIntegration.F90",
           "Spectral2D.F90"]


def compiler() -> str | None:
    for c in (os.environ.get("FC"), "/opt/rocm/bin/amdflang", shutil.which("amdflang"),
              shutil.which("flang")):
        if c and os.path.exists(c):
            return c
    return None


def build(verbose: bool = False) -> str:
    fc = compiler()
    if fc is None:
        raise RuntimeError("no Fortran compiler (amdflang) found")
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = os.path.join(BUILD, src.replace(".F90", ".o"))
        cmd = [fc, "-O2", "-c", os.path.join(FDIR, src), "-o", obj, "-module-dir", BUILD,
               "-I", BUILD]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"{fc} failed on {src}")
        objs.append(obj)
    cmd = [fc, *objs, "-o", EXE + ".tmp", "-L", HERE, "-ls2d", "-Wl,-rpath,$ORIGIN/.."]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link of the Fortran driver failed")
    os.replace(EXE + ".tmp", EXE)
    return EXE


if __name__ == "__main__":
    print(build(verbose=True))
