"""Build the Fortran drop-in (ISO_C_BINDING modules + driver) with amdflang.

Produces spectral2d_amd/fortran/spectral2d (the driver) and
spectral2d_routines (a harness applying every Spectral_m routine once to
fields from a file), linked against
spectral2d_amd/libs2d.so and libs2d_h5.so (rpath $ORIGIN/..; libs2d_h5
carries the HDF5 library's own rpath). Module files and objects go
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
EXE_ROUTINES = os.path.join(FDIR, "spectral2d_routines")
# dependency order of the modules; the two programs come last
MODULES = ["Parameters.F90", "s2d_c.F90", "Spectral.F90", "TimeIntegration.F90", "IO.F90",
           "Analysis.F90"]
PROGRAMS = {"Spectral2D.F90": EXE, "Routines.F90": EXE_ROUTINES}


def compiler() -> str | None:
    for c in (os.environ.get("FC"), "/opt/rocm/bin/amdflang", shutil.which("amdflang"),
              shutil.which("flang")):
        if c and os.path.exists(c):
            return c
    return None


def _run(cmd, verbose, what):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(what)


def build(verbose: bool = False) -> str:
    """Compile the modules and link the driver and the routine harness;
    returns the driver's path."""
    fc = compiler()
    if fc is None:
        raise RuntimeError("no Fortran compiler (amdflang) found")
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in MODULES + list(PROGRAMS):
        obj = os.path.join(BUILD, src.replace(".F90", ".o"))
        _run([fc, "-O2", "-c", os.path.join(FDIR, src), "-o", obj, "-module-dir", BUILD,
              "-I", BUILD], verbose, f"{fc} failed on {src}")
        objs.append(obj)
    mods = objs[:len(MODULES)]
    for src, exe in PROGRAMS.items():
        main = os.path.join(BUILD, src.replace(".F90", ".o"))
        _run([fc, *mods, main, "-o", exe + ".tmp", "-L", HERE, "-ls2d", "-ls2d_h5",
              "-Wl,-rpath,$ORIGIN/.."], verbose, f"link of {os.path.basename(exe)} failed")
        os.replace(exe + ".tmp", exe)
    return EXE


if __name__ == "__main__":
    print(build(verbose=True))
