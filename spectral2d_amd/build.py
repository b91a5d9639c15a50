"""Build libs2d.so (hand-written HIP for gfx950) in-tree with hipcc, and
libs2d_h5.so (the serial HDF5 writers of the reference's output files, C)
with gcc against the image's HDF5.

Usage: python -m spectral2d_amd.build  (or spectral2d_amd.build.build()).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libs2d.so")
LIB_H5 = os.path.join(HERE, "libs2d_h5.so")
SOURCES = [os.path.join(CSRC, "s2d_api.hip"), os.path.join(CSRC, "cnn.hip")]
ARCH = os.environ.get("S2D_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 instructions stay off (the compiler would otherwise form
    # them from ordinary float code)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    "-I", os.path.join(ROOT, "include"),
    "-I", CSRC,
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    """Compile every HIP source into spectral2d_amd/libs2d.so; returns its path."""
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, *SOURCES, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed ({r.returncode})")
    os.replace(tmp, LIB)
    return LIB


def hdf5_prefix() -> str:
    """Prefix holding include/hdf5.h and lib/libhdf5.so ($S2D_HDF5_PREFIX,
    else the conda prefix)."""
    for p in (os.environ.get("S2D_HDF5_PREFIX"), os.environ.get("CONDA_PREFIX"), "/opt/conda"):
        if p and os.path.exists(os.path.join(p, "include", "hdf5.h")) and \
                os.path.exists(os.path.join(p, "lib", "libhdf5.so")):
            return p
    raise RuntimeError("HDF5 (include/hdf5.h, lib/libhdf5.so) not found; set S2D_HDF5_PREFIX")


def build_h5(verbose: bool = False) -> str:
    """Compile csrc/s2d_h5.c into spectral2d_amd/libs2d_h5.so; returns its path."""
    pre = hdf5_prefix()
    tmp = LIB_H5 + ".tmp"
    cmd = [os.environ.get("CC", "gcc"), "-O2", "-fPIC", "-shared", "-std=c99", "-Wall",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(pre, "include"),
           os.path.join(CSRC, "s2d_h5.c"), "-o", tmp, "-L", os.path.join(pre, "lib"), "-lhdf5",
           "-Wl,-rpath," + os.path.join(pre, "lib")]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build of libs2d_h5.so failed ({r.returncode})")
    os.replace(tmp, LIB_H5)
    return LIB_H5


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_h5(verbose=True))
