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
    """Compile every HIP source into spectral2d_amd/libs2d.so (the
    translation units in parallel, then one link); returns its path."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    flags = [f for f in HIPCC_FLAGS if f != "-shared"]
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.basename(src).replace(".hip", ".o"))
        cmd = [hipcc(), *flags, "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                            text=True)))
        objs.append(obj)
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            failed.append(os.path.basename(src))
    if failed:
        raise RuntimeError(f"hipcc failed on {', '.join(failed)}")
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"link of libs2d.so failed ({r.returncode})")
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
