"""The reference's HDF5 output files, written and read through the HDF5 C
library (ctypes; serial HDF5, one process).

- ``REST_########.h5`` (src/IO.F90:306-424): dataset ``Header`` (null
  dataspace, native int) with scalar attributes ``nx``, ``ny``, ``nadv``
  (native int), ``time``, ``nu`` (native double); group ``FlowData`` with
  the physical padded-grid fields ``Omega``, ``Psi``, ``u``, ``v``, native
  double, Fortran dimensions (nxG, nyG) with x fastest (C dimensions
  (nyG, nxG)), chunked by the whole field (one process's block,
  src/IO.F90:439-496).
- ``REST_########.xmf`` (src/IO.F90:583-650), the Xdmf index of one restart.
- ``GRID.h5`` (src/IO.F90:158-285): ``Header`` with ``nx``, ``ny`` and
  group ``Domain_00001`` with ``x``, ``y`` at x_i = i 2 pi / nxG, i = 1..nxG.

:func:`write_restart` produces the fields as WriteRestart does
(src/IO.F90:98-143): the velocity by ComputeVelocity with the inverse
transform, omega and psi by TransformC2R. It works on copies, so the
context's state is left exactly as it was (the reference transforms omega
and psi back, a rounding-level change of the state). :func:`read_rest_file`
returns the header and fields; :func:`load_restart` puts omega and psi back
into a context with TransformR2C.

The library is ``libhdf5.so`` from ``$S2D_HDF5_LIB``, the conda prefix or
the loader path; without it these functions raise (nothing else in the
package needs HDF5).
"""
from __future__ import annotations

import ctypes as C
import ctypes.util
import math
import os

import numpy as np

_hid = C.c_int64
_hsize = C.c_uint64
H5F_ACC_RDONLY, H5F_ACC_TRUNC = 0x0000, 0x0002
H5P_DEFAULT = 0
H5S_ALL = 0
H5S_SCALAR, H5S_NULL = 0, 2

_lib = None


def _find() -> str:
    for c in (os.environ.get("S2D_HDF5_LIB"),
              os.path.join(os.environ.get("CONDA_PREFIX", "/opt/conda"), "lib", "libhdf5.so"),
              "/opt/conda/lib/libhdf5.so", ctypes.util.find_library("hdf5")):
        if c and os.path.exists(c):
            return c
    raise ImportError("libhdf5.so not found (set S2D_HDF5_LIB)")


def lib():
    global _lib
    if _lib is not None:
        return _lib
    L = C.CDLL(_find())
    P = C.POINTER
    sig = {
        "H5open": ([], C.c_int),
        "H5Fcreate": ([C.c_char_p, C.c_uint, _hid, _hid], _hid),
        "H5Fopen": ([C.c_char_p, C.c_uint, _hid], _hid),
        "H5Fclose": ([_hid], C.c_int),
        "H5Screate": ([C.c_int], _hid),
        "H5Screate_simple": ([C.c_int, P(_hsize), P(_hsize)], _hid),
        "H5Sclose": ([_hid], C.c_int),
        "H5Sget_simple_extent_ndims": ([_hid], C.c_int),
        "H5Sget_simple_extent_dims": ([_hid, P(_hsize), P(_hsize)], C.c_int),
        "H5Pcreate": ([_hid], _hid),
        "H5Pset_chunk": ([_hid, C.c_int, P(_hsize)], C.c_int),
        "H5Pclose": ([_hid], C.c_int),
        "H5Dcreate2": ([_hid, C.c_char_p, _hid, _hid, _hid, _hid, _hid], _hid),
        "H5Dopen2": ([_hid, C.c_char_p, _hid], _hid),
        "H5Dwrite": ([_hid, _hid, _hid, _hid, _hid, C.c_void_p], C.c_int),
        "H5Dread": ([_hid, _hid, _hid, _hid, _hid, C.c_void_p], C.c_int),
        "H5Dget_space": ([_hid], _hid),
        "H5Dclose": ([_hid], C.c_int),
        "H5Gcreate2": ([_hid, C.c_char_p, _hid, _hid, _hid], _hid),
        "H5Gclose": ([_hid], C.c_int),
        "H5Acreate2": ([_hid, C.c_char_p, _hid, _hid, _hid, _hid], _hid),
        "H5Aopen": ([_hid, C.c_char_p, _hid], _hid),
        "H5Awrite": ([_hid, _hid, C.c_void_p], C.c_int),
        "H5Aread": ([_hid, _hid, C.c_void_p], C.c_int),
        "H5Aclose": ([_hid], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    if L.H5open() < 0:
        raise RuntimeError("H5open failed")
    _lib = L
    return L


def _g(name: str) -> int:
    """An HDF5 global identifier (valid after H5open)."""
    return _hid.in_dll(lib(), name).value


def _ok(v, what):
    if v < 0:
        raise OSError(f"HDF5 call failed: {what}")
    return v


def _attr_write(obj, name: str, value, kind: str):
    L = lib()
    t = _g("H5T_NATIVE_INT_g") if kind == "int" else _g("H5T_NATIVE_DOUBLE_g")
    buf = C.c_int(int(value)) if kind == "int" else C.c_double(float(value))
    sp = _ok(L.H5Screate(H5S_SCALAR), "H5Screate")
    a = _ok(L.H5Acreate2(obj, name.encode(), t, sp, H5P_DEFAULT, H5P_DEFAULT), f"attribute {name}")
    _ok(L.H5Awrite(a, t, C.byref(buf)), f"write {name}")
    L.H5Aclose(a)
    L.H5Sclose(sp)


def _attr_read(obj, name: str, kind: str):
    L = lib()
    t = _g("H5T_NATIVE_INT_g") if kind == "int" else _g("H5T_NATIVE_DOUBLE_g")
    buf = C.c_int() if kind == "int" else C.c_double()
    a = _ok(L.H5Aopen(obj, name.encode(), H5P_DEFAULT), f"attribute {name}")
    _ok(L.H5Aread(a, t, C.byref(buf)), f"read {name}")
    L.H5Aclose(a)
    return buf.value


def _header(f, attrs):
    """The 'Header' dataset: null dataspace, native int, scalar attributes."""
    L = lib()
    sp = _ok(L.H5Screate(H5S_NULL), "H5Screate")
    d = _ok(L.H5Dcreate2(f, b"Header", _g("H5T_NATIVE_INT_g"), sp, H5P_DEFAULT, H5P_DEFAULT,
                         H5P_DEFAULT), "Header")
    for name, value, kind in attrs:
        _attr_write(d, name, value, kind)
    L.H5Dclose(d)
    L.H5Sclose(sp)


def _dataset(group, name: str, a: np.ndarray):
    """A 2D native-double dataset, C dimensions a.shape = (ny, nx), chunked
    by the whole array."""
    L = lib()
    a = np.ascontiguousarray(a, dtype=np.float64)
    dims = (_hsize * 2)(*a.shape)
    sp = _ok(L.H5Screate_simple(2, dims, None), "H5Screate_simple")
    pl = _ok(L.H5Pcreate(_g("H5P_CLS_DATASET_CREATE_ID_g")), "H5Pcreate")
    _ok(L.H5Pset_chunk(pl, 2, dims), "H5Pset_chunk")
    d = _ok(L.H5Dcreate2(group, name.encode(), _g("H5T_NATIVE_DOUBLE_g"), sp, H5P_DEFAULT, pl,
                         H5P_DEFAULT), name)
    _ok(L.H5Dwrite(d, _g("H5T_NATIVE_DOUBLE_g"), H5S_ALL, H5S_ALL, H5P_DEFAULT,
                   a.ctypes.data_as(C.c_void_p)), f"write {name}")
    L.H5Dclose(d)
    L.H5Pclose(pl)
    L.H5Sclose(sp)


def _read_dataset(f, path: str) -> np.ndarray:
    L = lib()
    d = _ok(L.H5Dopen2(f, path.encode(), H5P_DEFAULT), path)
    sp = L.H5Dget_space(d)
    nd = L.H5Sget_simple_extent_ndims(sp)
    dims = (_hsize * nd)()
    L.H5Sget_simple_extent_dims(sp, dims, None)
    out = np.empty(tuple(dims), dtype=np.float64)
    _ok(L.H5Dread(d, _g("H5T_NATIVE_DOUBLE_g"), H5S_ALL, H5S_ALL, H5P_DEFAULT,
                  out.ctypes.data_as(C.c_void_p)), f"read {path}")
    L.H5Sclose(sp)
    L.H5Dclose(d)
    return out


def rest_name(num: int, suffix: str = "h5") -> str:
    """REST_########.<suffix> (src/IO.F90:674-684)."""
    return f"REST_{int(num):08d}.{suffix}"


def write_rest_file(path: str, fields: dict, nadv: int, time: float, nu: float):
    """A REST file from physical fields {'Omega', 'Psi', 'u', 'v'}, each
    (nyG, nxG) with x fastest."""
    L = lib()
    ny, nx = np.shape(fields["Omega"])
    f = _ok(L.H5Fcreate(os.fsencode(path), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), path)
    try:
        _header(f, [("nx", nx, "int"), ("ny", ny, "int"), ("nadv", nadv, "int"),
                    ("time", time, "double"), ("nu", nu, "double")])
        g = _ok(L.H5Gcreate2(f, b"FlowData", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "FlowData")
        for name in ("Omega", "Psi", "u", "v"):
            _dataset(g, name, fields[name])
        L.H5Gclose(g)
    finally:
        L.H5Fclose(f)


def read_rest_file(path: str) -> dict:
    """Header attributes and the four fields of a REST file."""
    L = lib()
    f = _ok(L.H5Fopen(os.fsencode(path), H5F_ACC_RDONLY, H5P_DEFAULT), path)
    try:
        d = _ok(L.H5Dopen2(f, b"Header", H5P_DEFAULT), "Header")
        out = {"nx": _attr_read(d, "nx", "int"), "ny": _attr_read(d, "ny", "int"),
               "nadv": _attr_read(d, "nadv", "int"), "time": _attr_read(d, "time", "double"),
               "nu": _attr_read(d, "nu", "double")}
        L.H5Dclose(d)
        for name in ("Omega", "Psi", "u", "v"):
            out[name] = _read_dataset(f, "FlowData/" + name)
    finally:
        L.H5Fclose(f)
    return out


def write_grid_file(path: str, nxG: int, nyG: int):
    """GRID.h5: x(i, j) = i 2 pi / nxG, y(i, j) = j 2 pi / nyG, i, j from 1
    (src/IO.F90:158-285)."""
    L = lib()
    i = np.arange(1, nxG + 1, dtype=np.float64)
    j = np.arange(1, nyG + 1, dtype=np.float64)
    x = np.broadcast_to(i * 2.0 * math.pi / nxG, (nyG, nxG))
    y = np.broadcast_to((j * 2.0 * math.pi / nyG)[:, None], (nyG, nxG))
    f = _ok(L.H5Fcreate(os.fsencode(path), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), path)
    try:
        _header(f, [("nx", nxG, "int"), ("ny", nyG, "int")])
        g = _ok(L.H5Gcreate2(f, b"Domain_00001", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "group")
        _dataset(g, "x", x)
        _dataset(g, "y", y)
        L.H5Gclose(g)
    finally:
        L.H5Fclose(f)


def xmf_text(nxG: int, nyG: int, num: int) -> str:
    """REST_########.xmf as src/IO.F90:583-650 formats it (T edit
    descriptors as leading spaces, I4.4 dimensions)."""
    rest = rest_name(num)
    dims = f"{nxG:04d} {nyG:04d}"
    item = f'{" " * 12}<DataItem Dimensions="{dims}" NumberType="Float" Precision="8" Format="HDF">'
    lines = ['<?xml version="1.0" ?>', '<!DOCTYPE Xdmf SYSTEM "Xdmf.dtd" []>', '<Xdmf Version="2.1">',
             '   <Domain Name="Domain 1">', '      <Grid GridType="Uniform" Name="Block 1">',
             f'         <Topology TopologyType="2DSMesh" NumberOfElements="{dims}"/>',
             '         <Geometry GeometryType="X_Y">',
             item, "               GRID.h5:/Domain_00001/x", "            </DataItem>",
             item, "               GRID.h5:/Domain_00001/y", "            </DataItem>",
             "         </Geometry>"]
    for name, label in (("Omega", "Omega [1/s]"), ("Psi", "Psi [m^2/s]"), ("u", "u [m/s]"),
                        ("v", "v [m/s]")):
        lines += [f'         <Attribute AttributeType="Scalar" Center="Node" Name="{label}">', item,
                  f"               {rest}:/FlowData/{name}", "            </DataItem>",
                  "         </Attribute>"]
    lines += ["      </Grid>", "   </Domain>", "</Xdmf>"]
    return "\n".join(lines) + "\n"


def physical_fields(sim) -> dict:
    """Omega, Psi, u, v on the padded grid as WriteRestart computes them
    (src/IO.F90:118-121), on copies of the state: ComputeVelocity with the
    inverse transform, TransformC2R of omega and psi."""
    sim.copy_field("w", "wS")
    sim.copy_field("psi", "psiS")
    sim.ComputeVelocity("u", "v", "psi", trans=True)
    sim.TransformC2R("wS")
    sim.TransformC2R("psiS")
    nx = sim.nxG
    return {name: sim.rview(sim.get_field(slot))[:, :nx].copy()
            for name, slot in (("Omega", "wS"), ("Psi", "psiS"), ("u", "u"), ("v", "v"))}


def write_restart(sim, num: int, nu: float, directory: str = ".", xmf: bool = True) -> str:
    """REST_<num>.h5 (and its .xmf) of a context; returns the .h5 path."""
    t, nadv, _ = sim.get_time()
    path = os.path.join(directory, rest_name(num))
    write_rest_file(path, physical_fields(sim), nadv, t, nu)
    if xmf:
        with open(os.path.join(directory, rest_name(num, "xmf")), "w") as fh:
            fh.write(xmf_text(sim.nxG, sim.nyG, num))
    return path


def load_restart(sim, path: str) -> dict:
    """Omega and Psi of a REST file into the context's w and psi
    (TransformR2C of the physical fields), time and step count set from the
    header; returns the header."""
    r = read_rest_file(path)
    if (r["nx"], r["ny"]) != (sim.nxG, sim.nyG):
        raise ValueError(f"REST grid {r['nx']}x{r['ny']} != context {sim.nxG}x{sim.nyG}")
    for name, slot in (("Omega", "w"), ("Psi", "psi")):
        buf = sim.field()
        sim.rview(buf)[:, :sim.nxG] = r[name]
        sim.set_field(slot, buf)
        sim.TransformR2C(slot)
    sim.set_time(r["time"], r["nadv"])
    return {k: r[k] for k in ("nx", "ny", "nadv", "time", "nu")}


__all__ = ["rest_name", "write_rest_file", "read_rest_file", "write_grid_file", "xmf_text",
           "physical_fields", "write_restart", "load_restart"]
