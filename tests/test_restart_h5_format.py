"""The REST_########.h5 / GRID.h5 / .xmf writers and reader on host arrays
(no GPU): the layout of src/IO.F90:158-285, 306-496, 583-650, read back
through the same C library and, where the image has it, checked against
h5dump's own listing."""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

h5 = pytest.importorskip("spectral2d_amd.restart_h5")
try:
    h5.lib()
except ImportError as e:  # pragma: no cover
    pytest.skip(str(e), allow_module_level=True)


def _fields(nx, ny, seed=0):
    rng = np.random.default_rng(seed)
    return {k: rng.standard_normal((ny, nx)) for k in ("Omega", "Psi", "u", "v")}


def test_rest_round_trip(tmp_path):
    f = _fields(24, 18)
    p = tmp_path / h5.rest_name(7)
    assert p.name == "REST_00000007.h5"
    h5.write_rest_file(str(p), f, nadv=123, time=0.625, nu=1e-4)
    r = h5.read_rest_file(str(p))
    assert (r["nx"], r["ny"], r["nadv"]) == (24, 18, 123)
    assert r["time"] == 0.625 and r["nu"] == 1e-4
    for k, v in f.items():
        assert r[k].shape == (18, 24)
        assert np.array_equal(r[k], v)


@pytest.mark.skipif(shutil.which("h5dump") is None and not os.path.exists("/opt/conda/bin/h5dump"),
                    reason="no h5dump in this image")
def test_rest_layout_per_h5dump(tmp_path):
    """h5dump sees the reference's structure: Header (null dataspace, int)
    with five scalar attributes, FlowData/{Omega,Psi,u,v} doubles of C
    dimensions (nyG, nxG)."""
    p = tmp_path / h5.rest_name(1)
    h5.write_rest_file(str(p), _fields(12, 10), nadv=5, time=0.5, nu=2e-5)
    exe = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    out = subprocess.run([exe, "-H", str(p)], capture_output=True, text=True, check=True).stdout
    assert 'DATASET "Header"' in out and "DATASPACE  NULL" in out
    for a in ("nx", "ny", "nadv"):
        assert f'ATTRIBUTE "{a}"' in out
    for a in ("time", "nu"):
        assert f'ATTRIBUTE "{a}"' in out
    assert 'GROUP "FlowData"' in out
    for d in ("Omega", "Psi", "u", "v"):
        assert f'DATASET "{d}"' in out
    assert out.count("H5T_IEEE_F64LE") >= 6  # four fields, time, nu
    assert out.count("SIMPLE { ( 10, 12 ) / ( 10, 12 ) }") == 4


def test_grid_file(tmp_path):
    p = tmp_path / "GRID.h5"
    h5.write_grid_file(str(p), 8, 6)
    L = h5.lib()
    f = L.H5Fopen(os.fsencode(str(p)), h5.H5F_ACC_RDONLY, h5.H5P_DEFAULT)
    x = h5._read_dataset(f, "Domain_00001/x")
    y = h5._read_dataset(f, "Domain_00001/y")
    L.H5Fclose(f)
    assert x.shape == (6, 8) and y.shape == (6, 8)
    assert x[0, 0] == 2 * math.pi / 8 and x[3, 7] == 8 * 2 * math.pi / 8
    assert y[0, 5] == 2 * math.pi / 6 and y[5, 0] == 6 * 2 * math.pi / 6


def test_xmf_text():
    t = h5.xmf_text(1536, 1536, 3)
    lines = t.splitlines()
    assert lines[0] == '<?xml version="1.0" ?>'
    assert lines[3] == '   <Domain Name="Domain 1">'
    assert lines[5] == '         <Topology TopologyType="2DSMesh" NumberOfElements="1536 1536"/>'
    assert "               REST_00000003.h5:/FlowData/Omega" in lines
    assert "               GRID.h5:/Domain_00001/y" in lines
    assert lines[-1] == "</Xdmf>"


# ---- libs2d_h5 (the writers behind the Fortran IO_m): the same files


def _h5c():
    import ctypes as C
    from spectral2d_amd import build
    if not os.path.exists(build.LIB_H5):
        pytest.skip("libs2d_h5.so not built")
    L = C.CDLL(build.LIB_H5)
    dp = C.POINTER(C.c_double)
    L.s2d_h5_write_rest.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                    C.c_double, dp, dp, dp, dp]
    L.s2d_h5_write_grid.argtypes = [C.c_char_p, C.c_int, C.c_int]
    L.s2d_h5_write_xmf.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int]
    return L, dp


def test_c_rest_writer_matches_python_reader(tmp_path):
    """Padded real views (rISize, nyG): only the first nxG of each row land
    in the file, which the Python reader returns unchanged."""
    L, dp = _h5c()
    nxG, nyG, rI = 24, 18, 26
    f = _fields(nxG, nyG, seed=3)
    padded = {}
    for k, v in f.items():
        a = np.full((nyG, rI), np.nan)
        a[:, :nxG] = v
        padded[k] = np.ascontiguousarray(a)
    p = tmp_path / "REST_00000004.h5"
    rc = L.s2d_h5_write_rest(os.fsencode(str(p)), nxG, nyG, rI, 77, 1.25, 3e-4,
                             *[padded[k].ctypes.data_as(dp) for k in ("Omega", "Psi", "u", "v")])
    assert rc == 0
    r = h5.read_rest_file(str(p))
    assert (r["nx"], r["ny"], r["nadv"], r["time"], r["nu"]) == (nxG, nyG, 77, 1.25, 3e-4)
    for k, v in f.items():
        assert np.array_equal(r[k], v)
    # invalid sizes are refused, not written
    assert L.s2d_h5_write_rest(os.fsencode(str(p)), nxG, nyG, nxG - 1, 0, 0.0, 0.0,
                               *[padded[k].ctypes.data_as(dp) for k in ("Omega", "Psi", "u",
                                                                        "v")]) == 1


def test_c_grid_and_xmf_match_python(tmp_path):
    L, _ = _h5c()
    pc, pp = tmp_path / "GRID_c.h5", tmp_path / "GRID_py.h5"
    assert L.s2d_h5_write_grid(os.fsencode(str(pc)), 8, 6) == 0
    h5.write_grid_file(str(pp), 8, 6)
    H = h5.lib()
    for name in ("Domain_00001/x", "Domain_00001/y"):
        a, b = [H.H5Fopen(os.fsencode(str(q)), h5.H5F_ACC_RDONLY, h5.H5P_DEFAULT)
                for q in (pc, pp)]
        xa, xb = h5._read_dataset(a, name), h5._read_dataset(b, name)
        H.H5Fclose(a)
        H.H5Fclose(b)
        assert np.array_equal(xa, xb)
    px = tmp_path / "REST_00000003.xmf"
    assert L.s2d_h5_write_xmf(os.fsencode(str(px)), 1536, 1536, 3) == 0
    assert px.read_text() == h5.xmf_text(1536, 1536, 3)
