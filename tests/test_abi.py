"""C-ABI checks that need no GPU: libs2d.so loads, exports every function
include/s2d.h declares, its host-only set-up agrees with the oracle, and
invalid requests fail with status codes (never a crash)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle
from spectral2d_amd import _lib as L
from spectral2d_amd.spectral import GetGridSize, Wavenumbers

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header="s2d.h"):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return set(re.findall(r"\b(s2d_\w+)\s*\(", txt))


def test_every_header_is_checked():
    assert sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h")) == \
        ["s2d.h", "s2d_h5.h"]


def test_library_exports_every_declared_function():
    lib = C.CDLL(L.LIB_PATH)
    decl = declared_functions("s2d.h")
    assert len(decl) >= 30
    missing = [n for n in sorted(decl) if not hasattr(lib, n)]
    assert not missing, missing
    assert decl <= set(L.EXPORTS), sorted(decl - set(L.EXPORTS))


def test_h5_library_exports_every_declared_function():
    from spectral2d_amd import build
    lib = C.CDLL(build.LIB_H5)
    decl = declared_functions("s2d_h5.h")
    assert decl == {"s2d_h5_write_rest", "s2d_h5_write_grid", "s2d_h5_write_xmf"}
    assert all(hasattr(lib, n) for n in decl)


@pytest.mark.parametrize("nx,ny,d", [(96, 96, 1), (1024, 1024, 1), (64, 32, 1), (16, 16, 0)])
def test_grid_matches_oracle(nx, ny, d):
    g = GetGridSize(nx, ny, d)
    o = oracle.Oracle(nx, ny, dealias=d)
    for f in ("nxG", "nyG", "kxLG", "kyLG", "kxMG", "kyMG", "kxLU", "kyLU", "kxMU", "kyMU",
              "rISize", "cISize"):
        assert getattr(g, f) == getattr(o.g, f), f
    assert g.kMax == o.g.kMax
    assert (g.nxP, g.nyP, g.j1, g.j2, g.allocLocal) == (g.nxG, g.nyG, 1, g.nyG,
                                                         g.cISize * g.nyG)
    kx, ky = Wavenumbers(g)
    assert np.array_equal(kx, o.kx) and np.array_equal(ky, o.ky)


def test_undefined_dealiasing_is_rejected():
    for d in (2, 3, 7):
        g = L.Grid()
        rc = L.lib().s2d_get_grid_size(32, 32, d, C.byref(g))
        assert rc == L.S2D_ERR_UNSUPPORTED
        assert b"dealias" in L.lib().s2d_last_error(None)


def test_create_rejects_before_touching_a_device():
    h = C.c_void_p()
    # odd grid, non 2^a 3^b padded size, bad precision, bad dealiasing
    for args in ((33, 33, 1, 0), (40, 40, 1, 0), (32, 32, 1, 9), (32, 32, 2, 0)):
        rc = L.lib().s2d_create(*args, C.byref(h))
        assert rc in (L.S2D_ERR_UNSUPPORTED, L.S2D_ERR_INVALID_ARG), (args, rc)
        assert not h.value


def test_null_context_is_an_error_not_a_crash():
    assert L.lib().s2d_rk3_step(None, 1e-3, 1e-3) == L.S2D_ERR_INVALID_ARG
    assert L.lib().s2d_transform_r2c(None, 0) == L.S2D_ERR_INVALID_ARG
    assert L.lib().s2d_destroy(None) == L.S2D_OK


def test_ensemble_rejects_before_touching_a_device():
    h = C.c_void_p()
    for sizes, prec in (([128, 100], 1), ([128, 2048], 1), ([1024], 0), ([128], 7), ([], 1)):
        arr = (C.c_int * max(1, len(sizes)))(*sizes)
        rc = L.lib().s2d_ensemble_create(len(sizes), arr, prec, C.byref(h))
        assert rc in (L.S2D_ERR_UNSUPPORTED, L.S2D_ERR_INVALID_ARG), (sizes, prec, rc)
        assert not h.value
    assert L.lib().s2d_ensemble_steps(None, 1, 0.0, 0.5, 0) == L.S2D_ERR_INVALID_ARG
    assert L.lib().s2d_ensemble_destroy(None) == L.S2D_OK
    assert b"null ensemble" in L.lib().s2d_ensemble_last_error(None)


def test_cnn_create_rejects_before_touching_a_device():
    h = C.c_void_p()
    # W not a multiple of 128, H not a multiple of 8, no batch, and a batch
    # of 2^31 pixels or more (the kernels index pixels in 32 bits)
    for args in ((256, 200, 4), (252, 256, 4), (256, 256, 0), (256, 256, 1 << 15)):
        rc = L.lib().s2d_cnn_create(*args, C.byref(h))
        assert rc == L.S2D_ERR_UNSUPPORTED, (args, rc)
        assert not h.value
