"""CPU FP64 oracle for the Spectral2D TVD-RK3 time step.

TEST INFRASTRUCTURE ONLY: the checker for the HIP library. Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may import it.
The product path (``spectral2d_amd``) never does.

The C restatement lives in ``s2d_oracle.c`` (built by ``oracle/Makefile``
into ``libs2d_oracle.so``); this module wraps it with numpy buffers in the
reference layout: one float64 buffer of ``2*cISize*nyG`` doubles per field,
viewed as complex ``(nyG, cISize)`` or real ``(nyG, rISize)`` row-major
arrays (row j = the reference's Fortran column j; src/Spectral2D.F90:293-304).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libs2d_oracle.so")

NO_DEALIASING = 0
DEALIAS_3_2_RULE = 1
DEALIAS_2_3_RULE = 2
DEALIAS_GRID_SHIFT = 3
RK3_TVD_SHU = 1


class _Grid(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("dealias", C.c_int),
        ("nxG", C.c_int), ("nyG", C.c_int),
        ("kxLG", C.c_int), ("kyLG", C.c_int), ("kxMG", C.c_int), ("kyMG", C.c_int),
        ("kxLU", C.c_int), ("kyLU", C.c_int), ("kxMU", C.c_int), ("kyMU", C.c_int),
        ("kMax", C.c_double),
        ("rISize", C.c_int), ("cISize", C.c_int),
    ]


_lib = None


def build() -> str:
    """Compile the oracle with its Makefile (gcc); returns the .so path."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.POINTER
        dp, ip = P(C.c_double), P(C.c_int)
        vp = C.c_void_p
        L.or_get_grid_size.argtypes = [C.c_int, C.c_int, C.c_int, P(_Grid)]
        L.or_wavenumbers.argtypes = [P(_Grid), ip, ip]
        L.or_create.restype = vp
        L.or_destroy.argtypes = [vp]
        L.or_set_threads.argtypes = [C.c_int]
        L.or_setup_mask.argtypes = [vp, C.c_int, C.c_int, ip, ip, C.c_int, C.c_int]
        L.or_get_mask.argtypes = [vp, ip, ip, ip, ip]
        L.or_setup_truncation.argtypes = [vp, C.c_int, C.c_int, C.c_double, ip, ip]
        L.or_get_truncation.argtypes = [vp, ip]
        L.or_setup_plans.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, dp]
        L.or_transform_r2c.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, dp]
        L.or_transform_c2r.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp]
        L.or_dudx.argtypes = [vp, C.c_int, C.c_int, ip, ip, dp, dp]
        L.or_dudy.argtypes = [vp, C.c_int, C.c_int, ip, ip, dp, dp]
        L.or_truncate.argtypes = [vp, C.c_int, C.c_int, dp]
        L.or_compute_psi.argtypes = [C.c_int, C.c_int, ip, ip, dp, dp]
        L.or_compute_velocity.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int,
                                          ip, ip, dp, dp, dp, C.c_int]
        L.or_compute_vorticity.argtypes = [vp, C.c_int, C.c_int, C.c_int, ip, ip,
                                           dp, dp, dp]
        L.or_compute_rhs.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_int, ip, ip, C.c_double, dp, dp, dp, dp, dp]
        L.or_time_integration_setup.argtypes = [vp, C.c_int, C.c_int, C.c_int,
                                                C.c_int, C.c_int]
        L.or_integrate_one_step.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int,
                                            C.c_int, C.c_int, ip, ip, C.c_double,
                                            C.c_double, dp, dp, dp, dp]
        L.or_set_taylor_green.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int,
                                          C.c_int, C.c_int, C.c_int, C.c_int, dp, dp]
        L.or_uniform.argtypes = [C.c_uint64, C.c_int, C.c_int]
        L.or_uniform.restype = C.c_double
        L.or_set_turbulence.argtypes = [P(_Grid), ip, ip, C.c_uint64, C.c_double,
                                        C.c_double, dp, dp]
        L.or_cfl_dt.argtypes = [vp, P(_Grid), ip, ip, C.c_double, dp, dp, dp]
        L.or_cfl_dt.restype = C.c_double
        L.or_spectrum_bins.argtypes = [C.c_int, C.c_int]
        L.or_compute_spectrum.argtypes = [C.c_double, C.c_int, C.c_int, ip, ip,
                                          C.c_int, C.c_int, dp, dp, dp, ip, dp, dp, dp]
        L.or_fft_c2c.argtypes = [C.c_int, C.c_int, dp, dp]
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_int))


def set_threads(n: int) -> int:
    return lib().or_set_threads(int(n))


def fft_c2c(x: np.ndarray, sign: int = -1) -> np.ndarray:
    """The oracle's own 1D complex FFT (unnormalized)."""
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    rc = lib().or_fft_c2c(len(x), int(sign), _dp(x.view(np.float64)),
                          _dp(out.view(np.float64)))
    if rc != 0:
        raise ValueError(f"length {len(x)} is not 2^a 3^b")
    return out


def uniform(seed: int, kx: int, ky: int) -> float:
    return lib().or_uniform(seed, kx, ky)


def spectrum_bins(kxMG: int, kyMG: int) -> int:
    return lib().or_spectrum_bins(kxMG, kyMG)


class Oracle:
    """One simulation's Spectral_m + TimeIntegration_m state, set up in the
    reference's init order (src/Spectral2D.F90:263-307, P = 1)."""

    def __init__(self, nx: int, ny: int | None = None,
                 dealias: int = DEALIAS_3_2_RULE):
        ny = nx if ny is None else ny
        L = lib()
        g = _Grid()
        if L.or_get_grid_size(nx, ny, dealias, C.byref(g)) != 0:
            raise ValueError(f"dealiasing method {dealias} is not defined")
        self.g = g
        self.nx, self.ny = nx, ny
        self.nxG, self.nyG = g.nxG, g.nyG
        self.rISize, self.cISize = g.rISize, g.cISize
        self.kMax = g.kMax
        self.kx = np.zeros(g.cISize, np.int32)
        self.ky = np.zeros(g.nyG, np.int32)
        L.or_wavenumbers(C.byref(g), _ip(self.kx), _ip(self.ky))
        self.ctx = C.c_void_p(L.or_create())
        L.or_setup_mask(self.ctx, g.cISize, g.nyG, _ip(self.kx), _ip(self.ky),
                        g.kxMG, g.kyMG)
        L.or_setup_truncation(self.ctx, g.cISize, g.nyG, g.kMax, _ip(self.kx),
                              _ip(self.ky))
        scratch = self.field()
        L.or_setup_plans(self.ctx, g.nxG, g.nyG, g.nyG, g.rISize, g.cISize,
                         _dp(scratch))
        L.or_time_integration_setup(self.ctx, RK3_TVD_SHU, g.rISize, g.cISize,
                                    g.nyG, g.cISize * g.nyG)

    def __del__(self):
        try:
            lib().or_destroy(self.ctx)
        except Exception:
            pass

    # ---- buffers and views
    def field(self) -> np.ndarray:
        return np.zeros(2 * self.cISize * self.nyG, np.float64)

    def cview(self, buf: np.ndarray) -> np.ndarray:
        return buf.view(np.complex128).reshape(self.nyG, self.cISize)

    def rview(self, buf: np.ndarray) -> np.ndarray:
        return buf.reshape(self.nyG, self.rISize)

    def from_complex(self, c: np.ndarray) -> np.ndarray:
        c = np.ascontiguousarray(c, dtype=np.complex128).reshape(self.nyG, self.cISize)
        return c.reshape(-1).view(np.float64).copy()

    def from_real(self, r: np.ndarray) -> np.ndarray:
        buf = self.field()
        self.rview(buf)[:, : self.nxG] = r
        return buf

    # ---- Spectral_m
    def mask(self):
        a = [C.c_int() for _ in range(4)]
        lib().or_get_mask(self.ctx, *[C.byref(x) for x in a])
        return tuple(x.value for x in a)

    def truncation(self) -> np.ndarray:
        i1 = np.zeros(self.nyG, np.int32)
        lib().or_get_truncation(self.ctx, _ip(i1))
        return i1

    def kept_mask(self) -> np.ndarray:
        """Boolean (nyG, cISize): True where Truncate keeps the mode."""
        i1 = self.truncation()
        return np.arange(self.cISize)[None, :] < i1[:, None]

    def transform_r2c(self, buf):
        lib().or_transform_r2c(self.ctx, self.nxG, self.nyG, self.nyG,
                               self.rISize, self.cISize, _dp(buf))

    def transform_c2r(self, buf):
        lib().or_transform_c2r(self.ctx, self.rISize, self.cISize, self.nyG, _dp(buf))

    def dudx(self, src, dst):
        lib().or_dudx(self.ctx, self.cISize, self.nyG, _ip(self.kx), _ip(self.ky),
                      _dp(src), _dp(dst))

    def dudy(self, src, dst):
        lib().or_dudy(self.ctx, self.cISize, self.nyG, _ip(self.kx), _ip(self.ky),
                      _dp(src), _dp(dst))

    def truncate(self, buf):
        lib().or_truncate(self.ctx, self.cISize, self.nyG, _dp(buf))

    def compute_psi(self, w, psi):
        lib().or_compute_psi(self.cISize, self.nyG, _ip(self.kx), _ip(self.ky),
                             _dp(w), _dp(psi))

    def compute_velocity(self, u, v, psi, trans=False):
        lib().or_compute_velocity(self.ctx, self.nxG, self.nyG, self.rISize,
                                  self.cISize, _ip(self.kx), _ip(self.ky), _dp(u),
                                  _dp(v), _dp(psi), int(bool(trans)))

    def compute_vorticity(self, u, v, w):
        lib().or_compute_vorticity(self.ctx, self.nxG, self.nyG, self.cISize,
                                   _ip(self.kx), _ip(self.ky), _dp(u), _dp(v), _dp(w))

    def compute_rhs(self, nu, u, v, w, psi, dW):
        lib().or_compute_rhs(self.ctx, self.nxG, self.nyG, self.nxG, self.nyG,
                             self.rISize, self.cISize, _ip(self.kx), _ip(self.ky),
                             float(nu), _dp(u), _dp(v), _dp(w), _dp(psi), _dp(dW))

    # ---- TimeIntegration_m
    def integrate_one_step(self, nu, dt, u, v, w, psi):
        lib().or_integrate_one_step(self.ctx, self.nxG, self.nyG, self.nxG,
                                    self.nyG, self.rISize, self.cISize,
                                    _ip(self.kx), _ip(self.ky), float(nu),
                                    float(dt), _dp(u), _dp(v), _dp(w), _dp(psi))

    # ---- initial conditions and diagnostics
    def set_taylor_green(self, w, psi):
        lib().or_set_taylor_green(self.ctx, self.nxG, self.nyG, self.nxG,
                                  self.nyG, 1, self.nyG, self.rISize, self.cISize,
                                  _dp(w), _dp(psi))

    def set_turbulence(self, w, psi, seed=1, k0=10.0, ke=0.5):
        lib().or_set_turbulence(C.byref(self.g), _ip(self.kx), _ip(self.ky),
                                int(seed), float(k0), float(ke), _dp(w), _dp(psi))

    def cfl_dt(self, psi, cfl=0.5) -> float:
        u, v = self.field(), self.field()
        return lib().or_cfl_dt(self.ctx, C.byref(self.g), _ip(self.kx),
                               _ip(self.ky), float(cfl), _dp(psi), _dp(u), _dp(v))

    def spectrum(self, u, v, nu=0.0):
        nb = spectrum_bins(self.g.kxMG, self.g.kyMG)
        ek = np.zeros(nb, np.float64)
        cnt = np.zeros(nb, np.int32)
        t, e, z = C.c_double(), C.c_double(), C.c_double()
        lib().or_compute_spectrum(float(nu), self.cISize, self.nyG, _ip(self.kx),
                                  _ip(self.ky), self.g.kxMG, self.g.kyMG, _dp(u),
                                  _dp(v), _dp(ek), _ip(cnt), C.byref(t),
                                  C.byref(e), C.byref(z))
        return ek, cnt, t.value, e.value, z.value

    # ---- convenience: a run in the driver's order (src/Spectral2D.F90:338-350)
    def new_state(self):
        return {k: self.field() for k in ("u", "v", "w", "psi")}

    def step(self, st, nu, dt):
        self.integrate_one_step(nu, dt, st["u"], st["v"], st["w"], st["psi"])
