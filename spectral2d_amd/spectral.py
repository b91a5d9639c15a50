"""Host-side mirror of the reference's Spectral_m / TimeIntegration_m
interfaces over libs2d (src/Spectral.F90:83-88, src/TimeIntegration.F90:67-68).

A :class:`Spectral2D` object is one simulation: it replaces the module
singletons (plans, masks, truncation table, RK scratch) with one libs2d
context. Fields live on the GPU in the reference layout; host buffers are
float64 numpy arrays of ``2*cISize*nyG`` doubles (complex ``(nyG, cISize)`` /
real ``(nyG, rISize)`` row-major views, i.e. the Fortran ``(cISize, nyP)`` /
``(rISize, nyP)`` column-major arrays).

Errors raise :class:`S2DError` (the reference has no error returns; its
failures are fatal).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import (DEALIAS_3_2_RULE, FIELD_DW, FIELD_PSI, FIELD_PSIS, FIELD_TMP,
                   FIELD_U, FIELD_V, FIELD_W, FIELD_WS, FP32, FP64, S2DError, check)

FIELDS = {"w": FIELD_W, "psi": FIELD_PSI, "u": FIELD_U, "v": FIELD_V,
          "wS": FIELD_WS, "psiS": FIELD_PSIS, "dW": FIELD_DW, "tmp": FIELD_TMP}


def _fid(f) -> int:
    if isinstance(f, str):
        return FIELDS[f]
    return int(f)


def GetGridSize(nx: int, ny: int, dealias: int = DEALIAS_3_2_RULE) -> L.Grid:
    """GetGridSize + GetAllocSize for one process (src/Spectral.F90:117-219,
    src/Alloc.F90:62-122). Host only."""
    g = L.Grid()
    check(L.lib().s2d_get_grid_size(nx, ny, dealias, C.byref(g)))
    return g


def Wavenumbers(g: L.Grid):
    kx = np.zeros(g.cISize, np.int32)
    ky = np.zeros(g.nyG, np.int32)
    check(L.lib().s2d_wavenumbers(C.byref(g), kx.ctypes.data_as(C.POINTER(C.c_int)),
                                  ky.ctypes.data_as(C.POINTER(C.c_int))))
    return kx, ky


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Spectral2D:
    """One simulation on the GPU (a libs2d context)."""

    def __init__(self, nx: int, ny: int | None = None, dealias: int = DEALIAS_3_2_RULE,
                 precision: int = FP64):
        ny = nx if ny is None else ny
        h = C.c_void_p()
        check(L.lib().s2d_create(nx, ny, dealias, precision, C.byref(h)))
        self._h = h
        self.precision = precision
        g = L.Grid()
        check(L.lib().s2d_get_grid(h, C.byref(g)), h)
        self.g = g
        self.nx, self.ny = nx, ny
        self.nxG, self.nyG = g.nxG, g.nyG
        self.rISize, self.cISize = g.rISize, g.cISize
        self.kx, self.ky = Wavenumbers(g)

    # ---- lifetime
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            L.lib().s2d_destroy(self._h)
            self._h = C.c_void_p()

    SpectralFinalize = close
    TimeIntegrationFinalize = close

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _c(self, rc):
        check(rc, self._h)

    # ---- buffers
    @property
    def ndoubles(self) -> int:
        return 2 * self.cISize * self.nyG

    def field(self) -> np.ndarray:
        return np.zeros(self.ndoubles, np.float64)

    def cview(self, buf: np.ndarray) -> np.ndarray:
        return buf.view(np.complex128).reshape(self.nyG, self.cISize)

    def rview(self, buf: np.ndarray) -> np.ndarray:
        return buf.reshape(self.nyG, self.rISize)

    def set_field(self, f, buf: np.ndarray):
        buf = np.ascontiguousarray(buf, dtype=np.float64).reshape(-1)
        if buf.size != self.ndoubles:
            raise ValueError(f"field buffer must hold {self.ndoubles} doubles, got {buf.size}")
        self._c(L.lib().s2d_set_field_host(self._h, _fid(f), _dp(buf)))

    def get_field(self, f, out: np.ndarray | None = None) -> np.ndarray:
        """Slot f in the reference layout (into `out` when given: a
        contiguous float64 array of 2*cISize*nyG values)."""
        if out is None:
            out = self.field()
        elif out.dtype != np.float64 or out.size != 2 * self.cISize * self.nyG \
                or not out.flags.c_contiguous:
            raise ValueError("out must be contiguous float64 of 2*cISize*nyG values")
        self._c(L.lib().s2d_get_field_host(self._h, _fid(f), _dp(out)))
        return out

    def zero_field(self, f):
        self._c(L.lib().s2d_zero_field(self._h, _fid(f)))

    def copy_field(self, src, dst):
        self._c(L.lib().s2d_copy_field(self._h, _fid(src), _fid(dst)))

    def device_ptr(self, f) -> int:
        """Device address of slot f (brought up to date with the state);
        FP64 or FP32 values per the context's precision, in the reference
        layout (rows of rISize reals / cISize complex)."""
        ptr = C.c_void_p()
        self._c(L.lib().s2d_field_device_ptr(self._h, _fid(f), C.byref(ptr)))
        return int(ptr.value)

    def sync_fields(self):
        self._c(L.lib().s2d_sync_fields(self._h))

    # ---- Spectral_m (src/Spectral.F90:83-88)
    def mask(self):
        a = [C.c_int() for _ in range(4)]
        self._c(L.lib().s2d_get_mask(self._h, *[C.byref(x) for x in a]))
        return tuple(x.value for x in a)

    def truncation(self) -> np.ndarray:
        i1 = np.zeros(self.nyG, np.int32)
        self._c(L.lib().s2d_get_truncation(self._h, i1.ctypes.data_as(C.POINTER(C.c_int))))
        return i1

    def TransformR2C(self, f):
        self._c(L.lib().s2d_transform_r2c(self._h, _fid(f)))

    def TransformC2R(self, f):
        self._c(L.lib().s2d_transform_c2r(self._h, _fid(f)))

    def DuDx(self, sig_in, sig_out):
        self._c(L.lib().s2d_dudx(self._h, _fid(sig_in), _fid(sig_out)))

    def DuDy(self, sig_in, sig_out):
        self._c(L.lib().s2d_dudy(self._h, _fid(sig_in), _fid(sig_out)))

    def Truncate(self, f):
        self._c(L.lib().s2d_truncate(self._h, _fid(f)))

    def ComputePsi(self, w="w", psi="psi"):
        self._c(L.lib().s2d_compute_psi(self._h, _fid(w), _fid(psi)))

    def ComputeVelocity(self, u="u", v="v", psi="psi", trans=False):
        self._c(L.lib().s2d_compute_velocity(self._h, _fid(u), _fid(v), _fid(psi), int(bool(trans))))

    def ComputeVorticity(self, u="u", v="v", w="w"):
        self._c(L.lib().s2d_compute_vorticity(self._h, _fid(u), _fid(v), _fid(w)))

    def ComputeRHS(self, nu, u="u", v="v", w="w", psi="psi", dW="dW"):
        self._c(L.lib().s2d_compute_rhs(self._h, float(nu), _fid(u), _fid(v), _fid(w),
                                        _fid(psi), _fid(dW)))

    # ---- TimeIntegration_m (src/TimeIntegration.F90:67-68)
    def IntegrateOneStep(self, nu, dt):
        """The reference step composed literally from the per-routine kernels."""
        self._c(L.lib().s2d_integrate_one_step(self._h, float(nu), float(dt)))

    def rk3_step(self, nu, dt):
        """The fused device-resident step."""
        self._c(L.lib().s2d_rk3_step(self._h, float(nu), float(dt)))

    def rk3_steps(self, nsteps: int, nu: float, dt: float = 0.0, cfl: float = 0.0):
        self._c(L.lib().s2d_rk3_steps(self._h, int(nsteps), float(nu), float(dt), float(cfl)))

    def ComputeTimeStep(self, cfl: float = 0.5) -> float:
        """CFL dt of the current state (the reference's stub, filled in)."""
        dt = C.c_double()
        self._c(L.lib().s2d_cfl_dt(self._h, float(cfl), C.byref(dt)))
        return dt.value

    def diagnostics(self) -> dict:
        d = L.Diag()
        self._c(L.lib().s2d_diagnostics(self._h, C.byref(d)))
        return {k: getattr(d, k) for k, _ in L.Diag._fields_}

    def ComputeSpectrum(self, nu: float = 0.0):
        """Shell spectrum of the current state (src/Analysis.F90:267-460):
        (ek, cnt, tke, eps, enstrophy_dissipation), FP64 on the device."""
        nb = L.lib().s2d_spectrum_bins(C.byref(self.g))
        ek = np.zeros(nb, np.float64)
        cnt = np.zeros(nb, np.int32)
        t, e, z = C.c_double(), C.c_double(), C.c_double()
        self._c(L.lib().s2d_spectrum(self._h, float(nu), _dp(ek),
                                     cnt.ctypes.data_as(C.POINTER(C.c_int)),
                                     C.byref(t), C.byref(e), C.byref(z)))
        return ek, cnt, t.value, e.value, z.value

    def ComputeSpectrumUV(self, u, v, nu: float = 0.0):
        """ComputeSpectrum from the velocity spectra in slots u and v
        (src/Analysis.F90:267-460, its uC, vC arguments): (ek, cnt, tke, eps,
        enstrophy_dissipation), FP64 on the device."""
        nb = L.lib().s2d_spectrum_bins(C.byref(self.g))
        ek = np.zeros(nb, np.float64)
        cnt = np.zeros(nb, np.int32)
        t, e, z = C.c_double(), C.c_double(), C.c_double()
        self._c(L.lib().s2d_spectrum_uv(self._h, _fid(u), _fid(v), float(nu), _dp(ek),
                                        cnt.ctypes.data_as(C.POINTER(C.c_int)),
                                        C.byref(t), C.byref(e), C.byref(z)))
        return ek, cnt, t.value, e.value, z.value

    # ---- schedule of the fused step and its autotuner (s2d_tune)
    def schedule_variants(self):
        """Variants compiled for passes A, B, C at this size."""
        n = (C.c_int * 3)()
        self._c(L.lib().s2d_schedule_variants(self._h, n))
        return tuple(n)

    def get_schedule(self) -> dict:
        s = L.Schedule()
        self._c(L.lib().s2d_get_schedule(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in L.Schedule._fields_}

    def set_schedule(self, pass_a=0, pass_b=0, pass_c=0, step_graph=True):
        s = L.Schedule(pass_a, pass_b, pass_c, int(bool(step_graph)))
        self._c(L.lib().s2d_set_schedule(self._h, C.byref(s)))

    def tune(self, nu: float, dt: float = 0.0, cfl: float = 0.5, cache: str | None = None,
             measure: bool = False, cache_only: bool = False) -> dict:
        """Pick the fastest schedule by measurement from the current state
        (FFTW_MEASURE's part; the state is restored), caching the choice in a
        JSON-lines file."""
        s = L.Schedule()
        flags = (L.TUNE_MEASURE if measure else 0) | (L.TUNE_CACHE_ONLY if cache_only else 0)
        self._c(L.lib().s2d_tune(self._h, float(nu), float(dt), float(cfl),
                                 cache.encode() if cache else None, flags, C.byref(s)))
        return {k: getattr(s, k) for k, _ in L.Schedule._fields_}

    def restrict_to(self, f, dst: "Spectral2D", df, kcut: float):
        """dst's slot df = the modes |k| <= kcut of this context's slot f
        (sharp spectral cutoff, sampled on dst's grid)."""
        check(L.lib().s2d_restrict(self._h, _fid(f), dst._h, _fid(df), float(kcut)), dst._h)

    def axpby(self, a: float, x, b: float, y):
        """slot y = a * slot x + b * slot y."""
        self._c(L.lib().s2d_axpby(self._h, float(a), _fid(x), float(b), _fid(y)))

    def set_forcing(self, f):
        """Subtract the kept modes of slot f from every stage's RHS."""
        self._c(L.lib().s2d_set_forcing(self._h, _fid(f)))

    def clear_forcing(self):
        self._c(L.lib().s2d_clear_forcing(self._h))

    def WriteRestart(self, path: str):
        """Checkpoint of the state, time and step count (s2d_checkpoint_write)."""
        self._c(L.lib().s2d_checkpoint_write(self._h, str(path).encode()))

    def ReadRestart(self, path: str):
        """Restore a checkpoint written for this grid (s2d_checkpoint_read)."""
        self._c(L.lib().s2d_checkpoint_read(self._h, str(path).encode()))

    def SetTurbulence(self, seed: int = 1, k0: float = 10.0, ke: float = 0.5):
        """Seeded decaying-turbulence IC into w and psi."""
        self._c(L.lib().s2d_set_turbulence(self._h, int(seed), float(k0), float(ke)))

    def SetTaylorGreen(self):
        """SetTaylorGreen, src/SetIC.F90:70-101, into w and psi."""
        self._c(L.lib().s2d_set_taylor_green(self._h))

    def profile_begin(self, max_launches: int, every: int = 1):
        """Bracket fused launches with HIP events (one step in every `every`)."""
        self._c(L.lib().s2d_profile_begin(self._h, int(max_launches)))
        if every != 1:
            self._c(L.lib().s2d_profile_every(self._h, int(every)))

    def profile_end(self):
        """Per fused kernel (pass A, B, C): (total ms, launches)."""
        ms = (C.c_double * 3)()
        cnt = (C.c_longlong * 3)()
        self._c(L.lib().s2d_profile_end(self._h, ms, cnt))
        return [(ms[k], cnt[k]) for k in range(3)]

    def get_time(self):
        """(time, steps taken, dt of the last step)."""
        t, n, d = C.c_double(), C.c_longlong(), C.c_double()
        self._c(L.lib().s2d_get_time(self._h, C.byref(t), C.byref(n), C.byref(d)))
        return t.value, n.value, d.value

    def set_time(self, time: float = 0.0, nadv: int = 0):
        self._c(L.lib().s2d_set_time(self._h, float(time), int(nadv)))

    def set_end_time(self, t_end: float = 0.0):
        """Fused steps land on t_end (> 0) and then keep the state; 0: no end."""
        self._c(L.lib().s2d_set_end_time(self._h, float(t_end)))


__all__ = ["GetGridSize", "Wavenumbers", "Spectral2D", "S2DError", "FIELDS", "FP64", "FP32",
           "DEALIAS_3_2_RULE", "FIELD_W", "FIELD_PSI", "FIELD_U", "FIELD_V", "FIELD_WS",
           "FIELD_PSIS", "FIELD_DW", "FIELD_TMP"]
