"""An ensemble of independent decaying-turbulence runs of mixed sizes on one
GPU (the s2d_ensemble_* entry points of include/s2d.h).

The reference runs one simulation per process group (src/Spectral2D.F90);
there is no ensemble counterpart. Each member computes what a
:class:`~spectral2d_amd.spectral.Spectral2D` of its size computes alone with
the same initial condition, viscosity and step arguments; runs of one size
are batched, and every kernel launch advances all runs of all sizes.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import FIELD_PSI, FIELD_W, FP32, check_ens

FIELDS = {"w": FIELD_W, "psi": FIELD_PSI}


class Ensemble:
    """``len(n)`` runs, run r on an ``n[r] x n[r]`` grid (3/2 rule)."""

    def __init__(self, n, precision: int = FP32):
        n = np.ascontiguousarray(np.asarray(n, dtype=np.int32))
        h = C.c_void_p()
        check_ens(L.lib().s2d_ensemble_create(int(n.size), n.ctypes.data_as(C.POINTER(C.c_int)),
                                              precision, C.byref(h)))
        self._h = h
        self.n = n.copy()
        self.nruns = int(n.size)
        self.precision = precision
        nr, nc = C.c_int(), C.c_int()
        self._c(L.lib().s2d_ensemble_size(h, C.byref(nr), C.byref(nc)))
        self.nclasses = nc.value
        self.grids = []
        for r in range(self.nruns):
            g = L.Grid()
            self._c(L.lib().s2d_ensemble_run_grid(h, r, C.byref(g)))
            self.grids.append(g)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            L.lib().s2d_ensemble_destroy(self._h)
            self._h = C.c_void_p()

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
        check_ens(rc, self._h)

    @staticmethod
    def _d(a):
        return a.ctypes.data_as(C.POINTER(C.c_double))

    def set_viscosity(self, nu):
        nu = np.ascontiguousarray(np.broadcast_to(np.asarray(nu, np.float64), (self.nruns,)))
        self._c(L.lib().s2d_ensemble_set_viscosity(self._h, self._d(nu)))

    def set_end_time(self, t_end=None):
        """Per-run end times (> 0; each run lands on its own and then keeps its
        state); None clears them."""
        if t_end is None:
            self._c(L.lib().s2d_ensemble_set_end_time(self._h, None))
            return
        t = np.ascontiguousarray(np.broadcast_to(np.asarray(t_end, np.float64), (self.nruns,)))
        self._c(L.lib().s2d_ensemble_set_end_time(self._h, self._d(t)))

    def closure_steps(self, grid, net, nsteps: int, dt: float = 0.0, cfl: float = 0.5,
                      every: int = 8, guard: bool = True):
        """s2d_ensemble_closure_steps: the closure forcing of every run from one
        network call (grid: an FP32 unpadded Spectral2D of the runs' size;
        net: a ClosureNet of that grid with batch >= runs)."""
        self._c(L.lib().s2d_ensemble_closure_steps(self._h, grid._h, net._h, int(nsteps),
                                                   float(dt), float(cfl), int(every),
                                                   int(bool(guard))))

    def get_time(self):
        """(time[nruns], steps[nruns]) of every run."""
        t = np.zeros(self.nruns, np.float64)
        n = np.zeros(self.nruns, np.int64)
        self._c(L.lib().s2d_ensemble_get_time(self._h, self._d(t),
                                              n.ctypes.data_as(C.POINTER(C.c_longlong))))
        return t, n

    def clear_forcing(self):
        self._c(L.lib().s2d_ensemble_clear_forcing(self._h))

    def set_turbulence(self, seed, k0, ke=0.5):
        """Spectrum/phase recipe of ``Spectral2D.SetTurbulence`` for every run."""
        seed = np.ascontiguousarray(np.broadcast_to(np.asarray(seed, np.uint64), (self.nruns,)))
        k0 = np.ascontiguousarray(np.broadcast_to(np.asarray(k0, np.float64), (self.nruns,)))
        ke = np.ascontiguousarray(np.broadcast_to(np.asarray(ke, np.float64), (self.nruns,)))
        self._c(L.lib().s2d_ensemble_set_turbulence(
            self._h, seed.ctypes.data_as(C.POINTER(C.c_ulonglong)), self._d(k0), self._d(ke)))

    def ndoubles(self, run: int) -> int:
        g = self.grids[run]
        return 2 * g.cISize * g.nyG

    def set_vorticity(self, run: int, w: np.ndarray):
        w = np.ascontiguousarray(w, np.float64)
        if w.size != self.ndoubles(run):
            raise ValueError(f"run {run}: expected {self.ndoubles(run)} doubles, got {w.size}")
        self._c(L.lib().s2d_ensemble_set_vorticity(self._h, int(run), self._d(w)))

    def get_field(self, run: int, f="w") -> np.ndarray:
        """Reference-layout host copy of run ``run``'s vorticity or streamfunction."""
        buf = np.zeros(self.ndoubles(run), np.float64)
        fid = FIELDS[f] if isinstance(f, str) else int(f)
        self._c(L.lib().s2d_ensemble_get_field(self._h, int(run), fid, self._d(buf)))
        return buf

    def cview(self, run: int, buf: np.ndarray) -> np.ndarray:
        g = self.grids[run]
        return buf.view(np.complex128).reshape(g.nyG, g.cISize)

    def steps(self, nsteps: int, dt: float = 0.0, cfl: float = 0.0, graph: bool = False):
        self._c(L.lib().s2d_ensemble_steps(self._h, int(nsteps), float(dt), float(cfl),
                                           L.ENS_GRAPH if graph else 0))

    def set_schedule(self, per_class: bool = False, fused_a: bool = False,
                     stage_queue: bool = False):
        """One launch per pass for all sizes (the default; fused_a: pass A
        with the three fields of a column tile per block), per size class
        (per_class), or one kernel per RK stage over a work queue (stage_queue).
        Every schedule computes the same bits."""
        code = (L.ENS_PER_CLASS if per_class else L.ENS_STAGE_QUEUE if stage_queue
                else L.ENS_UNIFIED_FUSED_A if fused_a else L.ENS_UNIFIED)
        self._c(L.lib().s2d_ensemble_set_schedule(self._h, code))

    def tune(self, dt: float = 0.0, cfl: float = 0.5, cache: str | None = None,
             measure: bool = False, cache_only: bool = False):
        """s2d_ensemble_tune: the launch schedule (and the classes' kernel
        variants) by measurement from the current state, which is restored.
        Returns (schedule, microseconds per step)."""
        sched = C.c_int()
        us = C.c_double()
        flags = (L.TUNE_MEASURE if measure else 0) | (L.TUNE_CACHE_ONLY if cache_only else 0)
        self._c(L.lib().s2d_ensemble_tune(self._h, float(dt), float(cfl),
                                          cache.encode() if cache else None, flags,
                                          C.byref(sched), C.byref(us)))
        return sched.value, us.value

    def spectrum(self, run: int, nu: float = 0.0):
        """Shell spectrum of one run: (ek, cnt, tke, eps, enstrophy_dissipation)."""
        nb = L.lib().s2d_spectrum_bins(C.byref(self.grids[run]))
        ek = np.zeros(nb, np.float64)
        cnt = np.zeros(nb, np.int32)
        t, e, z = C.c_double(), C.c_double(), C.c_double()
        self._c(L.lib().s2d_ensemble_spectrum(self._h, int(run), float(nu), self._d(ek),
                                              cnt.ctypes.data_as(C.POINTER(C.c_int)),
                                              C.byref(t), C.byref(e), C.byref(z)))
        return ek, cnt, t.value, e.value, z.value

    def diagnostics(self) -> list:
        d = (L.Diag * self.nruns)()
        self._c(L.lib().s2d_ensemble_diagnostics(self._h, d))
        return [{k: getattr(x, k) for k, _ in L.Diag._fields_} for x in d]

    def profile_begin(self, max_launches: int):
        self._c(L.lib().s2d_ensemble_profile_begin(self._h, int(max_launches)))

    def profile_end(self):
        """Per ensemble pass (A, B, C): (total ms, launches)."""
        ms = (C.c_double * 3)()
        cnt = (C.c_longlong * 3)()
        self._c(L.lib().s2d_ensemble_profile_end(self._h, ms, cnt))
        return [(ms[k], cnt[k]) for k in range(3)]


def bytes_per_step(n, precision: int = FP32) -> int:
    """Algorithmic HBM bytes of one RK3 step of runs of sizes ``n`` (the
    two-pass fused model: per stage 2 s K (10 Np + 4 N) per run)."""
    s = 8 if precision == 0 else 4
    tot = 0
    for N in np.asarray(n).ravel():
        N = int(N)
        K, Np = N // 2 + 1, 3 * N // 2
        tot += 3 * 2 * s * K * (10 * Np + 4 * N)
    return tot


__all__ = ["Ensemble", "bytes_per_step"]
