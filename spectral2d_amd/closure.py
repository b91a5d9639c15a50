"""Filtered DNS data and the learned subgrid closure (no reference
counterpart; the reference solves one resolution at a time).

Pieces, all computing through libs2d:

- :class:`FilteredSampler` -- from a fine run (N_f, 3/2 rule): the sharp
  spectral cutoff |k| <= N_c/2 of omega and psi and the subgrid flux
  tau = filter(u.grad omega) - u_bar.grad omega_bar, where u_bar.grad
  omega_bar is the coarse solver's own dealiased operator (an N_c context
  with the 3/2 rule), all sampled on the N_c x N_c grid (a context without
  padding does the transforms).
- :class:`ClosedRun` -- a coarse run whose every RK3 stage subtracts
  tau_hat = FFT(CNN(omega_bar, psi_bar)), evaluated once per step at the
  step's start (s2d_set_forcing).
- :func:`enstrophy_spectrum` -- Z(k) = k^2 E(k) with the reference's shells
  (src/Analysis.F90:358-400).
"""
from __future__ import annotations

import time

import numpy as np

from ._lib import NO_DEALIASING
from .cnn import ClosureNet
from .spectral import FP32, Spectral2D


def enstrophy_spectrum(sim, nu: float = 0.0):
    ek = sim.ComputeSpectrum(nu)[0]
    k = np.arange(ek.size, dtype=np.float64)
    return k * k * ek


class FilteredSampler:
    """Filtered snapshots of a fine context on an Nc x Nc grid."""

    def __init__(self, fine: Spectral2D, Nc: int):
        self.fine = fine
        self.Nc = Nc
        self.kc = Nc // 2
        self.coarse = Spectral2D(Nc, precision=fine.precision)
        self.grid = Spectral2D(Nc, dealias=NO_DEALIASING, precision=fine.precision)

    def close(self):
        self.coarse.close()
        self.grid.close()

    def filtered_state(self, dst: Spectral2D):
        """dst's W slot = the filtered vorticity (dst: an Nc 3/2-rule context)."""
        self.fine.restrict_to("w", dst, "w", self.kc)

    def snapshot(self):
        """(w_bar, psi_bar, tau) as float32 arrays (Nc, Nc), physical."""
        f, c, g, kc = self.fine, self.coarse, self.grid, self.kc
        f.copy_field("w", "tmp")
        f.ComputePsi("tmp", "psiS")
        f.ComputeRHS(0.0, "u", "v", "tmp", "psiS", "dW")   # -N_fine
        f.restrict_to("w", c, "w", kc)
        f.restrict_to("dW", c, "tmp", kc)                  # -filter(N_fine)
        c.ComputePsi("w", "psi")
        c.copy_field("w", "wS")
        c.ComputeRHS(0.0, "u", "v", "wS", "psi", "dW")     # -N_coarse(w_bar)
        c.axpby(-1.0, "tmp", 1.0, "dW")                    # tau_hat
        c.restrict_to("w", g, "w", kc)
        c.restrict_to("psi", g, "psi", kc)
        c.restrict_to("dW", g, "tmp", kc)
        out = []
        for s in ("w", "psi", "tmp"):
            g.TransformC2R(s)
            out.append(g.rview(g.get_field(s))[:, : self.Nc].astype(np.float32))
        return tuple(out)


class ClosedRun:
    """A coarse (Nc, 3/2 rule) run with the CNN closure term."""

    def __init__(self, Nc: int, net: ClosureNet | None, precision: int = FP32,
                 every: int = 1, guard: bool = False):
        self.Nc = Nc
        self.kc = Nc // 2
        self.net = net
        self.every = every  # closure evaluations: every `every` steps (frozen between)
        self.sim = Spectral2D(Nc, precision=precision)
        self.grid = Spectral2D(Nc, dealias=NO_DEALIASING, precision=precision)
        if guard:
            from . import _lib as L
            L.check(L.lib().s2d_set_closure_guard(self.sim._h, 1), self.sim._h)
        if net is not None and precision != FP32:
            raise ValueError("the closure network reads FP32 planes")

    def close(self):
        self.sim.close()
        self.grid.close()

    def closure_forcing(self):
        """s2d_closure_forcing: the whole sequence in one library call."""
        from . import _lib as L
        L.check(L.lib().s2d_closure_forcing(self.sim._h, self.grid._h, self.net._h), self.sim._h)

    def closure_forcing_steps(self):
        """The same sequence through the per-routine entry points."""
        s, g, kc = self.sim, self.grid, self.kc
        s.restrict_to("w", g, "w", kc)
        s.restrict_to("psi", g, "psi", kc)
        g.TransformC2R("w")
        g.TransformC2R("psi")
        ld = g.rISize
        self.net.forward(1, g.device_ptr("w"), g.device_ptr("psi"), 0, ld,
                         g.device_ptr("tmp"), 0, ld)
        g.TransformR2C("tmp")
        g.restrict_to("tmp", s, "tmp", kc)
        s.set_forcing("tmp")

    def steps(self, n: int, nu: float, cfl: float = 0.5, dt: float = 0.0):
        """n steps, CFL (cfl > 0) or of fixed dt; with a network the
        forcing is refreshed once `every` steps have been taken with it, a
        period that runs on across calls (s2d_closure_steps; CFL steps replay
        a captured hipGraph)."""
        from . import _lib as L
        if self.net is None:
            self.sim.rk3_steps(n, nu, dt, cfl)
        else:
            L.check(L.lib().s2d_closure_steps(self.sim._h, self.grid._h, self.net._h, int(n),
                                              float(nu), float(dt), float(cfl), int(self.every)),
                    self.sim._h)

    def advance_to(self, t_end: float, nu: float, cfl: float = 0.5, chunk: int = 100,
                   max_steps: int = 20000, z_limit: float = 4.0, log=None):
        """CFL steps in chunks (halved near the end) until the next CFL step
        would reach t_end, then fixed steps that land on it. Raises
        RuntimeError when the enstrophy exceeds z_limit times its initial
        value or max_steps is reached (an unstable closure)."""
        s = self.sim
        z0 = s.diagnostics()["enstrophy"]

        def cfl_steps(n):
            self.steps(n, nu, cfl)
            z = s.diagnostics()["enstrophy"]
            if not (z <= z_limit * z0):
                raise RuntimeError(f"enstrophy grew {z / z0:.2f}x by t = {s.get_time()[0]:.4f}")

        def fixed_step(h):
            self.steps(1, nu, cfl=0.0, dt=h)

        n = _advance(s, t_end, cfl, chunk, max_steps, cfl_steps, fixed_step)
        if log:
            log(f"  coarse run reached t = {s.get_time()[0]:.4f} in {n} steps")
        return n


def _advance(sim: Spectral2D, t_end: float, cfl: float, chunk: int, max_steps: int, cfl_steps,
             fixed_step) -> int:
    """Chunks of CFL steps while a chunk stays short of t_end (the chunk
    halved as t_end nears, a step taken as at most 1.25 times the last
    dt), then fixed steps of min(CFL dt, t_end - t) that land on t_end.
    A chunk whose steps grew faster than that and passed t_end raises
    RuntimeError (the comparison at t_end would be made at a later time)."""
    n = 0
    t, _, dt = sim.get_time()
    if dt <= 0.0:
        dt = sim.ComputeTimeStep(cfl)
    c = chunk
    tol = 1e-12 * max(1.0, abs(t_end))
    while t < t_end - tol:
        if n >= max_steps:
            raise RuntimeError(f"no arrival at t = {t_end} within {max_steps} steps (t = {t})")
        while c > 1 and t + c * 1.25 * dt >= t_end:
            c //= 2
        if t + c * 1.25 * dt < t_end:
            cfl_steps(c)
            n += c
        else:
            fixed_step(min(sim.ComputeTimeStep(cfl), t_end - t))
            n += 1
        t, _, dt = sim.get_time()
        if t > t_end + tol:
            raise RuntimeError(f"a chunk of CFL steps passed t_end = {t_end} (t = {t}): dt "
                               f"grew faster than 1.25x per step")
    if abs(t - t_end) > tol:
        raise RuntimeError(f"stopped at t = {t}, not t_end = {t_end}")
    return n


def advance_to(sim: Spectral2D, t_end: float, nu: float, cfl: float = 0.5, chunk: int = 100):
    """Fused CFL steps in chunks, then fixed steps that land on t_end."""
    return _advance(sim, t_end, cfl, chunk, 1 << 30,
                    lambda k: sim.rk3_steps(k, nu, 0.0, cfl),
                    lambda h: sim.rk3_steps(1, nu, h, 0.0))


def spectrum_error(z, z_ref, kmax: int):
    """Relative L2 distance of enstrophy spectra over shells 1..kmax."""
    a, b = np.asarray(z)[1:kmax + 1], np.asarray(z_ref)[1:kmax + 1]
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


class Timer:
    def __init__(self):
        self.t = 0.0

    def __enter__(self):
        self.t0 = time.perf_counter()
        return self

    def __exit__(self, *a):
        self.t += time.perf_counter() - self.t0


__all__ = ["FilteredSampler", "ClosedRun", "enstrophy_spectrum", "advance_to", "spectrum_error",
           "Timer"]
