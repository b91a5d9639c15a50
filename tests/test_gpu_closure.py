"""The closure coupling on the GPU: the one-call s2d_closure_forcing equals
the per-routine sequence (restrict, transforms, network, transform,
restrict, s2d_set_forcing) bitwise, and a forcing of zero network output
leaves the run unchanged."""
import numpy as np
import pytest

from oracle import cnn as ocnn
from spectral2d_amd.closure import ClosedRun
from spectral2d_amd.cnn import ClosureNet
from spectral2d_amd.spectral import FP32, Spectral2D

pytestmark = pytest.mark.gpu


def _net(N, zero=False):
    net = ClosureNet(N, N, 1)
    p = ocnn.init_params(2).astype(np.float32)
    if zero:
        p[:] = 0.0
    net.set_params(p)
    net.set_scales(0.2, 5.0, 0.05)
    return net


def test_fused_forcing_equals_per_routine_sequence():
    N = 128
    print(f"[closure] fused vs per-routine forcing, N={N}", flush=True)
    net = _net(N)
    a, b = ClosedRun(N, net), ClosedRun(N, net)
    for r in (a, b):
        r.sim.SetTurbulence(11, 8.0, 0.5)
    for _ in range(3):
        a.closure_forcing()
        b.closure_forcing_steps()
        a.sim.rk3_steps(1, 1e-4, 0.0, 0.5)
        b.sim.rk3_steps(1, 1e-4, 0.0, 0.5)
    assert np.array_equal(a.sim.get_field("w"), b.sim.get_field("w"))
    for r in (a, b):
        r.close()
    net.close()


def test_zero_closure_is_the_plain_run():
    N = 128
    print("[closure] zero network output", flush=True)
    net = _net(N, zero=True)
    a = ClosedRun(N, net)
    p = Spectral2D(N, precision=FP32)
    for s in (a.sim, p):
        s.SetTurbulence(12, 8.0, 0.5)
    for _ in range(3):
        a.closure_forcing()
        a.sim.rk3_steps(1, 1e-4, 0.0, 0.5)
        p.rk3_steps(1, 1e-4, 0.0, 0.5)
    # the forcing is exactly zero: x - 0 == x in every stage
    assert np.array_equal(a.sim.get_field("w"), p.get_field("w"))
    a.close()
    p.close()
    net.close()


def test_captured_closure_steps_equal_eager_steps():
    """s2d_closure_steps (hipGraph replay, forcing refreshed every 2 steps)
    equals the same sequence of eager calls bitwise; the refresh period runs
    on across calls and into fixed-dt steps (steps 0, 2, 4, ... refresh)."""
    N = 128
    print("[closure] captured closure steps vs eager", flush=True)
    net = _net(N)
    a, b = ClosedRun(N, net, every=2), ClosedRun(N, net, every=2)
    for r in (a, b):
        r.sim.SetTurbulence(13, 8.0, 0.5)
    k = 0

    def eager(n, cfl=0.5, dt=0.0):
        nonlocal k
        for _ in range(n):
            if k % 2 == 0:
                b.closure_forcing()
            b.sim.rk3_steps(1, 1e-4, dt, cfl)
            k += 1

    a.steps(5, 1e-4, 0.5)
    eager(5)
    assert np.array_equal(a.sim.get_field("w"), b.sim.get_field("w"))
    assert a.sim.get_time() == b.sim.get_time()
    a.steps(3, 1e-4, 0.5)  # replays the cached graphs
    eager(3)
    assert np.array_equal(a.sim.get_field("w"), b.sim.get_field("w"))
    h = 0.5 * a.sim.get_time()[2]
    a.steps(3, 1e-4, cfl=0.0, dt=h)  # fixed steps keep the period
    eager(3, cfl=0.0, dt=h)
    assert np.array_equal(a.sim.get_field("w"), b.sim.get_field("w"))
    assert a.sim.get_time() == b.sim.get_time()
    for r in (a, b):
        r.close()
    net.close()


def test_ensemble_closure_equals_runs_alone():
    """s2d_ensemble_closure_steps (one network call at B = 3 per refresh,
    refreshed every 4 steps, shell guard on, CFL steps with per-run end
    times) gives each member bitwise the trajectory of s2d_closure_steps for
    that run alone, across two calls (the period runs on)."""
    from spectral2d_amd.ensemble import Ensemble
    N, M, every, nu = 128, 3, 4, 1e-4
    t_end = [0.02, 0.0, 0.015]
    print(f"[closure] ensemble of {M} x {N}^2 vs runs alone, every={every}", flush=True)
    p = ocnn.init_params(4).astype(np.float32)
    netb = ClosureNet(N, N, M)
    net1 = ClosureNet(N, N, 1)
    for n in (netb, net1):
        n.set_params(p)
        n.set_scales(0.2, 5.0, 0.05)
    e = Ensemble([N] * M, FP32)
    e.set_viscosity(nu)
    e.set_end_time(t_end)
    from spectral2d_amd._lib import NO_DEALIASING
    grid = Spectral2D(N, dealias=NO_DEALIASING, precision=FP32)
    alone = []
    for r in range(M):
        a = ClosedRun(N, net1, every=every, guard=True)
        a.sim.SetTurbulence(20 + r, 8.0, 0.5)
        a.sim.set_time(0.0, 0)
        a.sim.set_end_time(t_end[r])
        e.set_vorticity(r, a.sim.get_field("w"))
        alone.append(a)
    for n in (7, 6):
        e.closure_steps(grid, netb, n, cfl=0.5, every=every, guard=True)
        for a in alone:
            a.steps(n, nu, cfl=0.5)
    d = e.diagnostics()
    for r, a in enumerate(alone):
        assert np.array_equal(e.get_field(r, "w"), a.sim.get_field("w")), r
        assert d[r]["time"] == a.sim.get_time()[0]
        a.close()
    e.close()
    grid.close()
    netb.close()
    net1.close()
