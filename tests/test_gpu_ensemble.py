"""GPU tests of the mixed-size ensemble (s2d_ensemble_*).

Contract: every member computes exactly what a context of its size computes
alone with the same initial condition, viscosity and step arguments (the
same kernels' arithmetic per run: bitwise), graph replay equals eager
launches (bitwise), and an FP64 member matches the CPU oracle to the
drop-in's 1e-12. Each case prints what it runs before touching the GPU.
"""
import numpy as np
import pytest

import oracle
from spectral2d_amd.ensemble import Ensemble, bytes_per_step
from spectral2d_amd.spectral import FP32, FP64, S2DError, Spectral2D

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (n if n > 0 else 1.0)


def _members(sizes, seed0=1000):
    rng = np.random.default_rng(seed0)
    k0 = rng.uniform(4.0, 16.0, len(sizes)) * np.asarray(sizes) / 512.0 + 2.0
    nu = 10.0 ** rng.uniform(-4.0, -3.0, len(sizes))
    seeds = np.arange(seed0, seed0 + len(sizes), dtype=np.uint64)
    return seeds, k0, nu


def _ensemble(sizes, precision, seeds, k0, nu):
    e = Ensemble(sizes, precision)
    e.set_viscosity(nu)
    e.set_turbulence(seeds, k0, 0.5)
    return e


@pytest.mark.parametrize("precision", [FP32, FP64])
def test_member_equals_run_alone(precision):
    sizes = [64, 128, 64, 256, 128, 512]
    steps = 12
    print(f"[ensemble] member == alone, precision={precision} sizes={sizes} steps={steps}",
          flush=True)
    seeds, k0, nu = _members(sizes)
    e = _ensemble(sizes, precision, seeds, k0, nu)
    assert e.nclasses == 4
    e.steps(steps, cfl=0.5)
    diags = e.diagnostics()
    for r, N in enumerate(sizes):
        s = Spectral2D(N, precision=precision)
        s.SetTurbulence(int(seeds[r]), float(k0[r]), 0.5)
        s.rk3_steps(steps, float(nu[r]), 0.0, 0.5)
        a, b = e.get_field(r, "w"), s.get_field("w")
        assert np.array_equal(a, b), (r, N, rel_l2(a, b))
        assert np.array_equal(e.get_field(r, "psi"), s.get_field("psi")), (r, N)
        d = s.diagnostics()
        assert diags[r]["nadv"] == steps
        assert diags[r]["time"] == d["time"], (r, diags[r]["time"], d["time"])
        assert diags[r]["tke"] == d["tke"]
        s.close()
    e.close()


@pytest.mark.parametrize("cfl,dt", [(0.5, 0.0), (0.0, 2e-3)])
def test_graph_replay_equals_eager(cfl, dt):
    sizes = [128, 256, 128, 512, 256, 128, 64]
    print(f"[ensemble] graph == eager, sizes={sizes} cfl={cfl} dt={dt}", flush=True)
    seeds, k0, nu = _members(sizes, 7)
    a = _ensemble(sizes, FP32, seeds, k0, nu)
    b = _ensemble(sizes, FP32, seeds, k0, nu)
    # several replays, with eager work and synchronisation between them
    for chunk in (1, 4, 3):
        a.steps(chunk, dt=dt, cfl=cfl)
        b.steps(chunk, dt=dt, cfl=cfl, graph=True)
        for r in range(len(sizes)):
            assert np.array_equal(a.get_field(r, "w"), b.get_field(r, "w")), (chunk, r)
    da, db = a.diagnostics(), b.diagnostics()
    for r in range(len(sizes)):
        assert da[r] == db[r], r
        assert da[r]["nadv"] == 8
    a.close()
    b.close()


def test_fp64_member_matches_oracle():
    """An FP64 member against the CPU oracle's IntegrateOneStep with the
    oracle's own CFL dt (the drop-in's 1e-12 bar)."""
    sizes = [128, 64, 128]
    steps = 15
    print(f"[ensemble] FP64 members vs oracle, sizes={sizes} steps={steps}", flush=True)
    seeds, k0, nu = _members(sizes, 11)
    e = _ensemble(sizes, FP64, seeds, k0, nu)
    e.steps(steps, cfl=0.5)
    diags = e.diagnostics()
    for r in (1, 2):
        o = oracle.Oracle(sizes[r])
        st = o.new_state()
        o.set_turbulence(st["w"], st["psi"], seed=int(seeds[r]), k0=float(k0[r]), ke=0.5)
        total = 0.0
        for _ in range(steps):
            dt = o.cfl_dt(st["psi"], 0.5)
            total += dt
            o.step(st, float(nu[r]), dt)
        err = rel_l2(e.cview(r, e.get_field(r, "w")), o.cview(st["w"]))
        print(f"  run {r} N={sizes[r]} rel_l2(w) = {err:.3e}")
        assert err < 1e-12
        assert abs(diags[r]["time"] - total) < 1e-13 * total
    e.close()


def test_fp32_members_error_vs_fp64_oracle():
    """The FP32 ensemble's error against the FP64 oracle, per size."""
    sizes = [64, 128, 256]
    steps = 20
    print(f"[ensemble] FP32 members vs FP64 oracle, sizes={sizes} steps={steps}", flush=True)
    seeds, k0, nu = _members(sizes, 21)
    e = _ensemble(sizes, FP32, seeds, k0, nu)
    dt = 1e-3
    e.steps(steps, dt=dt)
    for r, N in enumerate(sizes):
        o = oracle.Oracle(N)
        st = o.new_state()
        o.set_turbulence(st["w"], st["psi"], seed=int(seeds[r]), k0=float(k0[r]), ke=0.5)
        for _ in range(steps):
            o.step(st, float(nu[r]), dt)
        err = rel_l2(e.cview(r, e.get_field(r, "w")), o.cview(st["w"]))
        print(f"  run {r} N={N} FP32 rel_l2(w) vs FP64 oracle = {err:.3e}")
        assert err < 2e-5
    e.close()


def test_set_vorticity_round_trip_and_nonfinite_run_is_named():
    sizes = [64, 128, 64]
    print(f"[ensemble] set/get and NaN report, sizes={sizes}", flush=True)
    seeds, k0, nu = _members(sizes, 3)
    e = _ensemble(sizes, FP64, seeds, k0, nu)
    w0 = e.get_field(2, "w")
    e.set_vorticity(2, w0)
    assert np.array_equal(e.get_field(2, "w"), w0)
    bad = w0.copy()
    bad[2 * (3 + 5 * e.grids[2].cISize)] = np.nan  # a kept mode (kx=3, ky=5)
    e.set_vorticity(2, bad)
    with pytest.raises(S2DError) as ei:
        e.steps(2, cfl=0.5)
    assert "run 2" in str(ei.value)
    e.close()


def test_c2_shape_smoke():
    """A C2-shaped draw (sizes from {128, 256, 512}, k0 in [4, 16], nu
    log-uniform in [1e-4, 1e-3], CFL 0.5) at 24 runs, a few graph steps:
    every run advances and stays finite; energy decays."""
    rng = np.random.default_rng(2024)
    sizes = rng.choice([128, 256, 512], 24)
    print(f"[ensemble] C2-shaped 24 runs: {np.bincount(sizes)[[128, 256, 512]]}", flush=True)
    k0 = rng.uniform(4.0, 16.0, 24)
    nu = 10.0 ** rng.uniform(-4.0, -3.0, 24)
    seeds = np.arange(1000, 1024, dtype=np.uint64)
    e = _ensemble(sizes, FP32, seeds, k0, nu)
    d0 = e.diagnostics()
    e.steps(5, cfl=0.5, graph=True)
    d1 = e.diagnostics()
    for r in range(24):
        assert abs(d0[r]["tke"] - 0.5) < 1e-5
        assert d1[r]["nadv"] == 5 and d1[r]["time"] > 0
        assert d1[r]["tke"] < d0[r]["tke"]
    assert bytes_per_step([512]) == 60_002_304  # the two-pass model, 512^2 FP32
    e.close()


def test_c2_full_draw_members_equal_runs_alone():
    """The measured C2 workload itself (bench.ensemble_draw: 256 runs of
    N in {128, 256, 512}), 20 graph-replayed CFL steps: every run advances,
    stays finite and loses energy, and the first run of each size equals the
    same run computed alone, bitwise."""
    from bench import ensemble_draw
    sizes, k0, nu, seeds = ensemble_draw()
    steps = 20
    print(f"[ensemble] C2 draw, {len(sizes)} runs, counts "
          f"{[int((sizes == n).sum()) for n in (128, 256, 512)]}, {steps} graph steps", flush=True)
    e = _ensemble(sizes, FP32, seeds, k0, nu)
    d0 = e.diagnostics()
    e.steps(steps, cfl=0.5, graph=True)
    d1 = e.diagnostics()
    for r in range(len(sizes)):
        assert d1[r]["nadv"] == steps and d1[r]["time"] > 0, r
        assert np.isfinite(d1[r]["tke"]) and d1[r]["tke"] < d0[r]["tke"], r
    for N in (128, 256, 512):
        r = int(np.flatnonzero(sizes == N)[0])
        s = Spectral2D(int(N), precision=FP32)
        s.SetTurbulence(int(seeds[r]), float(k0[r]), 0.5)
        s.rk3_steps(steps, float(nu[r]), 0.0, 0.5)
        assert np.array_equal(e.get_field(r, "w"), s.get_field("w")), (r, N)
        assert d1[r]["time"] == s.diagnostics()["time"], (r, N)
        s.close()
    e.close()


@pytest.mark.parametrize("precision", [FP32, FP64])
def test_every_schedule_computes_the_same_bits(precision):
    """The unified launches, the unified launches with the fused pass A (the
    three fields of a column tile in one block), the stage queue kernel
    (per-run pass dependencies through a work queue) and the per-class
    launches give bitwise the same states, eager and by graph replay."""
    sizes = [128, 64, 512, 256, 128, 512, 64]
    seeds, k0, nu = _members(sizes, seed0=1200)
    print(f"[ensemble] schedules, precision={precision} sizes={sizes}", flush=True)
    out = {}
    for name, kw in (("unified", {}), ("fused_a", {"fused_a": True}),
                     ("stage_queue", {"stage_queue": True}), ("per_class", {"per_class": True})):
        e = _ensemble(sizes, precision, seeds, k0, nu)
        e.set_schedule(**kw)
        e.steps(3, cfl=0.5)
        e.steps(4, cfl=0.5, graph=True)
        out[name] = [e.get_field(r, "w") for r in range(len(sizes))]
        e.close()
    for name in ("fused_a", "stage_queue", "per_class"):
        for r in range(len(sizes)):
            assert np.array_equal(out[name][r], out["unified"][r]), (name, r)
