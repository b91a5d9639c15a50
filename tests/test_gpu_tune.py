"""The fused step's schedules and the autotuner (s2d_tune,
s2d_ensemble_tune) on the GPU: every compiled variant of every pass agrees
with the default schedule to the arithmetic's rounding (and FP64 variants
with the oracle), graph replay equals eager launches bitwise, the tuner
restores the state it measured from and caches its choice as JSON lines."""
import json

import numpy as np
import pytest

import oracle
from spectral2d_amd.ensemble import Ensemble
from spectral2d_amd.spectral import FP32, FP64, Spectral2D

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def _run(N, prec, sched, steps=4, graph=True):
    s = Spectral2D(N, precision=prec)
    s.SetTurbulence(9, max(2.0, N / 32), 0.5)
    s.set_schedule(*sched, step_graph=graph)
    s.rk3_steps(steps, 1e-4, 0.0, 0.5)
    w = s.get_field("w")
    s.close()
    return w


@pytest.mark.parametrize("N,prec", [(512, FP64), (512, FP32), (1024, FP64), (1024, FP32),
                                    (2048, FP32)])
def test_every_variant_matches_default(N, prec):
    print(f"[tune] variants N={N} precision={prec}", flush=True)
    s = Spectral2D(N, precision=prec)
    n = s.schedule_variants()
    s.close()
    assert n[1] == 3 and n[0] == (4 if prec == FP32 else 3) and n[2] == (3 if prec == FP32 else 2)
    ref = _run(N, prec, (0, 0, 0))
    tol = 1e-12 if prec == FP64 else 1e-5
    for p in range(3):
        for v in range(1, n[p]):
            sched = [0, 0, 0]
            sched[p] = v
            print(f"  pass {'ABC'[p]} variant {v}", flush=True)
            e = rel_l2(_run(N, prec, tuple(sched)), ref)
            assert e < tol, (p, v, e)


def test_fp64_variants_match_oracle():
    """Non-default variants of every pass together at 512^2 FP64 (pass A
    with the three fields per block) against the oracle."""
    N, steps, nu = 512, 3, 1e-4
    oracle.set_threads(8)
    o = oracle.Oracle(N)
    st = o.new_state()
    o.set_turbulence(st["w"], st["psi"], seed=9, k0=N / 32, ke=0.5)
    for _ in range(steps):
        o.step(st, nu, o.cfl_dt(st["psi"], 0.5))
    got = _run(N, FP64, (2, 2, 1), steps=steps)
    assert rel_l2(got.view(np.complex128), o.cview(st["w"])) < 1e-12


@pytest.mark.parametrize("N,prec", [(64, FP64), (1024, FP64), (512, FP32)])
def test_graph_replay_equals_eager(N, prec):
    a = _run(N, prec, (0, 0, 0), steps=6, graph=True)
    b = _run(N, prec, (0, 0, 0), steps=6, graph=False)
    assert np.array_equal(a, b)


def test_tune_restores_state_and_caches(tmp_path):
    N = 1024
    cache = str(tmp_path / "tune.jsonl")
    s = Spectral2D(N, precision=FP64)
    s.SetTurbulence(3, 10.0, 0.5)
    s.rk3_steps(2, 1e-4, 0.0, 0.5)
    w0, t0 = s.get_field("w").copy(), s.get_time()
    r = s.tune(1e-4, cfl=0.5, cache=cache)
    print(f"[tune] N={N}: {r}", flush=True)
    # 3 x 3 x 2 variant combinations, each eager and by graph replay
    assert r["from_cache"] == 0 and r["candidates"] == 3 * 3 * 2 * 2
    assert r["step_us"] > 0 and r["step_us"] <= r["default_us"] * 1.02
    assert np.array_equal(s.get_field("w"), w0) and s.get_time() == t0
    assert s.get_schedule()["pass_a"] == r["pass_a"]
    lines = open(cache).read().splitlines()
    assert len(lines) == 1
    entry = json.loads(lines[0])
    assert entry["key"].startswith("single/fp64/1536x1536/1/")
    assert (entry["pass_a"], entry["pass_b"], entry["pass_c"]) == \
        (r["pass_a"], r["pass_b"], r["pass_c"])
    # a second context takes the cached choice without measuring
    s2 = Spectral2D(N, precision=FP64)
    r2 = s2.tune(1e-4, cfl=0.5, cache=cache)
    assert r2["from_cache"] == 1 and r2["pass_b"] == r["pass_b"]
    assert len(open(cache).read().splitlines()) == 1
    # the tuned context steps as the default one does, to rounding
    s2.SetTurbulence(3, 10.0, 0.5)
    s2.rk3_steps(2, 1e-4, 0.0, 0.5)
    s.rk3_steps(3, 1e-4, 0.0, 0.5)
    s2.rk3_steps(3, 1e-4, 0.0, 0.5)
    assert rel_l2(s.get_field("w"), s2.get_field("w")) < 1e-12
    for x in (s, s2):
        x.close()


def test_ensemble_tune(tmp_path):
    sizes = np.array([128, 512, 256, 512, 128, 256, 512])
    cache = str(tmp_path / "ens.jsonl")
    e = Ensemble(sizes, FP32)
    e.set_viscosity(np.full(len(sizes), 3e-4))
    e.set_turbulence(np.arange(len(sizes)) + 40, np.full(len(sizes), 6.0), 0.5)
    e.steps(2, cfl=0.5, graph=True)
    before = [e.get_field(r, "w").copy() for r in range(len(sizes))]
    sched, us = e.tune(cfl=0.5, cache=cache)
    print(f"[tune] ensemble schedule {sched}, {us:.1f} us/step", flush=True)
    assert sched in (0, 1, 2, 3, 4) and us > 0
    for r in range(len(sizes)):
        assert np.array_equal(e.get_field(r, "w"), before[r])
    # the chosen schedule steps as the unified one does
    f = Ensemble(sizes, FP32)
    f.set_viscosity(np.full(len(sizes), 3e-4))
    f.set_turbulence(np.arange(len(sizes)) + 40, np.full(len(sizes), 6.0), 0.5)
    f.steps(2, cfl=0.5, graph=True)
    e.steps(4, cfl=0.5, graph=True)
    f.steps(4, cfl=0.5, graph=True)
    for r in range(len(sizes)):
        a, b = e.get_field(r, "w"), f.get_field(r, "w")
        assert (np.array_equal(a, b) if sched != 2 else rel_l2(a, b) < 1e-5), r
    entry = json.loads(open(cache).read().splitlines()[-1])
    assert entry["schedule"] == sched and entry["key"].startswith("ensemble/fp32/192x2,384x2,768x3/")
    g = Ensemble(sizes, FP32)
    assert g.tune(cfl=0.5, cache=cache, cache_only=True)[0] == sched
    for x in (e, f, g):
        x.close()
