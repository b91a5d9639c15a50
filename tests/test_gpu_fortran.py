"""The Fortran drop-in (spectral2d_amd/fortran: Spectral_m, TimeIntegration_m
and the driver, over libs2d through ISO_C_BINDING) against the oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "spectral2d_amd", "fortran", "spectral2d")


def run_driver(tmp_path, **nml):
    cfg = tmp_path / "run.nml"
    out = tmp_path / "w.bin"
    items = []
    for k, v in nml.items():
        if isinstance(v, str):
            items.append(f"{k}='{v}'")
        elif isinstance(v, bool):
            items.append(f"{k}={'.TRUE.' if v else '.FALSE.'}")
        else:
            items.append(f"{k}={v}")
    items.append(f"out='{out}'")
    cfg.write_text("&spectral2d " + ", ".join(items) + " /\n")
    print(f"[fortran] {nml}", flush=True)
    r = subprocess.run([EXE, str(cfg)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    return np.fromfile(out, dtype=np.complex128), r.stdout


def rel_l2(a, b):
    return np.linalg.norm(a.ravel() - b.ravel()) / np.linalg.norm(b.ravel())


def test_fortran_taylor_green(tmp_path):
    N, nu, dt, n = 64, 0.01, 1e-3, 200
    w, out = run_driver(tmp_path, nx=N, ny=N, nu=nu, dt=dt, nsteps=n, ic="taylor-green",
                        printPeriod=50)
    o = oracle.Oracle(N)
    st = o.new_state()
    o.set_taylor_green(st["w"], st["psi"])
    w0 = o.cview(st["w"]).copy()
    for _ in range(n):
        o.step(st, nu, dt)
    got = w.reshape(o.nyG, o.cISize)
    assert rel_l2(got, o.cview(st["w"])) < 1e-12
    assert rel_l2(got, w0 * np.exp(-2 * nu * n * dt)) < 1e-12
    assert "Simulation step number: 00000200" in out


def test_fortran_end_time_fixed_dt(tmp_path):
    # DESIRED_END_TIME (src/Spectral2D.F90:426-430): after the step that
    # leaves time + dt >= tEnd, one last step of tEnd - time
    N, nu, dt, t_end = 64, 0.01, 3e-3, 0.01
    w, out = run_driver(tmp_path, nx=N, ny=N, nu=nu, dt=dt, tEnd=t_end, nsteps=1,
                        ic="taylor-green", printPeriod=100)
    o = oracle.Oracle(N)
    st = o.new_state()
    o.set_taylor_green(st["w"], st["psi"])
    w0 = o.cview(st["w"]).copy()
    t, steps, d = 0.0, 0, dt
    while True:
        o.step(st, nu, d)
        t += d
        steps += 1
        if t + d >= t_end:
            d = t_end - t
            o.step(st, nu, d)
            t += d
            steps += 1
            break
    assert steps == 4 and abs(t - t_end) < 1e-15
    got = w.reshape(o.nyG, o.cISize)
    assert rel_l2(got, o.cview(st["w"])) < 1e-12
    assert rel_l2(got, w0 * np.exp(-2 * nu * t_end)) < 1e-12
    assert "Simulation step number: 00000004" in out
    final = [l for l in out.splitlines() if l.startswith("Final TKE")][0]
    assert abs(float(final.split("time:")[1]) - t_end) < 1e-15


def test_fortran_end_time_cfl(tmp_path):
    # with a CFL dt the clamp follows ComputeTimeStep: min(CFL dt, tEnd - time)
    N, nu, t_end = 64, 1e-3, 0.05
    w, out = run_driver(tmp_path, nx=N, ny=N, nu=nu, cfl=0.5, tEnd=t_end, nsteps=1,
                        ic="turbulence", seed=5, k0=8.0, ke=0.5, printPeriod=1000)
    o = oracle.Oracle(N)
    st = o.new_state()
    o.set_turbulence(st["w"], st["psi"], seed=5, k0=8.0, ke=0.5)
    t, steps = 0.0, 0
    while True:
        d = o.cfl_dt(st["psi"], 0.5)
        last = t + d >= t_end
        if last:
            d = t_end - t
        o.step(st, nu, d)
        t += d
        steps += 1
        if last:
            break
    assert steps > 3
    assert rel_l2(w.reshape(o.nyG, o.cISize), o.cview(st["w"])) < 1e-12
    assert f"Simulation step number: {steps:08d}" in out


@pytest.mark.parametrize("resident", [True, False])
def test_fortran_turbulence_cfl(tmp_path, resident):
    N, nu, n = 64, 1e-3, 30
    w, _ = run_driver(tmp_path, nx=N, ny=N, nu=nu, cfl=0.5, nsteps=n, ic="turbulence",
                      seed=3, k0=8.0, ke=0.5, printPeriod=10, deviceResident=resident)
    o = oracle.Oracle(N)
    st = o.new_state()
    o.set_turbulence(st["w"], st["psi"], seed=3, k0=8.0, ke=0.5)
    for _ in range(n):
        o.step(st, nu, o.cfl_dt(st["psi"], 0.5))
    assert rel_l2(w.reshape(o.nyG, o.cISize), o.cview(st["w"])) < 1e-12


def test_library_ic_matches_oracle_bitwise():
    from spectral2d_amd.spectral import Spectral2D
    N = 96
    o = oracle.Oracle(N)
    w, psi = o.field(), o.field()
    o.set_turbulence(w, psi, seed=11, k0=7.0, ke=0.5)
    s = Spectral2D(N)
    s.SetTurbulence(seed=11, k0=7.0, ke=0.5)
    assert np.array_equal(s.get_field("w"), w)
    assert np.array_equal(s.get_field("psi"), psi)
    # Taylor-Green through the library's TransformR2C
    wt, pt = o.field(), o.field()
    o.set_taylor_green(wt, pt)
    s.SetTaylorGreen()
    assert np.max(np.abs(s.get_field("w") - wt)) < 1e-15
    s.close()
