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
    # output files (REST_*, GRID.h5, EULERIAN.DAT, SPECTRUM_*) land in tmp_path
    r = subprocess.run([EXE, str(cfg)], capture_output=True, text=True, timeout=120,
                       cwd=str(tmp_path))
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


# ---- every Spectral_m routine through the Fortran module interface

EXE_R = os.path.join(ROOT, "spectral2d_amd", "fortran", "spectral2d_routines")


def _rand(o, rng):
    c = rng.standard_normal((o.nyG, o.cISize)) + 1j * rng.standard_normal((o.nyG, o.cISize))
    return o.from_complex(c)


@pytest.mark.parametrize("N", [32, 64])
def test_fortran_every_spectral_routine(tmp_path, N):
    """spectral2d_routines (spectral2d_amd/fortran/Routines.F90) calls DuDx,
    DuDy, ComputePsi, Truncate, ComputeVelocity (both forms),
    ComputeVorticity, TransformC2R, TransformR2C and ComputeRHS once each
    through Spectral_m; every record against the oracle: pointwise routines
    bitwise, transforms at the FP64 rounding level."""
    nu = 1e-3
    o = oracle.Oracle(N)
    rng = np.random.default_rng(N)
    w, psi, u, v = (_rand(o, rng) for _ in range(4))
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    np.concatenate([w, psi, u, v]).tofile(fin)
    print(f"[fortran routines] N={N}", flush=True)
    r = subprocess.run([EXE_R, str(N), repr(nu), str(fin), str(fout)], capture_output=True,
                       text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    got = np.fromfile(fout, dtype=np.float64).reshape(-1, w.size)
    assert got.shape[0] == 17
    ref = []
    a = o.field(); o.dudx(w, a); ref.append(a)
    a = o.field(); o.dudy(w, a); ref.append(a)
    a = o.field(); o.compute_psi(w, a); ref.append(a)
    a = w.copy(); o.truncate(a); ref.append(a)
    a, b = o.field(), o.field()
    o.compute_velocity(a, b, psi, trans=False)
    ref += [a.copy(), b.copy()]
    c = o.field()
    o.compute_vorticity(a, b, c)
    ref += [c, a, b]
    a, b = o.field(), o.field()
    o.compute_velocity(a, b, psi, trans=True)
    ref += [a, b]
    c = w.copy()
    o.transform_c2r(c)
    ref.append(c.copy())
    o.transform_r2c(c)
    ref.append(c)
    a, b, c, d = u.copy(), v.copy(), w.copy(), o.field()
    o.compute_rhs(nu, a, b, c, psi, d)
    ref += [d, a, b, c]
    for k in range(9):  # pointwise: bitwise
        assert np.array_equal(got[k], ref[k]), f"record {k + 1}"
    tol = 1e-14 * np.log2(o.nxG * o.nyG)
    for k in (9, 10, 11):  # physical fields: the nxG points of each row
        g, e = o.rview(got[k])[:, :o.nxG], o.rview(ref[k])[:, :o.nxG]
        assert rel_l2(g, e) < tol, f"record {k + 1}"
    assert rel_l2(o.cview(got[12]), o.cview(ref[12])) < tol
    assert rel_l2(o.cview(got[13]), o.cview(ref[13])) < 1e-13  # ComputeRHS dW
    for k in (14, 15):  # its u, v after: the products' truncated spectra
        assert rel_l2(o.cview(got[k]), o.cview(ref[k])) < 1e-13, f"record {k + 1}"
    assert rel_l2(o.cview(got[16]), o.cview(ref[16])) < tol  # w round-tripped


# ---- the driver's write / statistics cadence (WriteRestart, Analysis)

def _oracle_analysis(o, st, nu, nadv):
    """x/y skewness and the spectrum as src/Analysis.F90:90-242 forms them."""
    def sk(a, b):
        n = o.nxG * o.nyG
        ar, br = o.rview(a)[:, :o.nxG], o.rview(b)[:, :o.nxG]
        return -2.0 * (np.sum(ar * br ** 2) / n) / (np.sqrt(np.sum(ar ** 2) / n) *
                                                    (np.sum(br ** 2) / n))
    u, v = o.field(), o.field()
    o.compute_velocity(u, v, st["psi"])
    o.dudx(u, u)
    o.dudx(st["w"], v)
    o.transform_c2r(u)
    o.transform_c2r(v)
    xs = sk(u, v)
    u, v = o.field(), o.field()
    o.compute_velocity(u, v, st["psi"])
    o.dudy(v, v)
    o.dudy(st["w"], u)
    o.transform_c2r(v)
    o.transform_c2r(u)
    ys = sk(v, u)
    u, v = o.field(), o.field()
    o.compute_velocity(u, v, st["psi"])
    ek, _, tke, eps, ediss = o.spectrum(u, v, nu)
    return {"nadv": nadv, "tke": tke, "eps": eps, "ediss": ediss, "xs": xs, "ys": ys, "ek": ek}


def _oracle_restart(o, st):
    """WriteRestart's fields, and its round trip of w and psi."""
    u, v = o.field(), o.field()
    o.compute_velocity(u, v, st["psi"], trans=True)
    o.transform_c2r(st["w"])
    o.transform_c2r(st["psi"])
    f = {"Omega": o.rview(st["w"])[:, :o.nxG].copy(), "Psi": o.rview(st["psi"])[:, :o.nxG].copy(),
         "u": o.rview(u)[:, :o.nxG].copy(), "v": o.rview(v)[:, :o.nxG].copy()}
    o.transform_r2c(st["w"])
    o.transform_r2c(st["psi"])
    return f


@pytest.mark.parametrize("resident", [True, False])
def test_fortran_write_and_stat_cadence(tmp_path, resident):
    """The reference's loop (src/Spectral2D.F90:330-431): REST files at the
    start, every writePeriod steps and at the end, each a WriteRestart round
    trip of w and psi that the trajectory then follows; statistics at the
    start, every statPeriod steps and at the end. The oracle runs the same
    sequence; the final state, the REST fields and the EULERIAN / SPECTRUM
    files are compared with it."""
    from spectral2d_amd import restart_h5 as h5
    N, nu, n, wp, sp, cfl = 32, 1e-3, 23, 5, 7, 0.5
    w, out = run_driver(tmp_path, nx=N, ny=N, nu=nu, cfl=cfl, nsteps=n, ic="turbulence", seed=4,
                        k0=6.0, ke=0.5, printPeriod=10, writePeriod=wp, statPeriod=sp,
                        deviceResident=resident)
    o = oracle.Oracle(N)
    st = o.new_state()
    o.set_turbulence(st["w"], st["psi"], seed=4, k0=6.0, ke=0.5)
    rest, stats = [], []
    t = 0.0
    rest.append((0, t, _oracle_restart(o, st)))
    stats.append(_oracle_analysis(o, st, nu, 0))
    dt = o.cfl_dt(st["psi"], cfl)
    for k in range(1, n + 1):
        o.step(st, nu, dt)
        t += dt
        dt = o.cfl_dt(st["psi"], cfl)
        if k % wp == 0 or k == n:
            rest.append((k, t, _oracle_restart(o, st)))
        if k % sp == 0 or k == n:
            stats.append(_oracle_analysis(o, st, nu, k))
    assert rel_l2(w.reshape(o.nyG, o.cISize), o.cview(st["w"])) < 1e-12
    # REST files: numbered from 0 in write order, header and fields
    assert sorted(p.name for p in tmp_path.glob("REST_*.h5")) == \
        [h5.rest_name(i) for i in range(len(rest))]
    assert (tmp_path / "GRID.h5").exists()
    assert (tmp_path / h5.rest_name(len(rest) - 1, "xmf")).read_text() == \
        h5.xmf_text(o.nxG, o.nyG, len(rest) - 1)
    for i, (k, tk, f) in enumerate(rest):
        r = h5.read_rest_file(str(tmp_path / h5.rest_name(i)))
        assert (r["nx"], r["ny"], r["nadv"]) == (o.nxG, o.nyG, k)
        assert abs(r["time"] - tk) <= 1e-12 * max(tk, 1e-30) and r["nu"] == nu
        for name in ("Omega", "Psi", "u", "v"):
            assert rel_l2(r[name], f[name]) < 1e-12, (i, name)
    # EULERIAN.DAT: header, then one line per statistics call
    lines = (tmp_path / "EULERIAN.DAT").read_text().splitlines()
    assert lines[0].split() == ["#", "NADV", "TIME", "TKE", "EPS", "EnstDiss", "xSkew", "ySkew"]
    rows = [l.split() for l in lines[1:]]
    assert [int(r[0]) for r in rows] == [s["nadv"] for s in stats]
    for r, s in zip(rows, stats):
        for j, key in ((2, "tke"), (3, "eps"), (4, "ediss"), (5, "xs"), (6, "ys")):
            assert abs(float(r[j]) - s[key]) <= 1e-7 * abs(s[key]) + 1e-300, (r[0], key)
    # SPECTRUM_<nadv>.DAT of the last call: k, E, k^2 E, k^4 E
    last = stats[-1]
    sp_lines = (tmp_path / f"SPECTRUM_{n:08d}.DAT").read_text().splitlines()
    assert sp_lines[0].split() == ["#", "k", "E(k)", "D(k)", "DEnst(k)"]
    vals = np.array([[float(x) for x in l.split()] for l in sp_lines[1:]])
    assert len(vals) == len(last["ek"]) and np.array_equal(vals[:, 0], np.arange(len(vals)))
    # shells past the truncation hold only the round trips' rounding noise
    # (~1e-32 here): compared to 1e-12 of the peak there
    ek = last["ek"]
    assert np.allclose(vals[:, 1], ek, rtol=1e-7, atol=1e-12 * ek.max())
    assert np.allclose(vals[:, 2], vals[:, 0] ** 2 * ek, rtol=1e-7,
                       atol=1e-12 * (vals[:, 0] ** 2 * ek).max())
    assert f"Writing REST_{len(rest) - 1:06d}.h5" in out
