"""GPU parity of libs2d against the CPU FP64 oracle (reference layout).

Each case prints what it runs before touching the GPU. Tolerances:
pointwise kernels are bitwise (the same expression order, no contraction);
transforms and everything built on them are compared in relative L2 / max
norm at the FP64 rounding level (<= 1e-13 per transform, <= 1e-12 after
multi-step runs, the bar stated for the drop-in).
"""
import numpy as np
import pytest

import oracle
from spectral2d_amd.spectral import Spectral2D

pytestmark = pytest.mark.gpu

SIZES = [8, 16, 24, 32, 64, 96, 128]


def rel_l2(a, b):
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (n if n > 0 else 1.0)


def _pair(N, seed=0):
    print(f"[parity] N={N} seed={seed}", flush=True)
    o = oracle.Oracle(N)
    s = Spectral2D(N)
    return o, s


def _random_spectral(o, seed):
    rng = np.random.default_rng(seed)
    c = rng.standard_normal((o.nyG, o.cISize)) + 1j * rng.standard_normal((o.nyG, o.cISize))
    return o.from_complex(c)


def _turbulence(o, seed=1, k0=None):
    w, psi = o.field(), o.field()
    o.set_turbulence(w, psi, seed=seed, k0=k0 if k0 else max(2.0, o.nx / 8), ke=0.5)
    return w, psi


@pytest.mark.parametrize("N", SIZES)
def test_transforms(N):
    o, s = _pair(N)
    rng = np.random.default_rng(N)
    # c2r of a random spectral field (the Im of kx=0 / Nyquist are ignored by both)
    buf = _random_spectral(o, N)
    ref = buf.copy()
    o.transform_c2r(ref)
    s.set_field("tmp", buf)
    s.TransformC2R("tmp")
    got = s.get_field("tmp")
    r_ref = o.rview(ref)[:, : o.nxG]
    r_got = s.rview(got)[:, : o.nxG]
    assert rel_l2(r_got, r_ref) < 1e-14 * max(1, np.log2(o.nxG * o.nyG))
    # r2c of a random real field
    real = rng.standard_normal((o.nyG, o.nxG))
    buf = o.from_real(real)
    ref = buf.copy()
    o.transform_r2c(ref)
    s.set_field("tmp", buf)
    s.TransformR2C("tmp")
    got = s.get_field("tmp")
    assert rel_l2(s.cview(got), o.cview(ref)) < 1e-14 * max(1, np.log2(o.nxG * o.nyG))
    s.close()


@pytest.mark.parametrize("N", [8, 32, 96])
def test_pointwise_bitwise(N):
    o, s = _pair(N)
    a = _random_spectral(o, 7)
    b = _random_spectral(o, 8)
    # DuDx / DuDy
    for name, fo, fs in (("dudx", o.dudx, s.DuDx), ("dudy", o.dudy, s.DuDy)):
        ref = o.field()
        fo(a, ref)
        s.set_field("u", a)
        fs("u", "v")
        assert np.array_equal(s.get_field("v"), ref), name
    # Truncate
    ref = a.copy()
    o.truncate(ref)
    s.set_field("tmp", a)
    s.Truncate("tmp")
    assert np.array_equal(s.get_field("tmp"), ref)
    # ComputePsi
    ref = o.field()
    o.compute_psi(a, ref)
    s.set_field("w", a)
    s.ComputePsi("w", "psi")
    assert np.array_equal(s.get_field("psi"), ref)
    # ComputeVelocity (spectral)
    u, v = o.field(), o.field()
    o.compute_velocity(u, v, a, trans=False)
    s.set_field("psi", a)
    s.ComputeVelocity("u", "v", "psi", trans=False)
    assert np.array_equal(s.get_field("u"), u)
    assert np.array_equal(s.get_field("v"), v)
    # ComputeVorticity (destroys u, v)
    uo, vo, wo = a.copy(), b.copy(), o.field()
    o.compute_vorticity(uo, vo, wo)
    s.set_field("u", a)
    s.set_field("v", b)
    s.ComputeVorticity("u", "v", "w")
    assert np.array_equal(s.get_field("w"), wo)
    assert np.array_equal(s.get_field("u"), uo)
    assert np.array_equal(s.get_field("v"), vo)
    s.close()


@pytest.mark.parametrize("N", [16, 64])
def test_compute_rhs(N):
    o, s = _pair(N)
    w, psi = _turbulence(o)
    u, v, dW = o.field(), o.field(), o.field()
    wo = w.copy()
    o.compute_rhs(1e-3, u, v, wo, psi, dW)
    s.set_field("w", w)
    s.set_field("psi", psi)
    s.ComputeRHS(1e-3)
    assert rel_l2(s.cview(s.get_field("dW")), o.cview(dW)) < 1e-13
    # side effect: w round-tripped
    assert rel_l2(s.cview(s.get_field("w")), o.cview(wo)) < 1e-14
    s.close()


def _run_pair(N, steps, nu, fused, seed=1, cfl=0.5):
    o, s = _pair(N, seed)
    w, psi = _turbulence(o, seed)
    dt = o.cfl_dt(psi, cfl)
    st = o.new_state()
    st["w"][:] = w
    st["psi"][:] = psi
    s.set_field("w", w)
    s.set_field("psi", psi)
    for _ in range(steps):
        o.step(st, nu, dt)
        if not fused:
            s.IntegrateOneStep(nu, dt)
    if fused:
        s.rk3_steps(steps, nu, dt, 0.0)
    return o, s, st


@pytest.mark.parametrize("N,steps", [(16, 20), (64, 20), (96, 10)])
def test_integrate_one_step_literal(N, steps):
    o, s, st = _run_pair(N, steps, 1e-3, fused=False)
    e = rel_l2(s.cview(s.get_field("w")), o.cview(st["w"]))
    print(f"  literal N={N} steps={steps} rel_l2(w)={e:.3e}")
    assert e < 1e-12
    assert rel_l2(s.cview(s.get_field("psi")), o.cview(st["psi"])) < 1e-12
    s.close()


@pytest.mark.parametrize("N,steps", [(16, 20), (32, 50), (64, 50), (96, 20), (128, 20)])
def test_fused_rk3(N, steps):
    o, s, st = _run_pair(N, steps, 1e-3, fused=True)
    e = rel_l2(s.cview(s.get_field("w")), o.cview(st["w"]))
    print(f"  fused N={N} steps={steps} rel_l2(w)={e:.3e}")
    assert e < 1e-12
    assert rel_l2(s.cview(s.get_field("psi")), o.cview(st["psi"])) < 1e-12
    # truncated modes are exactly zero, as after the reference's Truncate
    keep = o.kept_mask()
    assert np.all(s.cview(s.get_field("w"))[~keep] == 0)
    s.close()


@pytest.mark.parametrize("N,steps", [(32, 40), (64, 30)])
def test_fused_cfl_mode(N, steps):
    """Device CFL dt (cfl * 2pi/N / max|u|+|v|, from the state at each step's
    start) against the oracle recomputing the same dt on the CPU."""
    print(f"[parity] CFL mode N={N} steps={steps}", flush=True)
    o = oracle.Oracle(N)
    w, psi = _turbulence(o, 4)
    st = o.new_state()
    st["w"][:] = w
    st["psi"][:] = psi
    total = 0.0
    for _ in range(steps):
        dt = o.cfl_dt(st["psi"], 0.5)
        total += dt
        o.step(st, 1e-3, dt)
    s = Spectral2D(N)
    s.set_field("w", w)
    s.set_field("psi", psi)
    s.set_time(0.0, 0)
    s.rk3_steps(steps, 1e-3, 0.0, 0.5)
    d = s.diagnostics()
    assert d["nadv"] == steps
    assert abs(d["time"] - total) < 1e-13 * total
    assert rel_l2(s.cview(s.get_field("w")), o.cview(st["w"])) < 1e-12
    # ComputeTimeStep of the final state matches the oracle's
    assert abs(s.ComputeTimeStep(0.5) - o.cfl_dt(st["psi"], 0.5)) < 1e-13 * total
    s.close()


@pytest.mark.parametrize("N,steps", [(64, 50), (96, 30), (256, 40), (512, 20), (1024, 10),
                                     (2048, 3)])
def test_fp32_error_vs_fp64_oracle(N, steps):
    """FP32 transforms, products and updates (FP64 twiddles, FP64 CFL and
    diagnostics): relative L2 error against the FP64 oracle, reported."""
    import os
    from spectral2d_amd.spectral import FP32
    print(f"[parity] FP32 N={N} steps={steps}", flush=True)
    oracle.set_threads(min(16, os.cpu_count() or 1))
    o = oracle.Oracle(N)
    w, psi = _turbulence(o, 6)
    dt = o.cfl_dt(psi, 0.5)
    st = o.new_state()
    st["w"][:] = w
    st["psi"][:] = psi
    for _ in range(steps):
        o.step(st, 1e-3, dt)
    s = Spectral2D(N, precision=FP32)
    s.set_field("w", w)
    s.set_field("psi", psi)
    s.rk3_steps(steps, 1e-3, dt)
    e = rel_l2(s.cview(s.get_field("w")), o.cview(st["w"]))
    print(f"  FP32 N={N} steps={steps} rel_l2(w) vs FP64 oracle = {e:.3e}")
    assert e < 2e-5
    # FP64 diagnostics of the FP32 state: TKE agrees to FP32 rounding
    u, v = o.field(), o.field()
    o.compute_velocity(u, v, st["psi"])
    tke_ref = o.spectrum(u, v)[2]
    assert abs(s.diagnostics()["tke"] - tke_ref) < 1e-5 * tke_ref
    s.close()


@pytest.mark.parametrize("N,steps", [(512, 10), (2048, 3)])
def test_fused_large_sizes(N, steps):
    """The compiled 768- and 3072-point kernels against the oracle."""
    import os
    oracle.set_threads(min(16, os.cpu_count() or 1))
    o, s, st = _run_pair(N, steps, 1e-4, fused=True)
    e = rel_l2(s.cview(s.get_field("w")), o.cview(st["w"]))
    print(f"  fused N={N} steps={steps} rel_l2(w)={e:.3e}")
    assert e < 1e-12
    s.close()


@pytest.mark.parametrize("N,precision", [(512, "fp64"), (1024, "fp64"), (2048, "fp64"),
                                         (1024, "fp32")])
def test_last_column_mode(N, precision):
    """A strong mode at kx = N/2, ky = 0, the one kept mode of the last
    column, on top of a turbulent field: the fused kernels write that column
    without transforming it (pass A: a constant along y; pass C from
    Np = 1536 up: the column sum of H_uw). Against the oracle, which
    transforms every column."""
    import os
    from spectral2d_amd.spectral import FP32, FP64
    print(f"[parity] last column mode N={N} {precision}", flush=True)
    oracle.set_threads(min(16, os.cpu_count() or 1))
    o = oracle.Oracle(N)
    w, psi = _turbulence(o, 4)
    o.cview(w)[0, N // 2] += 40.0 - 25.0j
    o.compute_psi(w, psi)
    dt = o.cfl_dt(psi, 0.5)
    st = o.new_state()
    st["w"][:] = w
    st["psi"][:] = psi
    for _ in range(3):
        o.step(st, 1e-4, dt)
    s = Spectral2D(N, precision=FP64 if precision == "fp64" else FP32)
    s.set_field("w", w)
    s.set_field("psi", psi)
    s.rk3_steps(3, 1e-4, dt, 0.0)
    got, ref = s.cview(s.get_field("w")), o.cview(st["w"])
    e = rel_l2(got, ref)
    em = abs(got[0, N // 2] - ref[0, N // 2]) / abs(ref[0, N // 2])
    print(f"  rel_l2(w)={e:.3e}, mode (N/2, 0) rel err={em:.3e}")
    tol = 1e-12 if precision == "fp64" else 2e-5
    assert e < tol and em < tol
    s.close()


def test_generic_and_compiled_kernels_agree():
    """S2D_GENERIC_KERNELS=1 selects the runtime-plan kernels for a size that
    has compiled kernels; both must give the same trajectory to rounding."""
    import os
    N = 256
    print(f"[parity] generic vs compiled kernels N={N}", flush=True)
    o = oracle.Oracle(N)
    w, psi = _turbulence(o, 9)
    out = []
    for flag in ("1", "0"):
        os.environ["S2D_GENERIC_KERNELS"] = flag
        try:
            s = Spectral2D(N)
        finally:
            os.environ.pop("S2D_GENERIC_KERNELS", None)
        s.set_field("w", w)
        s.set_field("psi", psi)
        s.rk3_steps(20, 1e-3, 0.0, 0.5)
        out.append(s.cview(s.get_field("w")).copy())
        s.close()
    assert rel_l2(out[1], out[0]) < 1e-13


def test_nonfinite_state_is_reported():
    from spectral2d_amd import S2DError
    N = 16
    print(f"[parity] NaN injection N={N}", flush=True)
    o = oracle.Oracle(N)
    w, psi = _turbulence(o, 2)
    s = Spectral2D(N)
    w[2 * (3 + 2 * o.cISize)] = np.nan
    s.set_field("w", w)
    with pytest.raises(S2DError) as e:
        s.rk3_steps(2, 1e-3, 1e-3)
    assert e.value.code == 5
    s.close()


def test_fused_equals_literal_path():
    N = 48
    print(f"[parity] fused vs literal N={N}", flush=True)
    o = oracle.Oracle(N)
    w, psi = _turbulence(o, 3)
    a, b = Spectral2D(N), Spectral2D(N)
    for x in (a, b):
        x.set_field("w", w)
        x.set_field("psi", psi)
    for _ in range(10):
        a.IntegrateOneStep(1e-3, 2e-3)
    b.rk3_steps(10, 1e-3, 2e-3)
    assert rel_l2(b.cview(b.get_field("w")), a.cview(a.get_field("w"))) < 1e-13
    a.close()
    b.close()


def test_taylor_green_decay():
    """C0: 256^2 (384^2 padded), nu=0.01, dt=1e-3, 1000 steps vs the
    analytic omega0 exp(-2 nu t) (src/SetIC.F90:47-56)."""
    N, nu, dt, steps = 256, 0.01, 1e-3, 1000
    print(f"[parity] Taylor-Green N={N} steps={steps}", flush=True)
    o = oracle.Oracle(N)
    w, psi = o.field(), o.field()
    o.set_taylor_green(w, psi)
    s = Spectral2D(N)
    s.set_field("w", w)
    s.set_field("psi", psi)
    s.rk3_steps(steps, nu, dt)
    got = s.cview(s.get_field("w"))
    exact = o.cview(w) * np.exp(-2 * nu * steps * dt)
    e = rel_l2(got, exact)
    print(f"  TG rel_l2={e:.3e}")
    assert e < 1e-12
    d = s.diagnostics()
    assert abs(d["time"] - steps * dt) < 1e-12 and d["nadv"] == steps
    s.close()


@pytest.mark.slow
def test_c1_parity_100_steps():
    """C1 FP64 drop-in: N=1024 (1536^2 padded), seeded E(k)~k^4 exp(-(k/10)^2),
    KE=0.5, nu=1e-4, CFL 0.5 dt; 100 steps; <= 1e-12 relative L2 in w."""
    import os
    N, nu, steps = 1024, 1e-4, 100
    oracle.set_threads(min(16, os.cpu_count() or 1))
    print(f"[parity] C1 N={N} steps={steps}", flush=True)
    o = oracle.Oracle(N)
    w, psi = o.field(), o.field()
    o.set_turbulence(w, psi, seed=1, k0=10.0, ke=0.5)
    st = o.new_state()
    st["w"][:] = w
    st["psi"][:] = psi
    dts = []
    for k in range(steps):
        dt = o.cfl_dt(st["psi"], 0.5)
        dts.append(dt)
        o.step(st, nu, dt)
        if k % 20 == 0:
            print(f"  oracle step {k}", flush=True)
    s = Spectral2D(N)
    s.set_field("w", w)
    s.set_field("psi", psi)
    s.rk3_steps(steps, nu, 0.0, 0.5)
    d = s.diagnostics()
    e = rel_l2(s.cview(s.get_field("w")), o.cview(st["w"]))
    print(f"  C1 rel_l2(w)={e:.3e} time gpu={d['time']:.15e} cpu={sum(dts):.15e}")
    assert abs(d["time"] - sum(dts)) < 1e-12 * sum(dts)
    assert e < 1e-12
    s.close()


@pytest.mark.parametrize("nx,ny", [(64, 96), (96, 64), (48, 128)])
def test_rectangular_grids(nx, ny):
    """nx != ny (GetGridSize takes them apart, src/Spectral.F90:117-219; kMax
    = nx/2 truncates both directions): transforms and IntegrateOneStep of
    the per-routine path against the oracle, and rk3_steps (which takes
    that path for rectangular grids) equal to it."""
    print(f"[parity] rectangular {nx}x{ny}", flush=True)
    o = oracle.Oracle(nx, ny)
    s = Spectral2D(nx, ny)
    assert (s.nxG, s.nyG) == (o.nxG, o.nyG)
    assert np.array_equal(s.truncation(), o.truncation())
    w, psi = o.field(), o.field()
    o.set_turbulence(w, psi, seed=6, k0=4.0, ke=0.5)
    s.set_field("w", w)
    s.set_field("psi", psi)
    st = o.new_state()
    st["w"][:] = w
    st["psi"][:] = psi
    dt = 1e-3
    for _ in range(5):
        o.step(st, 1e-3, dt)
        s.IntegrateOneStep(1e-3, dt)
    for f in ("w", "psi"):
        e = rel_l2(s.cview(s.get_field(f)), o.cview(st[f]))
        print(f"  {f}: rel_l2 = {e:.2e}")
        assert e < 1e-12
    s2 = Spectral2D(nx, ny)
    s2.set_field("w", w)
    s2.set_field("psi", psi)
    s2.rk3_steps(5, 1e-3, dt, 0.0)
    assert np.array_equal(s2.get_field("w"), s.get_field("w"))
    s.close()
    s2.close()


# ---- the per-routine kernels at the full sizes the reference's IO and
# Analysis call them on (src/IO.F90:118-142, src/Analysis.F90:142-231):
# Np = 1536 (1024^2) and 3072 (2048^2), where the column and row kernels
# size their batches from the LDS budget.

@pytest.mark.parametrize("N", [1024, 2048])
def test_transforms_full_size(N):
    oracle.set_threads(8)
    o, s = _pair(N)
    rng = np.random.default_rng(N)
    tol = 1e-14 * np.log2(o.nxG * o.nyG)
    buf = _random_spectral(o, N)
    ref = buf.copy()
    o.transform_c2r(ref)
    s.set_field("tmp", buf)
    s.TransformC2R("tmp")
    got = s.get_field("tmp")
    e = rel_l2(s.rview(got)[:, : o.nxG], o.rview(ref)[:, : o.nxG])
    print(f"  c2r N={N} rel_l2={e:.3e}", flush=True)
    assert e < tol
    real = rng.standard_normal((o.nyG, o.nxG))
    buf = o.from_real(real)
    ref = buf.copy()
    o.transform_r2c(ref)
    s.set_field("tmp", buf)
    s.TransformR2C("tmp")
    e = rel_l2(s.cview(s.get_field("tmp")), o.cview(ref))
    print(f"  r2c N={N} rel_l2={e:.3e}", flush=True)
    assert e < tol
    s.close()


def test_pointwise_and_velocity_full_size():
    """1024^2: the pointwise routines bitwise, ComputeVelocity with the
    inverse transform (the WriteRestart form) at the rounding level."""
    oracle.set_threads(8)
    N = 1024
    o, s = _pair(N)
    a = _random_spectral(o, 3)
    for name, fo, fs in (("dudx", o.dudx, s.DuDx), ("dudy", o.dudy, s.DuDy)):
        ref = o.field()
        fo(a, ref)
        s.set_field("u", a)
        fs("u", "v")
        assert np.array_equal(s.get_field("v"), ref), name
    ref = o.field()
    o.compute_psi(a, ref)
    s.set_field("tmp", a)
    s.ComputePsi("tmp", "psiS")
    assert np.array_equal(s.get_field("psiS"), ref)
    ref = a.copy()
    o.truncate(ref)
    s.Truncate("tmp")
    assert np.array_equal(s.get_field("tmp"), ref)
    w, psi = _turbulence(o, 2)
    u, v = o.field(), o.field()
    o.compute_velocity(u, v, psi, trans=True)
    s.set_field("psiS", psi)
    s.ComputeVelocity("u", "v", "psiS", trans=True)
    for slot, r in (("u", u), ("v", v)):
        e = rel_l2(s.rview(s.get_field(slot))[:, : o.nxG], o.rview(r)[:, : o.nxG])
        assert e < 1e-14 * np.log2(o.nxG * o.nyG), slot
    s.close()


def test_compute_rhs_full_size():
    oracle.set_threads(8)
    N = 1024
    o, s = _pair(N)
    w, psi = _turbulence(o, 1, k0=10.0)
    u, v, dW = o.field(), o.field(), o.field()
    wo = w.copy()
    o.compute_rhs(1e-4, u, v, wo, psi, dW)
    s.set_field("w", w)
    s.set_field("psi", psi)
    s.ComputeRHS(1e-4)
    e = rel_l2(s.cview(s.get_field("dW")), o.cview(dW))
    print(f"  ComputeRHS N={N} rel_l2(dW)={e:.3e}", flush=True)
    assert e < 1e-13
    assert rel_l2(s.cview(s.get_field("w")), o.cview(wo)) < 1e-14
    s.close()


def test_integrate_one_step_literal_512():
    """The literal IntegrateOneStep (per-routine kernels, 18 transforms per
    step) at 512^2 (Np = 768) over a few CFL steps."""
    oracle.set_threads(8)
    N, steps = 512, 4
    o, s, st = _run_pair(N, steps, 1e-4, fused=False)
    e = rel_l2(s.cview(s.get_field("w")), o.cview(st["w"]))
    print(f"  literal N={N} steps={steps} rel_l2(w)={e:.3e}", flush=True)
    assert e < 1e-12
    assert rel_l2(s.cview(s.get_field("psi")), o.cview(st["psi"])) < 1e-12
    s.close()
