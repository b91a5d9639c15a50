"""End times of the fused steps (s2d_set_end_time, s2d_ensemble_set_end_time):
a run lands on its end time with one clamped step (bitwise the CFL steps
and then a fixed step of the remaining time), then keeps its state exactly;
ensemble members land on their own end times, each bitwise the run alone."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NU = 1e-3


def _run(N, precision, seed):
    from spectral2d_amd.spectral import Spectral2D
    s = Spectral2D(N, precision=precision)
    s.SetTurbulence(seed=seed, k0=6.0, ke=0.5)
    s.set_time(0.0, 0)
    return s


@pytest.mark.parametrize("precision", [0, 1])
def test_run_lands_on_its_end_time_and_stays(precision):
    from spectral2d_amd.spectral import FP32, FP64
    prec = FP64 if precision == 0 else FP32
    print(f"[end time] single run, precision={prec}", flush=True)
    a = _run(64, prec, 5)
    t_end = 0.0537
    a.set_end_time(t_end)
    a.rk3_steps(400, NU, 0.0, 0.5)
    ta, na, _ = a.get_time()
    assert abs(ta - t_end) <= 1e-12 * t_end and 0 < na < 400
    w_end = a.get_field("w")
    a.rk3_steps(5, NU, 0.0, 0.5)          # at the end: no change at all
    assert np.array_equal(a.get_field("w"), w_end)
    assert a.get_time()[:2] == (ta, na)
    # the same as na - 1 CFL steps and one fixed step of the remaining time
    b = _run(64, prec, 5)
    b.rk3_steps(na - 1, NU, 0.0, 0.5)
    tb = b.get_time()[0]
    b.rk3_steps(1, NU, t_end - tb, 0.0)
    assert np.array_equal(b.get_field("w"), w_end)
    a.set_end_time(0.0)                   # the limit removed: the run goes on
    a.rk3_steps(2, NU, 0.0, 0.5)
    assert a.get_time()[0] > t_end and not np.array_equal(a.get_field("w"), w_end)
    a.close()
    b.close()


def test_ensemble_members_land_on_their_own_end_times():
    from spectral2d_amd.ensemble import Ensemble
    from spectral2d_amd.spectral import FP32
    sizes = [64, 128, 64, 128]
    t_end = [0.031, 0.047, 0.0, 0.022]    # run 2: no end time
    print(f"[end time] ensemble sizes={sizes} t_end={t_end}", flush=True)
    e = Ensemble(sizes, FP32)
    e.set_viscosity(NU)
    seeds = np.arange(40, 44, dtype=np.uint64)
    e.set_turbulence(seeds, 6.0, 0.5)
    e.set_end_time(t_end)
    e.steps(60, cfl=0.5, graph=True)
    d = e.diagnostics()
    for r in range(len(sizes)):
        s = _run(sizes[r], FP32, 0)
        s.SetTurbulence(seed=int(seeds[r]), k0=6.0, ke=0.5)
        s.set_time(0.0, 0)
        s.set_end_time(t_end[r])
        s.rk3_steps(60, NU, 0.0, 0.5)
        t, n, _ = s.get_time()
        if t_end[r] > 0:
            assert abs(t - t_end[r]) <= 1e-12 * t_end[r] and n < 60
        else:
            assert n == 60
        assert d[r]["time"] == t
        assert np.array_equal(e.get_field(r, "w"), s.get_field("w")), r
        s.close()
    e.close()
