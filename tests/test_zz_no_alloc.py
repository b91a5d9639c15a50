"""s2d_diagnostics_noalloc (the entry point that allocates nothing) against
s2d_diagnostics through the C ABI: two calls in a row on one GPU, the second
on a smaller grid, each checked against the allocating entry point (FP64
sums to 1e-13 relative, max |w| and the non-finite count exactly)."""
import ctypes as C

import numpy as np
import pytest

from spectral2d_amd import _lib as L
from spectral2d_amd.spectral import FP32, FP64, Spectral2D

pytestmark = pytest.mark.gpu


def _device_sync():
    """The calls below enqueue on the null stream, which the contexts'
    non-blocking streams do not wait for: wait for the device."""
    hip = C.CDLL("libamdhip64.so")
    assert hip.hipDeviceSynchronize() == 0


def _noalloc(sim, out_sim):
    """Diagnostics of sim's W slot into the first 5 doubles of out_sim's
    TMP slot (an FP64 context), on the null stream, then read back."""
    sim.sync_fields()
    w = C.c_void_p(sim.device_ptr("w"))
    out = C.c_void_p(out_sim.device_ptr("tmp"))
    L.check(L.lib().s2d_diagnostics_noalloc(sim._h, None, w, out), sim._h)
    _device_sync()
    return out_sim.get_field("tmp")[:5]


@pytest.mark.parametrize("prec", [FP64, FP32])
def test_noalloc_diagnostics_match_allocating_entry(prec):
    out_sim = Spectral2D(32, precision=FP64)
    for N in (512, 96):  # the second call on a smaller input
        print(f"[noalloc] N={N} precision={prec}", flush=True)
        s = Spectral2D(N, precision=prec)
        s.SetTurbulence(7, max(2.0, N / 16), 0.5)
        s.rk3_steps(3, 1e-3, 0.0, 0.5)
        d = s.diagnostics()
        got = _noalloc(s, out_sim)
        for k, key in enumerate(("tke", "enstrophy", "palinstrophy")):
            assert abs(got[k] - d[key]) <= 1e-13 * abs(d[key]), key
        assert got[3] == d["max_abs_w"]
        assert got[4] == 0.0
        s.close()
    out_sim.close()


def test_noalloc_counts_nonfinite_and_refuses_null():
    s = Spectral2D(64)
    out_sim = Spectral2D(32)
    buf = s.field()
    buf[10] = np.nan
    buf[21] = np.inf
    s.set_field("tmp", buf)
    w = C.c_void_p(s.device_ptr("tmp"))
    out = C.c_void_p(out_sim.device_ptr("tmp"))
    L.check(L.lib().s2d_diagnostics_noalloc(s._h, None, w, out), s._h)
    _device_sync()
    assert out_sim.get_field("tmp")[4] == 2.0
    assert L.lib().s2d_diagnostics_noalloc(s._h, None, None, out) == 1
    s.close()
    out_sim.close()
