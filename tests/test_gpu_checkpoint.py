"""Checkpoint / restart on the GPU: a run restarted from a checkpoint
continues bitwise (fused contexts of one precision); the Python reader
parses what the library wrote; a mismatched grid is refused."""
import numpy as np
import pytest

from spectral2d_amd.checkpoint import read_checkpoint
from spectral2d_amd.spectral import FP32, FP64, S2DError, Spectral2D

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,prec", [(128, FP64), (256, FP32), (48, FP64)])
def test_restart_continues_bitwise(tmp_path, N, prec):
    print(f"[checkpoint] N={N} precision={prec}", flush=True)
    f = str(tmp_path / "r.ckpt")
    a = Spectral2D(N, precision=prec)
    a.SetTurbulence(3, max(2.0, N / 16), 0.5)
    a.rk3_steps(7, 1e-3, 0.0, 0.5)
    a.WriteRestart(f)
    a.rk3_steps(6, 1e-3, 0.0, 0.5)
    b = Spectral2D(N, precision=prec)
    b.ReadRestart(f)
    b.rk3_steps(6, 1e-3, 0.0, 0.5)
    assert np.array_equal(a.get_field("w"), b.get_field("w"))
    da, db = a.diagnostics(), b.diagnostics()
    assert da["time"] == db["time"] and da["nadv"] == db["nadv"] == 13
    d = read_checkpoint(f)
    assert (d["nx"], d["nxG"], d["precision"], d["nadv"]) == (N, 3 * N // 2, prec, 7)
    c = Spectral2D(N, precision=prec)
    c.ReadRestart(f)
    assert np.array_equal(c.cview(c.get_field("w")), d["w"])
    assert np.array_equal(c.cview(c.get_field("psi")), d["psi"])
    wrong = Spectral2D(N // 2, precision=prec)
    with pytest.raises(S2DError):
        wrong.ReadRestart(f)
    for s in (a, b, c, wrong):
        s.close()
