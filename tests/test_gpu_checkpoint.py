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


@pytest.mark.parametrize("prec", [FP64, FP32])
def test_forced_run_restarts_bitwise(tmp_path, prec):
    """The forcing term (s2d_set_forcing, the closure's entry point) is part
    of the saved state: a forced run restarted from its checkpoint continues
    bitwise; restoring it into a context without forcing turns it on, and a
    checkpoint without forcing turns a context's forcing off."""
    N = 64
    print(f"[checkpoint] forced N={N} precision={prec}", flush=True)
    f, g = str(tmp_path / "forced.ckpt"), str(tmp_path / "plain.ckpt")
    a = Spectral2D(N, precision=prec)
    a.SetTurbulence(5, 6.0, 0.5)
    a.copy_field("w", "tmp")
    a.axpby(0.0, "tmp", 0.05, "tmp")
    a.set_forcing("tmp")
    a.rk3_steps(4, 1e-3, 0.0, 0.5)
    a.WriteRestart(f)
    a.rk3_steps(5, 1e-3, 0.0, 0.5)
    b = Spectral2D(N, precision=prec)
    b.ReadRestart(f)
    b.rk3_steps(5, 1e-3, 0.0, 0.5)
    assert np.array_equal(a.get_field("w"), b.get_field("w"))
    d = read_checkpoint(f)
    assert d["version"] == 2 and d["forcing"] is not None and np.any(d["forcing"] != 0)
    # a plain checkpoint read into the forced context b clears its forcing
    c = Spectral2D(N, precision=prec)
    c.SetTurbulence(5, 6.0, 0.5)
    c.WriteRestart(g)
    b.ReadRestart(g)
    c.rk3_steps(3, 1e-3, 0.0, 0.5)
    b.rk3_steps(3, 1e-3, 0.0, 0.5)
    assert np.array_equal(b.get_field("w"), c.get_field("w"))
    # the forced checkpoint cannot go into a context of the other precision
    other = Spectral2D(N, precision=FP32 if prec == FP64 else FP64)
    with pytest.raises(S2DError):
        other.ReadRestart(f)
    for s in (a, b, c, other):
        s.close()


def test_corrupt_checkpoints_fail_with_status(tmp_path):
    """Damaged headers and sizes are refused with a status code before any
    allocation is sized from them (no abort across the C ABI)."""
    N = 32
    print(f"[checkpoint] corrupt headers N={N}", flush=True)
    f = tmp_path / "ok.ckpt"
    a = Spectral2D(N)
    a.SetTurbulence(2, 4.0, 0.5)
    a.WriteRestart(str(f))
    good = f.read_bytes()

    def patched(off, value: bytes, name):
        p = tmp_path / name
        raw = bytearray(good)
        raw[off:off + len(value)] = value
        p.write_bytes(bytes(raw))
        return str(p)

    bad = [patched(80, (2**31 - 1).to_bytes(4, "little"), "k.ckpt"),  # K huge
           patched(84, (-8).to_bytes(4, "little", signed=True), "kp.ckpt"),  # KP negative
           patched(28, (7).to_bytes(4, "little"), "prec.ckpt"),  # precision
           patched(8, (3).to_bytes(4, "little"), "ver.ckpt"),  # version
           patched(96, (1).to_bytes(4, "little"), "forc.ckpt"),  # forcing, no payload
           patched(200, bytes([good[200] ^ 1]), "sum.ckpt")]  # a payload bit
    trunc = tmp_path / "short.ckpt"
    trunc.write_bytes(good[:-16])
    bad.append(str(trunc))
    for p in bad:
        with pytest.raises(S2DError):
            a.ReadRestart(p)
    a.ReadRestart(str(f))  # the context still works
    a.rk3_steps(2, 1e-3, 0.0, 0.5)
    a.close()
