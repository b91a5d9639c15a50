"""The checkpoint format (s2d_checkpoint_write / read, include/s2d.h) on the
CPU: the Python writer and reader agree, the checksum catches a flipped
bit, and the checksum matches its definition on a tiny payload."""
import numpy as np
import pytest

from spectral2d_amd.checkpoint import checksum, read_checkpoint, write_checkpoint


def _hdr(**kw):
    h = dict(nx=16, ny=16, dealias=1, precision=0, nxG=24, nyG=24, cISize=13, rISize=26,
             time=0.375, nadv=17, dt=0.0125)
    h.update(kw)
    return h


def test_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    w = rng.standard_normal((24, 13)) + 1j * rng.standard_normal((24, 13))
    p = rng.standard_normal((24, 13)) + 1j * rng.standard_normal((24, 13))
    comp = (rng.standard_normal((9, 24)) + 1j * rng.standard_normal((9, 24))).astype(np.complex128)
    f = tmp_path / "a.ckpt"
    write_checkpoint(str(f), _hdr(), w, p, comp)
    d = read_checkpoint(str(f))
    assert d["time"] == 0.375 and d["nadv"] == 17 and d["dt"] == 0.0125
    assert np.array_equal(d["w"], w) and np.array_equal(d["psi"], p)
    assert np.array_equal(d["compact"], comp) and (d["K"], d["KP"]) == (9, 24)
    raw = bytearray(f.read_bytes())
    raw[200] ^= 1
    f.write_bytes(bytes(raw))
    with pytest.raises(ValueError):
        read_checkpoint(str(f))


def test_checksum_definition():
    words = np.array([1, 2, 3, 0xFFFFFFFFFFFFFFFF], dtype="<u8")
    s1 = (1 + 2 + 3 + 0xFFFFFFFFFFFFFFFF) % 2**64
    s2 = (1 * 1 + 2 * 2 + 3 * 3 + 4 * 0xFFFFFFFFFFFFFFFF) % 2**64
    rot = ((s2 << 32) | (s2 >> 32)) % 2**64
    assert checksum(words.tobytes()) == s1 ^ rot


def test_forcing_section_and_version_1(tmp_path):
    """Version 2 appends the forcing term after the compact state; a version-1
    file (no forcing, those header bytes zero) still reads."""
    rng = np.random.default_rng(1)
    w = rng.standard_normal((24, 13)) + 1j * rng.standard_normal((24, 13))
    comp = (rng.standard_normal((9, 24)) + 1j * rng.standard_normal((9, 24))).astype(np.complex64)
    forc = (rng.standard_normal((9, 24)) + 1j * rng.standard_normal((9, 24))).astype(np.complex64)
    f = tmp_path / "f.ckpt"
    write_checkpoint(str(f), _hdr(precision=1), w, w, comp, forc)
    d = read_checkpoint(str(f))
    assert d["version"] == 2
    assert np.array_equal(d["compact"], comp) and np.array_equal(d["forcing"], forc)
    raw = bytearray(f.read_bytes())
    assert raw[96:100] == b"\x01\x00\x00\x00"
    # the same state without forcing, relabelled version 1
    write_checkpoint(str(f), _hdr(precision=1), w, w, comp)
    raw = bytearray(f.read_bytes())
    raw[8:12] = (1).to_bytes(4, "little")
    f.write_bytes(bytes(raw))
    d = read_checkpoint(str(f))
    assert d["version"] == 1 and d["forcing"] is None and np.array_equal(d["compact"], comp)
    with pytest.raises(ValueError):
        write_checkpoint(str(f), _hdr(), w, w, None, forc)
