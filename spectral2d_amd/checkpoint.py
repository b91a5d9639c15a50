"""Reader (and writer) of the libs2d checkpoint format (s2d_checkpoint_write,
include/s2d.h): a 128-byte little-endian header, then W and PSI in the
reference spectral layout as FP64 complex (cISize, nyG) arrays (x fastest),
then optionally the compact kept-modes state of a fused context and, in
version 2, the RHS forcing term in the same compact layout.

This module only parses files; it needs no GPU.
"""
from __future__ import annotations

import struct

import numpy as np

MAGIC = b"S2DCKPT1"
_HDR = struct.Struct("<8sII4i4idqdII2iQI28x")
assert _HDR.size == 128


def checksum(payload: bytes) -> int:
    """s1 = sum w_i, s2 = sum (i+1) w_i over the payload's 64-bit words,
    mod 2^64; s1 xor rotl64(s2, 32)."""
    w = np.frombuffer(payload, "<u8")
    with np.errstate(over="ignore"):
        s1 = int(np.sum(w, dtype=np.uint64))
        s2 = int(np.sum(w * np.arange(1, w.size + 1, dtype=np.uint64), dtype=np.uint64))
    m = 0xFFFFFFFFFFFFFFFF
    return s1 ^ (((s2 << 32) | (s2 >> 32)) & m)


def read_checkpoint(path: str, verify: bool = True) -> dict:
    """Parse a checkpoint. Returns the header fields and `w`, `psi` as
    complex128 arrays of shape (nyG, cISize) (row j = FFTW ky index), plus
    `compact` (complex64/128 array (K, KP)) and `forcing` (the same shape)
    when present."""
    with open(path, "rb") as f:
        raw = f.read()
    (magic, version, hbytes, nx, ny, dealias, precision, nxG, nyG, cISize, rISize, time,
     nadv, dt, nfields, has_compact, K, KP, ck, has_forcing) = _HDR.unpack_from(raw, 0)
    if magic != MAGIC or version not in (1, 2) or hbytes != 128:
        raise ValueError(f"{path}: not a version-1 or -2 libs2d checkpoint")
    if version == 1:
        has_forcing = 0
    nd = 2 * cISize * nyG
    off = 128
    w = np.frombuffer(raw, np.float64, nd, off)
    off += 8 * nd
    psi = np.frombuffer(raw, np.float64, nd, off)
    off += 8 * nd
    compact = forcing = None
    dt_c = np.complex128 if precision == 0 else np.complex64
    if has_compact:
        n = K * KP
        compact = np.frombuffer(raw, dt_c, n, off).reshape(K, KP)
        off += n * np.dtype(dt_c).itemsize
    if has_forcing:
        if not has_compact:
            raise ValueError(f"{path}: a forcing section without the compact state")
        forcing = np.frombuffer(raw, dt_c, K * KP, off).reshape(K, KP)
        off += K * KP * np.dtype(dt_c).itemsize
    if off != len(raw):
        raise ValueError(f"{path}: {len(raw) - off} unexpected trailing bytes")
    if verify and checksum(raw[128:]) != ck:
        raise ValueError(f"{path}: checksum mismatch")
    return {
        "nx": nx, "ny": ny, "dealias": dealias, "precision": precision, "nxG": nxG,
        "nyG": nyG, "cISize": cISize, "rISize": rISize, "time": time, "nadv": nadv, "dt": dt,
        "nfields": nfields, "K": K, "KP": KP, "version": version,
        "w": w.view(np.complex128).reshape(nyG, cISize),
        "psi": psi.view(np.complex128).reshape(nyG, cISize),
        "compact": compact, "forcing": forcing,
    }


def write_checkpoint(path: str, hdr: dict, w: np.ndarray, psi: np.ndarray,
                     compact: np.ndarray | None = None, forcing: np.ndarray | None = None):
    """Write the same format, version 2 (used to pin the layout in tests)."""
    payload = (np.ascontiguousarray(w, np.complex128).tobytes() +
               np.ascontiguousarray(psi, np.complex128).tobytes())
    if compact is not None:
        payload += np.ascontiguousarray(compact).tobytes()
    if forcing is not None:
        if compact is None or np.shape(forcing) != np.shape(compact):
            raise ValueError("the forcing needs the compact state's layout")
        payload += np.ascontiguousarray(forcing).tobytes()
    ck = checksum(payload)
    K, KP = (compact.shape if compact is not None else (0, 0))
    head = _HDR.pack(MAGIC, 2, 128, hdr["nx"], hdr["ny"], hdr["dealias"], hdr["precision"],
                     hdr["nxG"], hdr["nyG"], hdr["cISize"], hdr["rISize"], hdr["time"],
                     hdr["nadv"], hdr["dt"], 2, 1 if compact is not None else 0, K, KP, ck,
                     1 if forcing is not None else 0)
    with open(path, "wb") as f:
        f.write(head + payload)


__all__ = ["read_checkpoint", "write_checkpoint", "checksum", "MAGIC"]
