"""CPU FP64 oracle of the closure CNN (test infrastructure only).

The network of the learned subgrid closure: six periodic 3x3 convolutions,
64 channels, GELU after all but the last; inputs (omega, psi) on the coarse
grid, output tau. It has no counterpart in the reference; this oracle states
the arithmetic the GPU kernels (spectral2d_amd/csrc/cnn.hip) perform, so
that their outputs and gradients can be checked:

- parameters are an FP32 master vector; the kernels use their BF16
  roundings (`bf16`: round to nearest even) in every product;
- the input channels are scaled, then rounded to BF16; every layer's
  pre-activation z is accumulated in FP32 (here FP64), stored rounded to
  BF16 (used by the backward pass), and its activation gelu(z) (of the
  unrounded z) is stored rounded to BF16;
- the output is FP32 (not rounded); the loss is the mean of
  (y - target / out_scale)^2;
- the backward pass rounds each dZ to BF16 where the kernels store it.

With round=False every rounding is the identity, which is the function the
finite-difference checks differentiate.

Layout: activations (B, H, W, C) (NHWC), tap t = 3 dy + dx with offset
(dy - 1, dx - 1), periodic wrap. Parameter vector, in order:
W0 (9, 2, 64) [tap, ci, co], b0 (64); for l = 1..4: Wl (9, 64, 64)
[tap, co, ci], bl (64); W5 (9, 64) [tap, ci], b5 (1).
"""
from __future__ import annotations

import math

import numpy as np

C = 64
LAYERS_MID = 4
SQRT_HALF = math.sqrt(0.5)


def param_count() -> int:
    return 9 * 2 * C + C + LAYERS_MID * (9 * C * C + C) + 9 * C + 1


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to the nearest BF16 (ties to even), returned as float64."""
    x32 = np.asarray(x, np.float32)
    u = x32.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    nan = np.isnan(x32)
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    out[nan] = np.nan
    return out


def _erf(x):
    from scipy.special import erf
    return erf(x)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z * SQRT_HALF))


def gelu_grad(z):
    return 0.5 * (1.0 + _erf(z * SQRT_HALF)) + z * np.exp(-0.5 * z * z) / math.sqrt(2 * math.pi)


def unpack(p: np.ndarray) -> dict:
    p = np.asarray(p, np.float64)
    o = 0
    out = {}

    def take(name, shape):
        nonlocal o
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n

    take("W0", (9, 2, C))
    take("b0", (C,))
    for l in range(1, LAYERS_MID + 1):
        take(f"W{l}", (9, C, C))
        take(f"b{l}", (C,))
    take("W5", (9, C))
    take("b5", (1,))
    assert o == p.size
    return out


def pack(d: dict) -> np.ndarray:
    parts = [d["W0"], d["b0"]]
    for l in range(1, LAYERS_MID + 1):
        parts += [d[f"W{l}"], d[f"b{l}"]]
    parts += [d["W5"], d["b5"]]
    return np.concatenate([np.asarray(x, np.float64).ravel() for x in parts])


def init_params(seed: int = 0) -> np.ndarray:
    """He-style initialisation (FP32 values)."""
    rng = np.random.default_rng(seed)
    d = {"W0": rng.standard_normal((9, 2, C)) * math.sqrt(2.0 / 18), "b0": np.zeros(C)}
    for l in range(1, LAYERS_MID + 1):
        d[f"W{l}"] = rng.standard_normal((9, C, C)) * math.sqrt(2.0 / (9 * C))
        d[f"b{l}"] = np.zeros(C)
    d["W5"] = rng.standard_normal((9, C)) * math.sqrt(1.0 / (9 * C))
    d["b5"] = np.zeros(1)
    return pack(d).astype(np.float32).astype(np.float64)


def _shift(a, dy, dx):
    """a[(y + dy) % H, (x + dx) % W] for a of shape (B, H, W, ...)."""
    return np.roll(np.roll(a, -dy, axis=1), -dx, axis=2)


def _taps():
    for t in range(9):
        yield t, t // 3 - 1, t % 3 - 1


def forward(params, x, in_scale=(1.0, 1.0), round=True):
    """x: (B, 2, H, W). Returns (y (B, H, W), cache)."""
    r = bf16 if round else (lambda v: np.asarray(v, np.float64))
    P = unpack(params)
    Wb = {k: (r(v) if k.startswith("W") else v) for k, v in P.items()}
    x = np.asarray(x, np.float64)
    x0 = np.stack([r(in_scale[0] * x[:, 0]), r(in_scale[1] * x[:, 1])], axis=-1)  # B,H,W,2
    cache = {"x0": x0, "Wb": Wb}
    # layer 0
    z = np.zeros(x0.shape[:3] + (C,))
    for t, dy, dx in _taps():
        z += _shift(x0, dy, dx) @ Wb["W0"][t]
    z += P["b0"]
    cache["z0"], cache["a0"] = r(z), r(gelu(z))
    a = cache["a0"]
    for l in range(1, LAYERS_MID + 1):
        z = np.zeros(a.shape)
        for t, dy, dx in _taps():
            z += _shift(a, dy, dx) @ Wb[f"W{l}"][t].T
        z += P[f"b{l}"]
        cache[f"z{l}"], cache[f"a{l}"] = r(z), r(gelu(z))
        a = cache[f"a{l}"]
    y = np.zeros(a.shape[:3])
    for t, dy, dx in _taps():
        y += _shift(a, dy, dx) @ Wb["W5"][t]
    y += P["b5"][0]
    return y, cache


def loss_and_grad(params, x, target, in_scale=(1.0, 1.0), out_scale=1.0, round=True):
    """Mean-square loss of y against target / out_scale and its gradient
    with respect to the parameter vector."""
    r = bf16 if round else (lambda v: np.asarray(v, np.float64))
    y, c = forward(params, x, in_scale, round)
    t = np.asarray(target, np.float64) / out_scale
    n = y.size
    res = y - t
    loss = float(np.mean(res * res))
    dy_ = 2.0 * res / n
    Wb = c["Wb"]
    g = {}
    a4 = c[f"a{LAYERS_MID}"]
    g["W5"] = np.stack([np.einsum("bhw,bhwc->c", dy_, _shift(a4, ddy, ddx))
                        for _, ddy, ddx in _taps()])
    g["b5"] = np.array([dy_.sum()])
    dA = np.zeros(a4.shape)
    for tt, ddy, ddx in _taps():
        dA += _shift(dy_, -ddy, -ddx)[..., None] * Wb["W5"][tt]
    dZ = r(dA * gelu_grad(c[f"z{LAYERS_MID}"]))
    for l in range(LAYERS_MID, 0, -1):
        ain = c[f"a{l - 1}"]
        g[f"W{l}"] = np.stack([np.einsum("bhwo,bhwi->oi", dZ, _shift(ain, ddy, ddx))
                               for _, ddy, ddx in _taps()])
        g[f"b{l}"] = dZ.sum(axis=(0, 1, 2))
        dA = np.zeros(ain.shape)
        for tt, ddy, ddx in _taps():
            dA += _shift(dZ, -ddy, -ddx) @ Wb[f"W{l}"][tt]
        dZ = r(dA * gelu_grad(c[f"z{l - 1}"]))
    x0 = c["x0"]
    g["W0"] = np.stack([np.einsum("bhwi,bhwo->io", _shift(x0, ddy, ddx), dZ)
                        for _, ddy, ddx in _taps()])
    g["b0"] = dZ.sum(axis=(0, 1, 2))
    return loss, pack(g), y


__all__ = ["param_count", "bf16", "gelu", "gelu_grad", "unpack", "pack", "init_params",
           "forward", "loss_and_grad"]
