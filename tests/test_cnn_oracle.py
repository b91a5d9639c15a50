"""The CNN oracle (oracle/cnn.py) pinned on the CPU: its gradient equals
central finite differences of its own loss (FP64, no rounding), an
independent convolution (scipy.ndimage, wrap mode) gives the same forward
pass, and its BF16 rounding is round-to-nearest-even."""
import math

import numpy as np
import pytest

from oracle import cnn


def _data(B=2, H=8, W=8, seed=0):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((B, 2, H, W)), rng.standard_normal((B, H, W))


def test_param_count():
    assert cnn.param_count() == 149505
    p = cnn.init_params(1)
    assert p.size == cnn.param_count()
    assert np.array_equal(cnn.pack(cnn.unpack(p)), p)


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -9, 3.0e38, -2.5])
    r = cnn.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0 + 2 ** -7 or r[1] == 1.0  # exactly half an ulp: ties to even -> 1.0
    assert r[1] == 1.0
    assert r[2] == 1.0 + 2 ** -7            # above the half: rounds up
    assert r[3] == 1.0                       # a quarter ulp: down
    assert r[5] == -2.5


def test_forward_matches_independent_convolution():
    from scipy.ndimage import correlate
    x, _ = _data()
    p = cnn.init_params(3)
    y, c = cnn.forward(p, x, round=False)
    P = cnn.unpack(p)
    # layer 0 by scipy: z0[co] = sum_ci correlate(x_ci, W0[:, ci, co] as 3x3) + b0
    z0 = np.zeros(c["z0"].shape)
    for co in range(cnn.C):
        for ci in range(2):
            k = P["W0"][:, ci, co].reshape(3, 3)
            for b in range(x.shape[0]):
                z0[b, :, :, co] += correlate(x[b, ci], k, mode="wrap")
        z0[..., co] += P["b0"][co]
    assert np.allclose(z0, c["z0"], rtol=1e-12, atol=1e-12)
    # last layer by scipy
    a4 = c["a4"]
    y2 = np.full(y.shape, P["b5"][0])
    for ci in range(cnn.C):
        k = P["W5"][:, ci].reshape(3, 3)
        for b in range(x.shape[0]):
            y2[b] += correlate(a4[b, :, :, ci], k, mode="wrap")
    assert np.allclose(y2, y, rtol=1e-12, atol=1e-12)


def test_gradient_matches_finite_differences():
    x, t = _data(B=1, H=6, W=6, seed=4)
    p = cnn.init_params(5)
    loss, g, _ = cnn.loss_and_grad(p, x, t, in_scale=(0.7, 1.3), out_scale=2.0, round=False)
    rng = np.random.default_rng(6)
    # probe parameters of every layer
    sizes = [9 * 2 * 64, 64] + [9 * 64 * 64, 64] * 4 + [9 * 64, 1]
    starts = np.cumsum([0] + sizes[:-1])
    idx = []
    for s, n in zip(starts, sizes):
        idx += list(s + rng.choice(n, size=min(3, n), replace=False))
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        lp = cnn.loss_and_grad(p + e, x, t, (0.7, 1.3), 2.0, round=False)[0]
        lm = cnn.loss_and_grad(p - e, x, t, (0.7, 1.3), 2.0, round=False)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_gelu_closed_forms():
    z = np.array([0.0, 1.0, -1.0, 3.0])
    assert cnn.gelu(z)[0] == 0.0
    assert abs(cnn.gelu(z)[1] - 0.8413447460685429) < 1e-15
    assert abs(cnn.gelu_grad(np.array([0.0]))[0] - 0.5) < 1e-15
    # derivative by differences
    for v in (-2.0, -0.3, 0.4, 2.5):
        fd = (cnn.gelu(np.array([v + 1e-6])) - cnn.gelu(np.array([v - 1e-6])))[0] / 2e-6
        assert abs(fd - cnn.gelu_grad(np.array([v]))[0]) < 1e-8


def test_every_layer_matches_torch_circular_convolution():
    """All six layers (97 % of the flops are the four 64 -> 64 ones) against
    an independent route: torch's cross-correlation on the CPU in FP64 with
    circular padding, the weights laid out from the documented parameter
    order (layer 0 [tap, ci, co], layers 1-4 [tap, co, ci], layer 5
    [tap, ci], tap t = 3 dy + dx). A transposed weight or a wrong tap offset
    in any layer fails it."""
    torch = pytest.importorskip("torch")
    import torch.nn.functional as F

    x, _ = _data(B=2, H=8, W=12, seed=7)
    p = cnn.init_params(11)
    P = cnn.unpack(p)
    rng = np.random.default_rng(2)
    for k in P:  # nonzero biases, so that a dropped bias shows
        if k.startswith("b"):
            P[k] = rng.standard_normal(P[k].shape)
    p = cnn.pack(P)
    y, c = cnn.forward(p, x, in_scale=(0.8, 1.1), round=False)

    def conv(a_nhwc, w_oikk, b):
        a = torch.from_numpy(np.ascontiguousarray(a_nhwc.transpose(0, 3, 1, 2)))
        a = F.pad(a, (1, 1, 1, 1), mode="circular")
        z = F.conv2d(a, torch.from_numpy(np.ascontiguousarray(w_oikk)),
                     torch.from_numpy(np.asarray(b, np.float64)))
        return z.numpy().transpose(0, 2, 3, 1)

    def w_from(taps_co_ci):  # (9, co, ci) -> (co, ci, 3, 3)
        return taps_co_ci.reshape(3, 3, *taps_co_ci.shape[1:]).transpose(2, 3, 0, 1)

    z = conv(c["x0"], w_from(P["W0"].transpose(0, 2, 1)), P["b0"])
    assert np.allclose(z, c["z0"], rtol=1e-12, atol=1e-12)
    for l in range(1, cnn.LAYERS_MID + 1):
        z = conv(c[f"a{l - 1}"], w_from(P[f"W{l}"]), P[f"b{l}"])
        assert np.allclose(z, c[f"z{l}"], rtol=1e-12, atol=1e-12), l
    z = conv(c[f"a{cnn.LAYERS_MID}"], w_from(P["W5"][:, None, :]), P["b5"])
    assert np.allclose(z[..., 0], y, rtol=1e-12, atol=1e-12)
    # the input scaling of the two channels
    assert np.allclose(c["x0"][..., 0], 0.8 * x[:, 0]) and np.allclose(c["x0"][..., 1], 1.1 * x[:, 1])
