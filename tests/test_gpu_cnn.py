"""The closure CNN's GPU kernels against the CPU FP64 oracle (oracle/cnn.py)
with the same BF16 rounding points. Accumulation order (FP32 matrix cores
vs FP64) may flip an occasional BF16 rounding, so the comparisons are in
relative L2 with stated bounds: forward <= 2e-3, each layer's weight
gradient <= 2e-2. Training on a fixed batch lowers the loss."""
import numpy as np
import pytest

from oracle import cnn as ocnn

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a, b = np.ravel(a), np.ravel(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _setup(B=2, H=8, W=128, ld=None, seed=0):
    from spectral2d_amd.cnn import ClosureNet
    rng = np.random.default_rng(seed)
    ld = ld or W
    x = rng.standard_normal((B, 2, H, W)).astype(np.float32)
    t = (0.3 * rng.standard_normal((B, H, W))).astype(np.float32)
    # device planes with row pitch ld (the padded real layout of a transform)
    dev = torch.device("cuda:0")
    xw = torch.zeros((B, H, ld), dtype=torch.float32, device=dev)
    xp = torch.zeros((B, H, ld), dtype=torch.float32, device=dev)
    tg = torch.zeros((B, H, ld), dtype=torch.float32, device=dev)
    xw[:, :, :W] = torch.from_numpy(x[:, 0])
    xp[:, :, :W] = torch.from_numpy(x[:, 1])
    tg[:, :, :W] = torch.from_numpy(t)
    net = ClosureNet(H, W, B)
    p = ocnn.init_params(seed + 1).astype(np.float32)
    net.set_params(p)
    net.set_scales(0.8, 1.25, 0.5)
    return net, p, x, t, xw, xp, tg, H * ld, ld


# W = 128 and 384 take the 128-pixel conv tiles, W = 256 the 256-pixel ones
@pytest.mark.parametrize("W,ld", [(128, 128), (128, 130), (256, 256), (384, 384)])
def test_forward_matches_oracle(W, ld):
    print(f"[cnn] forward vs oracle, W={W} ld={ld}", flush=True)
    net, p, x, t, xw, xp, tg, stride, ld = _setup(W=W, ld=ld)
    out = torch.zeros_like(tg)
    net.forward(2, xw.data_ptr(), xp.data_ptr(), stride, ld, out.data_ptr(), stride, ld)
    torch.cuda.synchronize()
    y_ref, _ = ocnn.forward(p.astype(np.float64), x, in_scale=(0.8, 1.25))
    got = out.cpu().numpy()[:, :, :W] / 0.5
    e = rel_l2(got, y_ref)
    print(f"  forward rel_l2 = {e:.2e}")
    assert e < 2e-3
    net.close()


@pytest.mark.parametrize("W", [128, 256])
def test_gradient_matches_oracle(W):
    print(f"[cnn] gradient vs oracle, W={W}", flush=True)
    net, p, x, t, xw, xp, tg, stride, ld = _setup(W=W, seed=3)
    loss, g = net.gradient(2, xw.data_ptr(), xp.data_ptr(), stride, ld, tg.data_ptr(), stride, ld)
    loss_ref, g_ref, _ = ocnn.loss_and_grad(p.astype(np.float64), x, t, in_scale=(0.8, 1.25),
                                            out_scale=0.5)
    assert abs(loss - loss_ref) <= 2e-3 * loss_ref
    G, R = ocnn.unpack(g.astype(np.float64)), ocnn.unpack(g_ref)
    for k in R:
        e = rel_l2(G[k], R[k])
        print(f"  {k}: rel_l2 = {e:.2e}")
        assert e < 2e-2, k
    net.close()


def test_training_lowers_the_loss():
    print("[cnn] Adam steps on one batch", flush=True)
    net, p, x, t, xw, xp, tg, stride, ld = _setup(B=4, seed=5)
    losses = [net.train_step(4, xw.data_ptr(), xp.data_ptr(), stride, ld, tg.data_ptr(), stride,
                             ld, 1e-3) for _ in range(60)]
    print(f"  loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    assert losses[-1] < 0.5 * losses[0]
    q = net.get_params()
    assert q.shape == p.shape and not np.array_equal(q, p)
    net.close()


@pytest.mark.parametrize("H,W", [(64, 128), (32, 256)])
def test_narrow_launches_take_the_multi_row_paths(H, W):
    """Launch width 1: the 64->64 kernels walk all H rows per block (the
    conv ring refilled H/8 - 1 times, the weight-gradient ring over every
    row), the thin-layer gradients take every row chunk and the grid-stride
    kernel loops over every pixel. The forward is bitwise the default
    launches' (same accumulation order per output), the gradient equal to
    rounding (other partial-sum grouping), both against the oracle."""
    print(f"[cnn] launch width 1 vs default, H={H} W={W}", flush=True)
    net, p, x, t, xw, xp, tg, stride, ld = _setup(B=2, H=H, W=W, seed=11)
    out0 = torch.zeros_like(tg)
    out1 = torch.zeros_like(tg)
    net.forward(2, xw.data_ptr(), xp.data_ptr(), stride, ld, out0.data_ptr(), stride, ld)
    loss0, g0 = net.gradient(2, xw.data_ptr(), xp.data_ptr(), stride, ld, tg.data_ptr(), stride, ld)
    net.set_launch_width(1)
    net.forward(2, xw.data_ptr(), xp.data_ptr(), stride, ld, out1.data_ptr(), stride, ld)
    loss1, g1 = net.gradient(2, xw.data_ptr(), xp.data_ptr(), stride, ld, tg.data_ptr(), stride, ld)
    torch.cuda.synchronize()
    assert torch.equal(out0, out1)
    assert loss1 == loss0
    e = rel_l2(g1, g0)
    print(f"  gradient, width 1 vs default: rel_l2 = {e:.2e}")
    assert e < 1e-4
    y_ref, _ = ocnn.forward(p.astype(np.float64), x, in_scale=(0.8, 1.25))
    assert rel_l2(out1.cpu().numpy()[:, :, :W] / 0.5, y_ref) < 2e-3
    _, g_ref, _ = ocnn.loss_and_grad(p.astype(np.float64), x, t, in_scale=(0.8, 1.25),
                                     out_scale=0.5)
    G, R = ocnn.unpack(g1.astype(np.float64)), ocnn.unpack(g_ref)
    for k in R:
        assert rel_l2(G[k], R[k]) < 2e-2, k
    with pytest.raises(Exception):
        net.set_launch_width(513)
    net.set_launch_width(0)
    net.close()
