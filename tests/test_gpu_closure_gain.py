"""What the trained closure gains (the shipped network,
spectral2d_amd/data/closure_cnn_256.npz): on held-out seeds never used in
training or in its selection, 256^2 runs started from the filtered 2048^2
DNS state at t0 = 0.5 and advanced to t1 = 0.8 are compared with the
filtered DNS by the enstrophy spectrum Z(k) = k^2 E(k) over shells 1..128
(relative L2 of the ensemble mean). The closure run must stay within 5 %,
and below half the plain 256^2 run's error; the GPU-time ratio to the DNS
over the same interval is printed.

Ablations, on the same members with the same shell guard (which zeroes the
forcing on every shell it would feed enstrophy) and refresh period: the
same network with random weights (He initialisation) and with zero weights
(the guard alone: zero forcing, bitwise the plain run). The trained
network's error must be below both: the gain is the training's, not the
guard's."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_closure_beats_plain_coarse_run_and_ablations():
    from oracle import cnn as cnn_oracle
    from spectral2d_amd.closure_run import eval_args, evaluate, load_net
    a = eval_args(eval_runs=2, eval_seed0=3100, t1=0.8)
    print(f"[closure gain] seeds 3100, 3101: 2048^2 DNS vs 256^2 plain / closure / ablations, "
          f"t {a.t0} -> {a.t1}", flush=True)
    net = load_net()
    rnd, zero = load_net(), load_net()  # the trained scales, other weights
    rnd.set_params(cnn_oracle.init_params(17).astype(np.float32))
    zero.set_params(np.zeros(net.n, np.float32))
    e = evaluate(a, net, log=lambda *x: print(*x, flush=True),
                 extra={"random": rnd, "zero": zero})
    print(f"  error plain {e['err_plain']:.4f}  closure {e['err_closure']:.4f}  "
          f"random weights {e['err_random']:.4f}  zero weights {e['err_zero']:.4f}  "
          f"GPU time DNS {e['gpu_s_dns']:.2f} s / closure {e['gpu_s_closure']:.3f} s = "
          f"{e['speedup_vs_dns']:.1f}x")
    assert not e["unstable"].get("closure")
    assert e["err_closure"] < 0.05
    assert e["err_closure"] < 0.5 * e["err_plain"]
    assert e["err_zero"] == e["err_plain"]
    assert e["err_closure"] < e["err_random"]
    assert e["err_closure"] < e["err_zero"]
    for n in (net, rnd, zero):
        n.close()
