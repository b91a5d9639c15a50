"""What the trained closure gains (the shipped network,
spectral2d_amd/data/closure_cnn_256.npz): on held-out seeds never used in
training or in its selection, 256^2 runs started from the filtered 2048^2
DNS state at t0 = 0.5 and advanced to t1 = 0.8 are compared with the
filtered DNS by the enstrophy spectrum Z(k) = k^2 E(k) over shells 1..128
(relative L2 of the ensemble mean). The closure run must stay within 5 %,
and below half the plain 256^2 run's error; the GPU-time ratio to the DNS
over the same interval is printed."""
import pytest

pytestmark = pytest.mark.gpu


def test_closure_beats_plain_coarse_run():
    from spectral2d_amd.closure_run import eval_args, evaluate, load_net
    a = eval_args(eval_runs=2, eval_seed0=3100, t1=0.8)
    print(f"[closure gain] seeds 3100, 3101: 2048^2 DNS vs 256^2 plain / closure, "
          f"t {a.t0} -> {a.t1}", flush=True)
    net = load_net()
    e = evaluate(a, net, log=lambda *x: print(*x, flush=True))
    print(f"  error plain {e['err_plain']:.4f}  closure {e['err_closure']:.4f}  "
          f"GPU time DNS {e['gpu_s_dns']:.2f} s / closure {e['gpu_s_closure']:.3f} s = "
          f"{e['speedup_vs_dns']:.1f}x")
    assert not e["unstable"]
    assert e["err_closure"] < 0.05
    assert e["err_closure"] < 0.5 * e["err_plain"]
    net.close()
