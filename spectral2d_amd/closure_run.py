"""The learned-closure pipeline end to end, on one GPU:

1. training data: seeded decaying-turbulence DNS runs (seeds 2000 + r,
   N_f^2, FP32 with FP64 reductions, k0 = 12, CFL 0.5); every `every` steps
   after `t_spin`, the filtered omega, psi and the subgrid flux tau on the
   N_c^2 grid (FilteredSampler);
2. training: the closure CNN on those samples (the library's forward,
   backward and Adam kernels), inputs and output normalised by their
   standard deviations;
3. evaluation: held-out seeds 3000 + m; each DNS runs to t0, its filtered
   state starts a plain N_c run and a closure run, and the DNS continues to
   t1; the ensemble-mean enstrophy spectra Z(k) = k^2 E(k) of the two coarse
   runs are compared with the filtered DNS over shells 1..N_c/2 (relative
   L2); the GPU time of the DNS over [t0, t1] is divided by that of the
   closure run.

Writes <out>/closure_cnn.npz (parameters and scales) and <out>/metrics.json.
Run: python -m spectral2d_amd.closure_run --out DIR [options]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np


def log(*a):
    print(time.strftime("%H:%M:%S"), *a, flush=True)


def make_dns(Nf, seed, k0, nu):
    from .spectral import FP32, Spectral2D
    s = Spectral2D(Nf, precision=FP32)
    s.SetTurbulence(seed, k0, 0.5)
    s.set_time(0.0, 0)
    return s


def gen_data(a):
    from .closure import FilteredSampler, advance_to
    X, T = [], []
    for r in range(a.train_runs):
        s = make_dns(a.nf, 2000 + r, a.k0, a.nu)
        sm = FilteredSampler(s, a.nc)
        advance_to(s, a.t_spin, a.nu, a.cfl)
        n = 0
        while s.get_time()[0] < a.t_end:
            s.rk3_steps(a.every, a.nu, 0.0, a.cfl)
            w, p, tau = sm.snapshot()
            X.append(np.stack([w, p]))
            T.append(tau)
            n += 1
        log(f"data run {r}: {n} samples, t = {s.get_time()[0]:.3f}")
        sm.close()
        s.close()
    return np.stack(X).astype(np.float32), np.stack(T).astype(np.float32)


def train_net(a, X, T, params0=None):
    """Adam on the dataset held by the library; cosine learning rate."""
    from .cnn import ClosureNet, param_count
    n, _, H, W = X.shape
    sw, sp, st = (float(X[:, 0].std()), float(X[:, 1].std()), float(T.std()))
    B = min(a.batch, n)
    net = ClosureNet(H, W, B)
    net.set_params(params0 if params0 is not None else init_params(a.seed, param_count()))
    net.set_scales(1.0 / sw, 1.0 / sp, st)
    net.set_dataset(X[:, 0], X[:, 1], T)
    rng = np.random.default_rng(a.seed)
    hist = []
    t0 = time.perf_counter()
    for it in range(a.iters):
        lr = a.lr * (0.5 * (1 + np.cos(np.pi * it / a.iters)))
        hist.append(net.train_batch(rng.choice(n, B, replace=False), lr))
        if it % 200 == 0 or it == a.iters - 1:
            log(f"train it {it}: loss {np.mean(hist[-50:]):.4f} (lr {lr:.2e})")
    el = time.perf_counter() - t0
    return net, (sw, sp, st), hist, el


def init_params(seed, n):
    """He-style initialisation of the network parameter vector (layout of
    include/s2d.h, s2d_cnn_*), small last layer."""
    rng = np.random.default_rng(seed)
    C = 64
    parts = [rng.standard_normal(9 * 2 * C) * np.sqrt(2.0 / 18), np.zeros(C)]
    for _ in range(4):
        parts += [rng.standard_normal(9 * C * C) * np.sqrt(2.0 / (9 * C)), np.zeros(C)]
    parts += [rng.standard_normal(9 * C) * np.sqrt(0.1 / (9 * C)), np.zeros(1)]
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == n
    return p


SHIPPED = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                       "closure_cnn_256.npz")


def load_net(path: str = SHIPPED, nc: int = 256):
    """A batch-1 closure network with stored parameters and scales."""
    from .cnn import ClosureNet
    d = np.load(path)
    net = ClosureNet(nc, nc, 1)
    net.set_params(d["params"])
    net.set_scales(*[float(v) for v in d["scales"]])
    return net


def eval_args(**kw):
    """The evaluation settings of main() (C4 defaults), overridable."""
    base = dict(nf=2048, nc=256, nu=2e-5, k0=12.0, cfl=0.5, eval_runs=64, eval_seed0=3000,
                t0=0.5, t1=1.0, closure_every=8, guard=1, batched=1)
    base.update(kw)
    return argparse.Namespace(**base)


def evaluate(a, net, log=log, extra=None):
    """Held-out members: DNS to t0, then from its filtered state a plain
    256^2 run, a closure run with `net` and, for each (name, network) of
    `extra`, a run with that network (same guard and refresh period), all
    to t1 while the DNS continues; errors of each kind's ensemble-mean
    enstrophy spectrum against the filtered DNS. a.batched (the default):
    the coarse runs of each kind go as one ensemble (one network call per
    refresh for all members, every member landing on t1 by its own end
    time); otherwise member by member."""
    if getattr(a, "batched", 1):
        return _evaluate_batched(a, net, log, extra)
    return _evaluate_per_member(a, net, log, extra)


def _batched_net(net, M):
    """A ClosureNet of batch M with `net`'s parameters and scales."""
    from .cnn import ClosureNet
    b = ClosureNet(net.H, net.W, M)
    b.set_params(net.get_params())
    b.set_scales(*net.scales)
    b.set_clip(net.clip)
    return b


def _evaluate_batched(a, net, log=log, extra=None):
    from ._lib import NO_DEALIASING
    from .closure import FilteredSampler, Timer, advance_to, enstrophy_spectrum
    from .closure import spectrum_error
    from .ensemble import Ensemble
    from .spectral import FP32, Spectral2D
    extra = dict(extra or {})
    kmax, M, span = a.nc // 2, a.eval_runs, a.t1 - a.t0
    t_dns = Timer()
    steps = {"dns": 0}
    z_dns, w0 = [], []
    coarse = Spectral2D(a.nc, precision=FP32)
    for m in range(M):
        s = make_dns(a.nf, a.eval_seed0 + m, a.k0, a.nu)
        advance_to(s, a.t0, a.nu, a.cfl)
        s.restrict_to("w", coarse, "w", kmax)
        w0.append(coarse.get_field("w"))
        with t_dns:
            steps["dns"] += advance_to(s, a.t1, a.nu, a.cfl)
        fs = FilteredSampler(s, a.nc)
        fs.filtered_state(fs.coarse)
        fs.coarse.ComputePsi("w", "psi")
        z_dns.append(enstrophy_spectrum(fs.coarse))
        fs.close()
        s.close()
        log(f"eval member {m}: DNS reached t1 ({steps['dns']} steps so far)")
    coarse.close()
    grid = Spectral2D(a.nc, dealias=NO_DEALIASING, precision=FP32)
    nets = {"plain": None, "closure": net, **extra}
    Z, times, unstable = {}, {}, {}
    for kind, kn in nets.items():
        e = Ensemble([a.nc] * M, FP32)
        e.set_viscosity(a.nu)
        for m in range(M):
            e.set_vorticity(m, w0[m])
        e.set_end_time(span)  # members start at 0 and land on t1 - t0
        bn = _batched_net(kn, M) if kn is not None else None
        t, n = Timer(), 0
        try:
            with t:
                while min(e.get_time()[0]) < span * (1 - 1e-12):
                    if n > 20000:
                        raise RuntimeError("end time not reached in 20000 steps")
                    if bn is None:
                        e.steps(50, cfl=a.cfl, graph=True)
                    else:
                        e.closure_steps(grid, bn, 48, cfl=a.cfl, every=a.closure_every,
                                        guard=bool(a.guard))
                    n += 48 if bn is not None else 50
            Z[kind] = [enstrophy_spectrum_of(e, m) for m in range(M)]
            steps[kind] = int(sum(e.get_time()[1]))
        except RuntimeError as ex:
            log(f"  {kind} ensemble unstable: {ex}")
            unstable[kind] = M
        times[kind] = t.t
        log(f"{kind}: {M} members to t1 in {t.t:.3f} s")
        if bn is not None:
            bn.close()
        e.close()
    grid.close()
    zd = np.mean(z_dns, 0)
    if "closure" not in Z:
        return {"members": M, "unstable": unstable, "steps": steps}
    err = {k: (spectrum_error(np.mean(Z[k], 0), zd, kmax) if k in Z else float("inf"))
           for k in nets}
    zp, zc = np.mean(Z["plain"], 0), np.mean(Z["closure"], 0)
    return {
        **{f"err_{k}": err[k] for k in extra},
        "unstable": unstable,
        "members": M, "batched": True,
        "t0": a.t0, "t1": a.t1,
        "err_plain": err["plain"], "err_closure": err["closure"],
        "gpu_s_dns": t_dns.t, "gpu_s_closure": times["closure"], "gpu_s_plain": times["plain"],
        "speedup_vs_dns": t_dns.t / max(times["closure"], 1e-12),
        "steps": steps,
        "Z_dns": zd[: kmax + 1].tolist(), "Z_plain": zp[: kmax + 1].tolist(),
        "Z_closure": zc[: kmax + 1].tolist(),
    }


def enstrophy_spectrum_of(e, run: int, nu: float = 0.0):
    """Z(k) = k^2 E(k) of one ensemble member (enstrophy_spectrum's)."""
    ek = e.spectrum(run, nu)[0]
    k = np.arange(ek.size, dtype=np.float64)
    return k * k * ek


def _evaluate_per_member(a, net, log=log, extra=None):
    from .closure import ClosedRun, FilteredSampler, Timer, advance_to, enstrophy_spectrum
    from .closure import spectrum_error
    extra = dict(extra or {})
    kmax = a.nc // 2
    Z = {"dns": [], "plain": [], "closure": []}
    Z.update({k: [] for k in extra})
    t_dns, t_clo, t_pln = Timer(), Timer(), Timer()
    steps = {"dns": 0, "plain": 0, "closure": 0}
    unstable = {}
    for m in range(a.eval_runs):
        s = make_dns(a.nf, a.eval_seed0 + m, a.k0, a.nu)
        advance_to(s, a.t0, a.nu, a.cfl)
        runs = {}
        nets = {"plain": None, "closure": net, **extra}
        for kind in ("plain", "closure", *extra):
            r = ClosedRun(a.nc, nets[kind], every=a.closure_every, guard=bool(a.guard))
            s.restrict_to("w", r.sim, "w", kmax)
            r.sim.ComputePsi("w", "psi")
            r.sim.set_time(s.get_time()[0], 0)
            runs[kind] = r
        with t_dns:
            steps["dns"] += advance_to(s, a.t1, a.nu, a.cfl)
        log(f"eval member {m}: DNS reached t1 ({steps['dns']} steps so far)")
        fs = FilteredSampler(s, a.nc)
        fs.filtered_state(fs.coarse)
        fs.coarse.ComputePsi("w", "psi")
        Z["dns"].append(enstrophy_spectrum(fs.coarse))
        fs.close()
        s.close()
        for kind, r in runs.items():
            try:
                with (t_clo if kind == "closure" else t_pln if kind == "plain" else Timer()):
                    steps[kind] = steps.get(kind, 0) + r.advance_to(a.t1, a.nu, a.cfl, log=log)
                Z[kind].append(enstrophy_spectrum(r.sim))
            except RuntimeError as ex:
                log(f"  {kind} run of member {m} unstable: {ex}")
                unstable[kind] = unstable.get(kind, 0) + 1
            r.close()
        if Z["plain"] and Z["closure"]:
            e_p = spectrum_error(np.mean(Z["plain"], 0), np.mean(Z["dns"], 0), kmax)
            e_c = spectrum_error(np.mean(Z["closure"], 0), np.mean(Z["dns"], 0), kmax)
            log(f"eval member {m}: ensemble errors so far plain {e_p:.4f} closure {e_c:.4f}")
    if not Z["closure"]:
        return {"members": a.eval_runs, "unstable": unstable, "steps": steps}
    zd, zp, zc = (np.mean(Z[k], 0) for k in ("dns", "plain", "closure"))
    for k in extra:
        if len(Z[k]) == len(Z["dns"]):
            out_extra = spectrum_error(np.mean(Z[k], 0), zd, kmax)
        else:  # a member of this kind went unstable
            out_extra = float("inf")
        extra[k] = out_extra
    return {
        **{f"err_{k}": v for k, v in extra.items()},
        "unstable": unstable,
        "members": a.eval_runs,
        "t0": a.t0, "t1": a.t1,
        "err_plain": spectrum_error(zp, zd, kmax),
        "err_closure": spectrum_error(zc, zd, kmax),
        "gpu_s_dns": t_dns.t, "gpu_s_closure": t_clo.t, "gpu_s_plain": t_pln.t,
        "speedup_vs_dns": t_dns.t / max(t_clo.t, 1e-12),
        "steps": steps,
        "Z_dns": zd[: kmax + 1].tolist(), "Z_plain": zp[: kmax + 1].tolist(),
        "Z_closure": zc[: kmax + 1].tolist(),
    }


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--out", required=True)
    ap.add_argument("--nf", type=int, default=2048)
    ap.add_argument("--nc", type=int, default=256)
    ap.add_argument("--nu", type=float, default=2e-5)
    ap.add_argument("--k0", type=float, default=12.0)
    ap.add_argument("--cfl", type=float, default=0.5)
    ap.add_argument("--train-runs", type=int, default=32)
    ap.add_argument("--t-spin", type=float, default=0.4)
    ap.add_argument("--t-end", type=float, default=1.0)
    ap.add_argument("--every", type=int, default=10)
    ap.add_argument("--iters", type=int, default=2500)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--eval-runs", type=int, default=64)
    ap.add_argument("--eval-seed0", type=int, default=3000)
    ap.add_argument("--t0", type=float, default=0.5)
    ap.add_argument("--t1", type=float, default=1.0)
    ap.add_argument("--clip", type=int, default=0,
                    help="1: keep only the enstrophy-sink part of tau (omega tau > 0)")
    ap.add_argument("--guard", type=int, default=1,
                    help="1: zero the closure forcing on backscattering shells")
    ap.add_argument("--per-member", action="store_true",
                    help="evaluate the coarse runs member by member instead of as ensembles")
    ap.add_argument("--closure-every", type=int, default=8,
                    help="closure evaluations every this many coarse steps (8: chosen on "
                         "validation seeds 4000.. among 4, 8, 12)")
    ap.add_argument("--weights", help="skip data and training: evaluate these weights (e.g. "
                    "spectral2d_amd/data/closure_cnn_256.npz)")
    a = ap.parse_args(argv)
    os.makedirs(a.out, exist_ok=True)
    from .cnn import ClosureNet
    metrics = {"config": vars(a)}
    if a.weights:
        net = load_net(a.weights, a.nc)
    else:
        log("generating filtered DNS data")
        t0 = time.perf_counter()
        X, T = gen_data(a)
        metrics["data"] = {"samples": int(len(X)), "seconds": time.perf_counter() - t0,
                           "std": [float(X[:, 0].std()), float(X[:, 1].std()), float(T.std())]}
        log(f"{len(X)} samples in {metrics['data']['seconds']:.1f} s")
        net, scales, hist, el = train_net(a, X, T)
        # a-priori: correlation of the prediction with tau on a few samples
        k = min(32, len(X))
        pred = np.stack([net.predict(i) for i in range(k)])
        cc = float(np.corrcoef(pred.ravel(), T[:k].ravel())[0, 1])
        metrics["train"] = {"iters": a.iters, "seconds": el, "loss_first": float(np.mean(hist[:20])),
                            "loss_last": float(np.mean(hist[-50:])), "corr_apriori": cc}
        log(f"trained: {el:.1f} s, loss {metrics['train']['loss_first']:.4f} -> "
            f"{metrics['train']['loss_last']:.4f}, a-priori correlation {cc:.3f}")
        params = net.get_params()
        np.savez(os.path.join(a.out, "closure_cnn.npz"), params=params,
                 scales=np.array(net.scales, np.float32))
        net.close()
        net = ClosureNet(a.nc, a.nc, 1)
        net.set_params(params)
        net.set_scales(*scales)
    net.set_clip(a.clip)
    a.batched = 0 if a.per_member else 1
    metrics["eval"] = evaluate(a, net)
    e = metrics["eval"]
    if "err_closure" in e:
        log(f"spectrum error plain {e['err_plain']:.4f} closure {e['err_closure']:.4f}; "
            f"DNS {e['gpu_s_dns']:.2f} s vs closure {e['gpu_s_closure']:.2f} s "
            f"({e['speedup_vs_dns']:.1f}x); unstable {e['unstable']}")
    else:
        log(f"every closure run unstable: {e}")
    with open(os.path.join(a.out, "metrics.json"), "w") as f:
        json.dump(metrics, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
