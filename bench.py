#!/usr/bin/env python3
"""Benchmark of the fused TVD-RK3 time step on one MI355X.

Workload (decaying 2D turbulence, FP64 drop-in): N = 1024 resolved
(1536^2 padded for the 3/2 rule), seeded E(k) ~ k^4 exp(-(k/10)^2) scaled to
kinetic energy 0.5, nu = 1e-4, CFL 0.5 time step recomputed on the device
every step. A "step" is one full RK3 step (3 stages, 9 kernels) over the
whole field; inputs are resident in HBM before the timed region.

Prints ONE JSON line. `value` = resolved Mpts x RK3 steps per second.
`roofline` is for the dominant kernel, from HIP events recorded on the
library's stream around every launch in the timed region. `cpu_baseline`
times the CPU oracle (a C restatement of the reference algorithm) on the
host cores for a bounded number of steps of the same workload.

It also carries `closure`: the learned-closure speedup at the 5 % spectrum
bound. Held-out members start 256^2 runs (the shipped closure network,
refreshed every 8 steps) from the filtered 2048^2 DNS at t = 0.5; the
members' coarse runs go as one ensemble (one network call per refresh for
all of them, each landing on t = 1 by its own end time); at t = 1 the
ensemble-mean enstrophy spectrum is compared with the filtered DNS, and the
GPU time of the 2048^2 solves over [0.5, 1] is divided by the closure
ensemble's (both wall-clock on this GPU, one stream).

The same line carries `ensemble`: the batched FP32 ensemble of 256
independent decaying-turbulence runs (N drawn from {128, 256, 512}, k0 in
[4, 16], nu log-uniform in [1e-4, 1e-3], seeds 1000 + m, per-run CFL 0.5),
advanced by hipGraph replay of one captured step; its value is
sum_r N_r^2 x steps / s, with the roofline of its dominant kernel.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = ("Mpts·RK3-steps/s (1024² FP64, FP32 ensemble); "
          "closure speedup @5% spectrum err")
UNIT = "Mpts·RK3-steps/s"
HBM_PEAK_GBS = 8000.0

KERNELS = ("passA_col_inverse", "passB_rows", "passC_col_forward_rk3")


def alg_bytes_per_launch(N: int, s: int):
    """Algorithmic HBM bytes of one launch of each fused kernel (two-pass
    fused pseudo-spectral model; s = bytes per real, K = N/2+1, Np = 3N/2):
      A: read N*K complex + write 3*Np*K complex
      B: read 3*Np*K complex + write 2*Np*K complex
      C: read 2*Np*K + 2*N*K complex + write N*K complex."""
    K, Np, c = N // 2 + 1, 3 * N // 2, 2 * s
    return (c * K * (N + 3 * Np), c * K * (5 * Np), c * K * (2 * Np + 3 * N))


TRAFFIC_JSON = os.path.join(ROOT, "profiles", "traffic.json")
# the autotuner's choices (s2d_tune / s2d_ensemble_tune), JSON lines keyed by
# device, precision and grid: an entry is used as is, a missing one measured
TUNE_CACHE = os.path.join(ROOT, "spectral2d_amd", "data", "tune_mi355x.jsonl")
PROFILE_EVERY = 100  # kernel timing: one step in a hundred of the timed region
# kernel-name patterns of the three passes per workload
KERNEL_TAGS = {"c1": ("k2_passA<", "k2_passB<", "k2_passC<"),
               "ensemble": ("k3_ens_passA<", "k3_ens_passB<", "k3_ens_passC<")}


# the sources the three pass kernels and their launch shapes are compiled from
PASS_SOURCES = ("fft_block.hpp", "fft_wave.hpp", "kernels_ref.hpp", "kernels_fused.hpp",
                "kernels_fused_v2.hpp", "kernels_ensemble.hpp", "s2d_internal.hpp",
                "s2d_ensemble.inc")


def kernel_source_hash() -> str:
    """Hash of the pass kernels' sources (code lines, without // comments and
    blank lines): measured traffic is only reported for the kernels it was
    measured on."""
    import hashlib
    h = hashlib.sha256()
    d = os.path.join(ROOT, "spectral2d_amd", "csrc")
    for fn in PASS_SOURCES:
        with open(os.path.join(d, fn)) as f:
            code = [ln.split("//", 1)[0].rstrip() for ln in f]
        h.update(fn.encode() + b"\0" + "\n".join(c for c in code if c.strip()).encode())
    return h.hexdigest()[:16]


def pmc_summary(fetch_dir: str, write_dir: str, copy_mib: int, key: str, tags) -> dict:
    """Per-dispatch HBM bytes of the three passes from two rocprofv3 --pmc
    runs (FETCH_SIZE, WRITE_SIZE, KiB per dispatch), each scaled by what the
    counter reads for the plain copy of copy_mib MiB in the same run; stored
    under `key` in profiles/traffic.json."""
    import csv
    import glob

    def load(d, counter):
        (path,) = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        rows = [r for r in csv.DictReader(open(path)) if r["Counter_Name"] == counter]
        copies = [float(r["Counter_Value"]) for r in rows if "copyBuffer" in r["Kernel_Name"]
                  or "copy" in r["Kernel_Name"].lower()]
        factor = (copy_mib * 1024.0) / max(copies)
        per = {}
        for tag in tags:
            v = [float(r["Counter_Value"]) for r in rows if tag in r["Kernel_Name"]]
            per[tag] = sum(v) / len(v) * 1024.0 * factor if v else None
        return per, factor

    fetch, ff = load(fetch_dir, "FETCH_SIZE")
    write, wf = load(write_dir, "WRITE_SIZE")
    entry = {"source_hash": kernel_source_hash(), "fetch_factor": ff, "write_factor": wf,
             "copy_mib": copy_mib,
             "kernels": {t: {"read_bytes": fetch[t], "write_bytes": write[t],
                             "traffic": (fetch[t] + write[t]) if fetch[t] and write[t] else None}
                         for t in tags}}
    try:
        with open(TRAFFIC_JSON) as f:
            allt = json.load(f)
        if "kernels" in allt:
            allt = {}
    except (OSError, ValueError):
        allt = {}
    allt[key] = entry
    os.makedirs(os.path.dirname(TRAFFIC_JSON), exist_ok=True)
    with open(TRAFFIC_JSON, "w") as f:
        json.dump(allt, f, indent=1)
    return entry


def measured_traffic(key: str):
    """Per-kernel traffic from profiles/traffic.json when it was measured on
    these kernels and this workload; {} otherwise."""
    try:
        with open(TRAFFIC_JSON) as f:
            t = json.load(f).get(key)
    except (OSError, ValueError, AttributeError):
        return {}
    if not t or t.get("source_hash") != kernel_source_hash():
        return {}
    return {k: v.get("traffic") for k, v in t["kernels"].items()}


def ensemble_draw(nruns: int = 256, seed: int = 2024):
    """The C2 draw: sizes uniform from {128, 256, 512}, k0 ~ U[4, 16],
    nu = 10^U[-4, -3], member m seeded 1000 + m."""
    import numpy as np
    rng = np.random.default_rng(seed)
    sizes = rng.choice([128, 256, 512], nruns)
    k0 = rng.uniform(4.0, 16.0, nruns)
    nu = 10.0 ** rng.uniform(-4.0, -3.0, nruns)
    seeds = np.arange(1000, 1000 + nruns, dtype=np.uint64)
    return sizes, k0, nu, seeds


def ens_alg_bytes(sizes, s: int):
    """Per-launch algorithmic bytes of the ensemble passes (the sums over
    runs of the per-run figures of alg_bytes_per_launch)."""
    tot = [0, 0, 0]
    for N in sizes:
        for k, b in enumerate(alg_bytes_per_launch(int(N), s)):
            tot[k] += b
    return tuple(tot)


def run_ensemble(steps: int, warmup: int, cfl: float):
    import numpy as np
    from spectral2d_amd.ensemble import Ensemble
    from spectral2d_amd.spectral import FP32
    sizes, k0, nu, seeds = ensemble_draw()
    e = Ensemble(sizes, FP32)
    e.set_viscosity(nu)
    e.set_turbulence(seeds, k0, 0.5)
    sched, tuned_us = e.tune(cfl=cfl, cache=TUNE_CACHE)  # plan-time, state restored
    if warmup > 0:
        e.steps(warmup, cfl=cfl, graph=True)
    t0 = time.perf_counter()
    e.steps(steps, cfl=cfl, graph=True)
    el = time.perf_counter() - t0
    # per-kernel durations: eager launches of the same step, events on the
    # ensemble's stream (a graph replay records no per-kernel events)
    kt = max(2, min(20, steps))
    if sched == 4:  # one kernel per stage: the per-pass figures are the unified launches'
        e.set_schedule()
    e.profile_begin(9 * kt)
    e.steps(kt, cfl=cfl)
    prof = e.profile_end()
    d = e.diagnostics()
    e.close()
    per = ens_alg_bytes(sizes, 4)
    avg_ms = [ms / max(n, 1) for ms, n in prof]
    dom = max(range(3), key=lambda k: prof[k][0])
    achieved = per[dom] / (avg_ms[dom] * 1e-3) / 1e9
    step_bytes = 3 * sum(per)
    traffic = measured_traffic("ensemble")
    tags = KERNEL_TAGS["ensemble"]
    pts = float(np.sum(sizes.astype(np.float64) ** 2))
    return {
        "metric": "sum_r N_r^2 x RK3 steps / s (FP32 ensemble)",
        "value": round(pts / 1e6 * steps / el, 3),
        "unit": UNIT,
        "steps": steps, "warmup": warmup,
        "ms_per_step": round(el / steps * 1e3, 5),
        "dtype": "f32",
        "config": {"workload": "256 independent decaying-turbulence runs, N in {128, 256, 512} "
                               "(counts %s), k0 U[4,16], nu log-U[1e-4,1e-3], per-run CFL 0.5, "
                               "hipGraph replay" % [int((sizes == n).sum()) for n in (128, 256, 512)],
                   "runs": int(len(sizes))},
        "alg_bytes_per_step": step_bytes,
        "alg_gbs": round(step_bytes / (el / steps) / 1e9, 1),
        "hbm_frac_step": round(step_bytes / (el / steps) / 1e9 / HBM_PEAK_GBS, 4),
        "roofline": {"bound": "hbm", "kernel": KERNELS[dom], "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": round(traffic[tags[dom]]) if traffic.get(tags[dom]) else None},
        "kernels": {KERNELS[k]: {"avg_us": round(avg_ms[k] * 1e3, 2), "alg_bytes": per[k],
                                 "gbs": round(per[k] / (avg_ms[k] * 1e-3) / 1e9, 1),
                                 "traffic": traffic.get(tags[k])} for k in range(3)},
        "sim_time_range": [min(x["time"] for x in d), max(x["time"] for x in d)],
        "schedule": {0: "unified", 1: "per-class", 2: "per-class tuned", 3: "unified, fused pass A",
                     4: "stage queue"}[sched],
        "err_vs_fp64": ensemble_err_vs_fp64(sizes, k0, nu, seeds, cfl),
    }


def ensemble_err_vs_fp64(sizes, k0, nu, seeds, cfl: float, steps: int = 100):
    """Relative L2 distance of omega between FP32 and FP64 runs of one member
    per size (the first drawn of each), from the same initial condition over
    the same fixed steps (dt: the FP64 run's initial CFL step). An FP32 run
    alone is bitwise the ensemble member (tests/test_gpu_ensemble.py)."""
    import numpy as np
    from spectral2d_amd.spectral import FP32, FP64, Spectral2D
    out = {}
    for n in (128, 256, 512):
        m = int(np.flatnonzero(sizes == n)[0])
        w = {}
        for prec in (FP64, FP32):
            s = Spectral2D(n, precision=prec)
            s.SetTurbulence(int(seeds[m]), float(k0[m]), 0.5)
            if prec == FP64:
                dt = s.ComputeTimeStep(cfl)
            s.rk3_steps(steps, float(nu[m]), dt, 0.0)
            w[prec] = s.get_field("w")
            s.close()
        out[f"{n}^2 member {m}"] = float(np.linalg.norm(w[FP32] - w[FP64]) /
                                          np.linalg.norm(w[FP64]))
    return {"rel_l2_omega": out, "steps": steps, "dt": "fixed, the FP64 run's initial CFL step"}


def run_c1_fp32(N: int, steps: int, warmup: int, nu: float, cfl: float):
    """The same decaying-turbulence workload with an FP32 context."""
    from spectral2d_amd.spectral import FP32, Spectral2D
    sim = Spectral2D(N, precision=FP32)
    sim.SetTurbulence(seed=1, k0=10.0, ke=0.5)
    sim.set_time(0.0, 0)
    sched = sim.tune(nu, cfl=cfl, cache=TUNE_CACHE)
    sim.rk3_steps(warmup, nu, 0.0, cfl)
    t0 = time.perf_counter()
    sim.rk3_steps(steps, nu, 0.0, cfl)
    el = time.perf_counter() - t0
    sim.close()
    bps = 3 * sum(alg_bytes_per_launch(N, 4))
    return {"metric": "Mpts·RK3-steps/s (FP32 context)", "value": round(N * N / 1e6 * steps / el, 3),
            "unit": UNIT, "steps": steps, "ms_per_step": round(el / steps * 1e3, 5), "dtype": "f32",
            "alg_gbs": round(bps * steps / el / 1e9, 1),
            "schedule": schedule_summary(sched),
            "err_vs_fp64": c1_err_vs_fp64(N, nu, cfl),
            "config": {"workload": f"decaying 2D turbulence {N}^2 (padded {3 * N // 2}^2), FP32, "
                                   f"CFL {cfl}, nu {nu}"}}


def c1_err_vs_fp64(N: int, nu: float, cfl: float, steps: int = 200):
    """Relative L2 distance of omega between the FP32 and FP64 GPU paths
    from the same initial condition over the same fixed steps (dt: the FP64
    run's initial CFL step)."""
    import numpy as np
    from spectral2d_amd.spectral import FP32, FP64, Spectral2D
    w = {}
    for prec in (FP64, FP32):
        s = Spectral2D(N, precision=prec)
        s.SetTurbulence(seed=1, k0=10.0, ke=0.5)
        if prec == FP64:
            dt = s.ComputeTimeStep(cfl)
        s.rk3_steps(steps, nu, dt, 0.0)
        w[prec] = s.get_field("w")
        s.close()
    return {"rel_l2_omega": float(np.linalg.norm(w[FP32] - w[FP64]) / np.linalg.norm(w[FP64])),
            "steps": steps, "dt": "fixed, the FP64 run's initial CFL step"}


def schedule_summary(r: dict) -> dict:
    """The tuner's choice as the bench line reports it."""
    return {"passes": [r["pass_a"], r["pass_b"], r["pass_c"]], "step_graph": bool(r["step_graph"]),
            "from_cache": bool(r["from_cache"]), "tuned_us_per_step": round(r["step_us"], 2),
            "default_us_per_step": round(r["default_us"], 2)}


def run_host_exchange(N: int, steps: int, nu: float, cfl: float):
    """Steps as the Fortran IntegrateOneStep makes them by default: the host
    wC uploaded (the host arrays are the state), one step, then wC and psiC
    copied back to host buffers in the reference layout (the only fields
    valid after a step)."""
    import numpy as np
    from spectral2d_amd.spectral import FP64, Spectral2D
    sim = Spectral2D(N, precision=FP64)
    sim.SetTurbulence(seed=1, k0=10.0, ke=0.5)
    sim.set_time(0.0, 0)
    w = np.empty(2 * sim.cISize * sim.nyG)
    p = np.empty_like(w)
    sim.rk3_steps(2, nu, 0.0, cfl)
    sim.get_field("w", out=w)
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.set_field("w", w)
        sim.rk3_steps(1, nu, 0.0, cfl)
        sim.get_field("w", out=w)
        sim.get_field("psi", out=p)
    el = time.perf_counter() - t0
    sim.close()
    return {"metric": "Mpts·RK3-steps/s with wC uploaded before and wC, psiC copied back "
                      "after every step",
            "value": round(N * N / 1e6 * steps / el, 3), "unit": UNIT, "steps": steps,
            "ms_per_step": round(el / steps * 1e3, 5), "dtype": "f64",
            "host_bytes_per_step": 3 * w.nbytes}


def run_cnn_train(steps: int, warmup: int, batch: int = 16, n: int = 256):
    """The closure CNN's training step (forward, backward, Adam) at batch x n^2
    on synthetic planes with the shipped weights; the FLOP count is that of
    the six convolutions (forward, backward data except layer 0, weight
    gradients)."""
    import numpy as np
    from spectral2d_amd.closure_run import SHIPPED
    from spectral2d_amd.cnn import ClosureNet
    rng = np.random.default_rng(7)
    m = 2 * batch
    w = rng.standard_normal((m, n, n)).astype(np.float32)
    p = rng.standard_normal((m, n, n)).astype(np.float32)
    t = (0.3 * rng.standard_normal((m, n, n))).astype(np.float32)
    net = ClosureNet(n, n, batch)
    with np.load(SHIPPED) as z:
        net.set_params(z["params"].astype(np.float32))
    net.set_scales(1.0, 1.0, 1.0)
    net.set_dataset(w, p, t)
    for _ in range(warmup):
        net.train_batch(rng.choice(m, batch, replace=False), 1e-5)
    idx = [rng.choice(m, batch, replace=False) for _ in range(steps)]
    t0 = time.perf_counter()
    for i in idx:
        loss = net.train_batch(i, 1e-5)  # synchronous (returns the loss)
    el = time.perf_counter() - t0
    net.close()
    px = float(batch) * n * n
    mac = px * 9 * (2 * 64 + 4 * 64 * 64 + 64)          # forward
    flops = 2 * mac + 2 * px * 9 * (4 * 64 * 64 + 64)    # + backward data (layers 1..5)
    flops += 2 * mac                                      # + weight gradients
    ms = el / steps * 1e3
    return {"metric": "closure CNN training step (forward, backward, Adam)",
            "value": round(ms, 4), "unit": "ms per step", "higher_is_better": False,
            "batch": batch, "grid": n, "steps": steps, "dtype": "bf16 (FP32 accumulation)",
            "gflop_per_step": round(flops / 1e9, 1),
            "tflops": round(flops / (ms * 1e-3) / 1e12, 1), "loss_last": round(loss, 5)}


def run_closure(members: int):
    from spectral2d_amd.closure_run import eval_args, evaluate, load_net
    a = eval_args(eval_runs=members)
    net = load_net()
    e = evaluate(a, net, log=lambda *x: print(*x, file=sys.stderr, flush=True))
    net.close()
    if "err_closure" not in e:
        return {"metric": "closure speedup @5% spectrum err", "value": None, "unstable": e}
    return {
        "metric": "closure speedup @5% spectrum err",
        "value": round(e["speedup_vs_dns"], 2),
        "unit": "x (GPU time, 2048^2 DNS / 256^2 closure run, t 0.5 -> 1.0)",
        "spectrum_err_closure": round(e["err_closure"], 5),
        "spectrum_err_plain": round(e["err_plain"], 5),
        "within_5pct": bool(e["err_closure"] <= 0.05),
        "members": members,
        "gpu_s_dns": round(e["gpu_s_dns"], 3), "gpu_s_closure": round(e["gpu_s_closure"], 4),
        "steps": e["steps"],
        "config": {"workload": "held-out seeds 3000.., 2048^2 DNS (FP32, nu 2e-5, k0 12, CFL 0.5) "
                               "vs 256^2 with the shipped CNN closure, refreshed every 8 steps, "
                               "shell guard on, the members as one ensemble; Z(k) = k^2 E(k) "
                               "over shells 1..128"},
    }


def host_cpu() -> dict:
    """The host's CPU model, physical core count and the CPUs this process
    may use (/proc/cpuinfo, the affinity mask)."""
    model, cores = "unknown", set()
    try:
        phys = core = None
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "model name" and model == "unknown":
                model = v
            elif k == "physical id":
                phys = v
            elif k == "core id":
                core = v
            elif not k and phys is not None:
                cores.add((phys, core))
                phys = core = None
    except OSError:
        pass
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"model": model, "physical_cores": len(cores) or None, "available_cpus": avail}


def _oracle_rate(N: int, nu: float, cfl: float, threads: int, budget_s: float, max_steps: int):
    import oracle
    oracle.set_threads(threads)
    o = oracle.Oracle(N)
    st = o.new_state()
    o.set_turbulence(st["w"], st["psi"], seed=1, k0=10.0, ke=0.5)
    # one untimed step, then as many as fit the budget (at least 2)
    dt = o.cfl_dt(st["psi"], cfl)
    t0 = time.perf_counter()
    o.step(st, nu, dt)
    one = time.perf_counter() - t0
    n = int(max(2, min(max_steps, budget_s / max(one, 1e-3))))
    t0 = time.perf_counter()
    for _ in range(n):
        dt = o.cfl_dt(st["psi"], cfl)
        o.step(st, nu, dt)
    el = time.perf_counter() - t0
    return (N * N / 1e6) * n / el, n


def cpu_baseline(N: int, nu: float, cfl: float, budget_s: float, gpu_value: float):
    """The C oracle (a restatement of the reference algorithm with its own
    radix-2/3 FFT and OpenMP loops, one process) on the host: with the CPU
    share of one GPU (16 threads on these boxes, at most the CPUs available)
    and with one thread (P = 1)."""
    cpu = host_cpu()
    threads = max(1, min(16, cpu["available_cpus"] or 1))
    value, n = _oracle_rate(N, nu, cfl, threads, budget_s, 50)
    p1, n1 = _oracle_rate(N, nu, cfl, 1, min(budget_s, 5.0), 10)
    return {
        "value": value,
        "unit": UNIT,
        "cores": threads,
        "kind": "port",
        "sample": (f"{n} RK3 steps (CFL dt each step) of the {N}^2 FP64 workload on "
                   f"{threads} host threads, C oracle with its own radix-2/3 FFT"),
        "cpu_model": cpu["model"],
        "host_physical_cores": cpu["physical_cores"],
        "host_available_cpus": cpu["available_cpus"],
        "p1": {"value": p1, "sample": f"{n1} steps on 1 thread"},
        "gpu_speedup": round(gpu_value / value, 1),
    }


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--precision", choices=("fp64", "fp32"), default="fp64")
    ap.add_argument("--nu", type=float, default=1e-4)
    ap.add_argument("--cfl", type=float, default=0.5)
    ap.add_argument("--cpu-budget", type=float, default=15.0,
                    help="seconds of CPU-oracle stepping for cpu_baseline (0: skip)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run a plain device copy of MIB MiB (torch) so a counter "
                         "profile of this command holds a copy kernel of known size")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="parse two rocprofv3 --pmc output directories (FETCH_SIZE, "
                         "WRITE_SIZE runs made with --copy-calibration) into "
                         "profiles/traffic.json and exit")
    ap.add_argument("--workload", choices=("all", "c1", "ensemble", "cnn"), default="all",
                    help="c1: the 1024^2 run only; ensemble: the ensemble only (its JSON "
                         "object is printed alone); cnn: the closure CNN's training step at "
                         "256^2 x 16 alone; all: the bench line")
    ap.add_argument("--closure-members", type=int, default=4,
                    help="held-out members of the closure evaluation (0: skip)")
    ap.add_argument("--ens-steps", type=int, default=500)
    ap.add_argument("--ens-warmup", type=int, default=10)
    args = ap.parse_args(argv)

    if args.pmc_summary:
        key = "ensemble" if args.workload == "ensemble" else f"c1_{args.precision}_{args.n}"
        tags = KERNEL_TAGS["ensemble" if args.workload == "ensemble" else "c1"]
        out = pmc_summary(args.pmc_summary[0], args.pmc_summary[1],
                          args.copy_calibration or 256, key, tags)
        print(json.dumps(out, indent=1))
        return 0

    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported here "
              "(a single run does not shard across GPUs in this build)", file=sys.stderr)
        return 2

    from spectral2d_amd.spectral import FP32, FP64, Spectral2D

    if args.copy_calibration > 0:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(f"copy calibration: 3 x {args.copy_calibration} MiB read + write", file=sys.stderr)

    if args.workload == "ensemble":
        print(json.dumps(run_ensemble(args.ens_steps, args.ens_warmup, args.cfl)), flush=True)
        return 0
    if args.workload == "cnn":
        print(json.dumps(run_cnn_train(min(args.steps, 200), min(args.warmup, 10))), flush=True)
        return 0

    N = args.n
    prec = FP64 if args.precision == "fp64" else FP32
    s_bytes = 8 if prec == FP64 else 4
    sim = Spectral2D(N, precision=prec)
    sim.SetTurbulence(seed=1, k0=10.0, ke=0.5)
    sim.set_time(0.0, 0)
    # plan-time schedule choice (FFTW_MEASURE's part): cached, or measured
    # from this state, which the tuner restores
    sched = sim.tune(args.nu, cfl=args.cfl, cache=TUNE_CACHE)
    if args.warmup > 0:
        sim.rk3_steps(args.warmup, args.nu, 0.0, args.cfl)

    # timed region: K steps, replays of one captured step (hipGraph); one
    # step in every PROFILE_EVERY runs eagerly with each fused launch
    # bracketed by HIP events on the library's stream (each event pair adds
    # a dispatch gap of a few microseconds, so timing every launch would
    # slow the measured steps); rk3_steps synchronises before returning
    sim.profile_begin(9 * ((args.steps + PROFILE_EVERY - 1) // PROFILE_EVERY), PROFILE_EVERY)
    t0 = time.perf_counter()
    sim.rk3_steps(args.steps, args.nu, 0.0, args.cfl)
    t1 = time.perf_counter()
    prof = sim.profile_end()
    el = t1 - t0
    diag = sim.diagnostics()
    sim.close()

    value = (N * N / 1e6) * args.steps / el
    per = alg_bytes_per_launch(N, s_bytes)
    avg_ms = [ms / max(n, 1) for ms, n in prof]
    dom = max(range(3), key=lambda k: prof[k][0])
    achieved = per[dom] / (avg_ms[dom] * 1e-3) / 1e9
    traffic = measured_traffic(f"c1_{args.precision}_{N}")
    tags = KERNEL_TAGS["c1"]
    kernel_table = {
        KERNELS[k]: {"avg_us": round(avg_ms[k] * 1e3, 3), "timed_launches": prof[k][1],
                     "alg_bytes": per[k],
                     "gbs": round(per[k] / (avg_ms[k] * 1e-3) / 1e9, 1) if avg_ms[k] else None,
                     "traffic": traffic.get(tags[k])}
        for k in range(3)
    }
    dom_traffic = traffic.get(tags[dom])
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": UNIT,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == FP64 else "f32",
        "data": "synthetic (seeded decaying-turbulence IC generated by the library)",
        "config": {"workload": f"decaying 2D turbulence {N}^2 (padded {3 * N // 2}^2), "
                               f"{args.precision.upper()} drop-in, CFL {args.cfl}, nu {args.nu}",
                   "N": N, "padded": 3 * N // 2, "nu": args.nu, "cfl": args.cfl,
                   "sim_time": diag["time"]},
        "schedule": schedule_summary(sched),
        "roofline": {"bound": "hbm", "kernel": KERNELS[dom],
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": round(dom_traffic) if dom_traffic else None,
                     "timing": f"HIP events around every fused launch of one step in "
                               f"{PROFILE_EVERY} of the timed steps"},
        "kernels": kernel_table,
    }
    if args.workload == "all":
        out["fp32"] = run_c1_fp32(N, args.steps, args.warmup, args.nu, args.cfl)
        out["host_exchange"] = run_host_exchange(N, 100, args.nu, args.cfl)
        out["ensemble"] = run_ensemble(args.ens_steps, args.ens_warmup, args.cfl)
        if args.closure_members > 0:
            out["closure"] = run_closure(args.closure_members)
    if args.cpu_budget > 0:
        out["cpu_baseline"] = cpu_baseline(N, args.nu, args.cfl, args.cpu_budget, value)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
