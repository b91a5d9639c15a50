"""bench.py's host-side contract (no GPU): one GPU only, the algorithmic
byte model of the fused step, and the traffic record keyed by the pass
kernels' sources."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


@pytest.mark.parametrize("N,s,total", [(256, 8, 30_117_888), (1024, 8, 479_084_544),
                                       (1024, 4, 239_542_272), (2048, 4, 957_235_200)])
def test_algorithmic_bytes_per_step(N, s, total):
    """3 x (A + B + C) = 6 s K (10 Np + 4 N) per RK3 step (the two-pass model)."""
    assert 3 * sum(bench.alg_bytes_per_launch(N, s)) == total


def test_traffic_record_is_reported_only_for_its_kernels(monkeypatch):
    """profiles/traffic.json is keyed by the hash of the pass kernels' code
    lines: bench.py reports a record only while the hash matches."""
    with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
        t = json.load(f)
    for key in ("c1_fp64_1024", "ensemble"):
        for k, v in t[key]["kernels"].items():
            assert v["traffic"] == v["read_bytes"] + v["write_bytes"], k
    monkeypatch.setattr(bench, "kernel_source_hash", lambda: t["c1_fp64_1024"]["source_hash"])
    assert set(bench.measured_traffic("c1_fp64_1024")) == set(bench.KERNEL_TAGS["c1"])
    monkeypatch.setattr(bench, "kernel_source_hash", lambda: "0" * 16)
    assert bench.measured_traffic("c1_fp64_1024") == {}
