"""The oracle against the committed golden fixtures (tests/golden/golden.npz,
written by tests/golden/make_golden.py): a change to the oracle's
arithmetic shows here as a bitwise difference."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
import make_golden  # noqa: E402

GOLDEN = os.path.join(ROOT, "tests", "golden", "golden.npz")


def test_oracle_reproduces_golden_bitwise():
    g = np.load(GOLDEN)
    new = make_golden.cases()
    assert sorted(new) == sorted(g.files)
    for k in g.files:
        assert new[k].dtype == g[k].dtype and new[k].shape == g[k].shape, k
        assert np.array_equal(new[k], g[k]), k


def test_golden_taylor_green_is_the_analytic_decay():
    # the frozen 10-step TG run decays as w0 exp(-2 nu t) (nu 0.01, t 0.01)
    g = np.load(GOLDEN)
    exact = g["tg16_w0"] * np.exp(-2 * 0.01 * 10 * 1e-3)
    assert np.linalg.norm(g["tg16_w"] - exact) <= 1e-13 * np.linalg.norm(exact)
