"""The HIP library against the committed golden fixtures
(tests/golden/golden.npz, oracle results frozen by
tests/golden/make_golden.py): the literal and the fused step, the CFL
steps, ComputeRHS and ComputeSpectrum on the same inputs, to the drop-in's
1e-12 (1e-13 for one RHS evaluation)."""
import os

import numpy as np
import pytest

from spectral2d_amd.spectral import Spectral2D

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden.npz")


def rel_l2(a, b):
    return np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel())


@pytest.mark.parametrize("path", ["literal", "fused"])
def test_golden_taylor_green(path):
    g = np.load(GOLDEN)
    print(f"[golden] tg16 {path}", flush=True)
    s = Spectral2D(16)
    s.SetTaylorGreen()
    assert np.max(np.abs(s.cview(s.get_field("w")) - g["tg16_w0"])) < 1e-15
    if path == "literal":
        for _ in range(10):
            s.IntegrateOneStep(0.01, 1e-3)
    else:
        s.rk3_steps(10, 0.01, 1e-3, 0.0)
    assert rel_l2(s.cview(s.get_field("w")), g["tg16_w"]) < 1e-12
    assert rel_l2(s.cview(s.get_field("psi")), g["tg16_psi"]) < 1e-12
    s.close()


def test_golden_turbulence_cfl_steps():
    g = np.load(GOLDEN)
    print("[golden] turb32 CFL steps", flush=True)
    s = Spectral2D(32)
    s.SetTurbulence(7, 6.0, 0.5)
    assert np.array_equal(s.cview(s.get_field("w")), g["turb32_w0"])
    assert np.array_equal(s.cview(s.get_field("psi")), g["turb32_psi0"])
    s.rk3_steps(10, 1e-3, 0.0, 0.5)
    t = float(np.sum(g["turb32_dts"]))
    assert abs(s.diagnostics()["time"] - t) < 1e-13 * t
    assert rel_l2(s.cview(s.get_field("w")), g["turb32_w"]) < 1e-12
    assert rel_l2(s.cview(s.get_field("psi")), g["turb32_psi"]) < 1e-12
    s.close()


def test_golden_rhs_and_spectrum():
    g = np.load(GOLDEN)
    print("[golden] rhs16, spec32", flush=True)
    s = Spectral2D(16)
    s.set_field("w", np.ascontiguousarray(g["rhs16_w"]).view(np.float64).ravel())
    s.set_field("psi", np.ascontiguousarray(g["rhs16_psi"]).view(np.float64).ravel())
    s.ComputeRHS(1e-3)
    assert rel_l2(s.cview(s.get_field("dW")), g["rhs16_dW"]) < 1e-13
    assert rel_l2(s.cview(s.get_field("w")), g["rhs16_w_out"]) < 1e-14
    s.close()
    s = Spectral2D(32)
    s.set_field("w", np.ascontiguousarray(g["turb32_w0"]).view(np.float64).ravel())
    s.set_field("psi", np.ascontiguousarray(g["turb32_psi0"]).view(np.float64).ravel())
    ek, cnt, tke, eps, z = s.ComputeSpectrum(1e-3)
    assert np.array_equal(cnt, g["spec32_cnt"])
    assert np.max(np.abs(ek - g["spec32_ek"])) <= 1e-13 * np.max(np.abs(g["spec32_ek"]))
    for got, ref in zip((tke, eps, z), g["spec32_tke_eps_z"]):
        assert abs(got - ref) <= 1e-12 * abs(ref)
    s.close()
