"""Device ComputeSpectrum (src/Analysis.F90:267-460) against the oracle.

Shell energies agree to FP64 summation-order rounding (1e-13 of the largest
shell), the mode counts and bin count exactly; the total equals the TKE of
s2d_diagnostics; an ensemble member's spectrum equals the same run's alone.
"""
import numpy as np
import pytest

import oracle
from spectral2d_amd.ensemble import Ensemble
from spectral2d_amd.spectral import DEALIAS_3_2_RULE, FP32, FP64, Spectral2D

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,dealias", [(32, DEALIAS_3_2_RULE), (128, DEALIAS_3_2_RULE),
                                       (48, 0), (512, DEALIAS_3_2_RULE)])
def test_spectrum_matches_oracle(N, dealias):
    print(f"[spectrum] N={N} dealias={dealias}", flush=True)
    o = oracle.Oracle(N, dealias=dealias)
    w, psi = o.field(), o.field()
    o.set_turbulence(w, psi, seed=5, k0=max(2.0, N / 10), ke=0.5)
    if dealias == 0:  # every mode populated, Nyquist row / column included
        rng = np.random.default_rng(N)
        c = rng.standard_normal((o.nyG, o.cISize)) + 1j * rng.standard_normal((o.nyG, o.cISize))
        psi = o.from_complex(c)
    u, v = o.field(), o.field()
    o.compute_velocity(u, v, psi)
    ek_ref, cnt_ref, tke_ref, eps_ref, z_ref = o.spectrum(u, v, nu=1e-3)
    s = Spectral2D(N, dealias=dealias)
    s.set_field("psi", psi)
    s.set_field("w", w)
    ek, cnt, tke, eps, z = s.ComputeSpectrum(nu=1e-3)
    assert ek.shape == ek_ref.shape
    assert np.array_equal(cnt, cnt_ref)
    assert np.max(np.abs(ek - ek_ref)) <= 1e-13 * np.max(np.abs(ek_ref))
    assert abs(tke - tke_ref) <= 1e-13 * tke_ref
    assert abs(eps - eps_ref) <= 1e-12 * abs(eps_ref)
    assert abs(z - z_ref) <= 1e-12 * abs(z_ref)
    if dealias == DEALIAS_3_2_RULE:
        assert abs(s.diagnostics()["tke"] - tke) <= 1e-13 * tke
    s.close()


def test_ensemble_member_spectrum_equals_alone():
    sizes = [128, 64, 256]
    print(f"[spectrum] ensemble members {sizes}", flush=True)
    e = Ensemble(sizes, FP32)
    e.set_viscosity([1e-3, 2e-3, 5e-4])
    e.set_turbulence([7, 8, 9], [6.0, 3.0, 10.0], 0.5)
    e.steps(5, cfl=0.5)
    for r, N in enumerate(sizes):
        s = Spectral2D(N, precision=FP32)
        s.SetTurbulence(7 + r, [6.0, 3.0, 10.0][r], 0.5)
        s.rk3_steps(5, [1e-3, 2e-3, 5e-4][r], 0.0, 0.5)
        a, b = e.spectrum(r, 1e-3), s.ComputeSpectrum(1e-3)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        assert a[2:] == b[2:]
        s.close()
    e.close()
