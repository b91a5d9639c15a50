"""Writes golden.npz, the fixtures of this folder, from the CPU oracle.

The reference holds no golden vectors or known-answer files for the time
step (its one known answer, the analytic Taylor-Green decay, is checked by
the tests directly). These fixtures freeze small oracle cases, inputs and
outputs, in the reference layout (complex (nyG, cISize) arrays, row j = the
reference's Fortran column j): the CPU tests hold the oracle to them
bitwise, the GPU tests hold the HIP library to them within 1e-12.

    tg16_*      Taylor-Green at N = 16 (Np = 24): w0, then w and psi after
                10 IntegrateOneStep calls (nu 0.01, dt 1e-3)
    turb32_*    seeded turbulence at N = 32 (seed 7, k0 6, TKE 0.5): the
                initial w0 and psi0, the 10 CFL(0.5) time steps dts, then w
                and psi (nu 1e-3)
    rhs16_*     ComputeRHS at N = 16 of a seeded field (seed 3, k0 4,
                nu 1e-3): inputs w and psi, output dW and the round-tripped w
    spec32_*    ComputeSpectrum of turb32's initial field (nu 1e-3): shell
                energies ek, mode counts cnt, and tke, eps, enstrophy

Run from the repository root after building the oracle:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.npz")


def cases() -> dict:
    g = {}
    o = oracle.Oracle(16)
    st = o.new_state()
    o.set_taylor_green(st["w"], st["psi"])
    g["tg16_w0"] = o.cview(st["w"]).copy()
    for _ in range(10):
        o.step(st, 0.01, 1e-3)
    g["tg16_w"] = o.cview(st["w"]).copy()
    g["tg16_psi"] = o.cview(st["psi"]).copy()

    o = oracle.Oracle(32)
    st = o.new_state()
    o.set_turbulence(st["w"], st["psi"], seed=7, k0=6.0, ke=0.5)
    g["turb32_w0"] = o.cview(st["w"]).copy()
    g["turb32_psi0"] = o.cview(st["psi"]).copy()
    u, v = o.field(), o.field()
    o.compute_velocity(u, v, st["psi"].copy())
    ek, cnt, tke, eps, z = o.spectrum(u, v, nu=1e-3)
    g["spec32_ek"], g["spec32_cnt"] = ek, cnt
    g["spec32_tke_eps_z"] = np.array([tke, eps, z])
    dts = []
    for _ in range(10):
        dts.append(o.cfl_dt(st["psi"], 0.5))
        o.step(st, 1e-3, dts[-1])
    g["turb32_dts"] = np.array(dts)
    g["turb32_w"] = o.cview(st["w"]).copy()
    g["turb32_psi"] = o.cview(st["psi"]).copy()

    o = oracle.Oracle(16)
    w, psi = o.field(), o.field()
    o.set_turbulence(w, psi, seed=3, k0=4.0, ke=0.5)
    g["rhs16_w"] = o.cview(w).copy()
    g["rhs16_psi"] = o.cview(psi).copy()
    u, v, dW = o.field(), o.field(), o.field()
    o.compute_rhs(1e-3, u, v, w, psi, dW)
    g["rhs16_dW"] = o.cview(dW).copy()
    g["rhs16_w_out"] = o.cview(w).copy()
    return g


if __name__ == "__main__":
    np.savez_compressed(OUT, **cases())
    print("wrote", OUT, os.path.getsize(OUT), "bytes")
