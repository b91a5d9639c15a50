"""Pins the CPU oracle (oracle/s2d_oracle.c) with checks that do not
involve the oracle's own formulas: brute-force DFTs, numpy's pocketfft,
closed-form tables, single-mode identities, the 3/2-rule convolution
identity, conservation laws, the RK3 amplification factor, third-order
convergence and the analytic Taylor-Green decay. CPU only."""
import itertools

import numpy as np
import pytest

import oracle


def brute_dft(x, sign):
    n = len(x)
    j = np.arange(n)
    return np.exp(sign * 2j * np.pi * np.outer(j, j) / n) @ x


# ---------------------------------------------------------------- FFT

@pytest.mark.parametrize("n", [1, 2, 3, 4, 6, 8, 9, 12, 16, 18, 24, 27, 32, 36, 48, 54, 96])
def test_fft_brute_force(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for sign in (-1, 1):
        ref = brute_dft(x, sign)
        got = oracle.fft_c2c(x, sign)
        assert np.max(np.abs(got - ref)) < 1e-13 * max(1, n)


@pytest.mark.parametrize("n", [192, 384, 768, 1536, 3072])
def test_fft_vs_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    s = np.sqrt(n)
    assert np.max(np.abs(oracle.fft_c2c(x, -1) - np.fft.fft(x))) < 2e-15 * s * np.log2(n)
    assert np.max(np.abs(oracle.fft_c2c(x, 1) - np.fft.ifft(x) * n)) < 2e-15 * s * np.log2(n)


def test_fft_rejects_other_primes():
    with pytest.raises(ValueError):
        oracle.fft_c2c(np.ones(10), -1)


@pytest.mark.parametrize("N", [8, 16, 24, 32, 64, 96])
def test_transform_r2c_vs_numpy(N):
    o = oracle.Oracle(N)
    rng = np.random.default_rng(N)
    r = rng.standard_normal((o.nyG, o.nxG))
    buf = o.from_real(r)
    o.transform_r2c(buf)
    ref = np.fft.rfft2(r) / (o.nxG * o.nyG)  # Spectral.F90:760-762
    assert np.max(np.abs(o.cview(buf) - ref)) < 1e-16 * np.log2(o.nxG * o.nyG) * 4


@pytest.mark.parametrize("N", [8, 16, 32, 64])
def test_transform_c2r_vs_numpy(N):
    """Unnormalized c2r; Im of the kx=0 and Nyquist columns is ignored after
    the column transform, as numpy's irfft does."""
    o = oracle.Oracle(N)
    rng = np.random.default_rng(N + 1)
    c = rng.standard_normal((o.nyG, o.cISize)) + 1j * rng.standard_normal((o.nyG, o.cISize))
    buf = o.from_complex(c)
    o.transform_c2r(buf)
    ref = np.fft.irfft2(c, s=(o.nyG, o.nxG)) * (o.nxG * o.nyG)
    assert np.max(np.abs(o.rview(buf)[:, : o.nxG] - ref)) < 1e-13 * np.sqrt(o.nxG * o.nyG)


def test_transform_roundtrip_identity():
    o = oracle.Oracle(32)
    rng = np.random.default_rng(3)
    r = rng.standard_normal((o.nyG, o.nxG))
    buf = o.from_real(r)
    o.transform_r2c(buf)
    o.transform_c2r(buf)
    assert np.max(np.abs(o.rview(buf)[:, : o.nxG] - r)) < 1e-14


def test_single_mode_coefficients():
    """cos(3x) cos(2y) -> four coefficients of 1/4 (1/2 after Hermitian
    folding in x is not applied: r2c stores kx >= 0 only)."""
    o = oracle.Oracle(16)
    x = 2 * np.pi * np.arange(o.nxG) / o.nxG
    y = 2 * np.pi * np.arange(o.nyG) / o.nyG
    r = np.cos(3 * x)[None, :] * np.cos(2 * y)[:, None]
    buf = o.from_real(r)
    o.transform_r2c(buf)
    c = o.cview(buf)
    expect = np.zeros_like(c)
    expect[2, 3] = 0.25
    expect[o.nyG - 2, 3] = 0.25
    assert np.max(np.abs(c - expect)) < 1e-15


# ------------------------------------------------------- grid / tables

@pytest.mark.parametrize("nx,ny", [(96, 96), (64, 32), (256, 256), (1024, 1024), (2048, 2048)])
def test_grid_closed_form(nx, ny):
    o = oracle.Oracle(nx, ny)
    g = o.g
    assert (g.nxG, g.nyG) == (3 * nx // 2, 3 * ny // 2)
    assert (g.kxMG, g.kyMG, g.kyLG) == (g.nxG // 2, g.nyG // 2, -g.nyG // 2 + 1)
    assert (g.kxMU, g.kyMU, g.kyLU) == (nx // 2, ny // 2, -ny // 2 + 1)
    assert g.kMax == nx // 2
    assert (g.cISize, g.rISize) == (g.nxG // 2 + 1, 2 * (g.nxG // 2 + 1))
    assert np.array_equal(o.kx, np.arange(g.cISize))
    assert np.array_equal(o.ky, np.fft.fftfreq(g.nyG, 1.0 / g.nyG).astype(int) * 1
                          + np.where(np.arange(g.nyG) == g.nyG // 2, g.nyG, 0))


def test_no_dealiasing_and_undefined_methods():
    o = oracle.Oracle(16, dealias=oracle.NO_DEALIASING)
    assert (o.nxG, o.nyG) == (16, 16)
    for d in (oracle.DEALIAS_2_3_RULE, oracle.DEALIAS_GRID_SHIFT):
        with pytest.raises(ValueError):
            oracle.Oracle(16, dealias=d)


@pytest.mark.parametrize("N", [8, 16, 96, 256])
def test_truncation_is_the_disk(N):
    o = oracle.Oracle(N)
    kx, ky = np.meshgrid(o.kx, o.ky)
    disk = np.sqrt(kx.astype(float) ** 2 + ky.astype(float) ** 2) <= N / 2 + 1e-10
    keep = o.kept_mask()
    # SetupTruncation zeroes from the first |k| > kMax on, to the row end
    assert np.array_equal(keep, disk)
    i1 = o.truncation()
    assert i1.min() == 0 and i1.max() == N // 2 + 1


def test_masks():
    o = oracle.Oracle(32)
    xb, xi, yb, yi = o.mask()
    assert (xb, xi) == (1, o.cISize - 1)       # kx = nxG/2
    assert (yb, yi) == (1, o.nyG // 2)         # ky = nyG/2


# ---------------------------------------------------- spectral operators

def _mode_field(o, kx, ky, amp=1.0 + 0.5j):
    c = np.zeros((o.nyG, o.cISize), complex)
    c[ky % o.nyG, kx] = amp
    return o.from_complex(c)


def test_derivatives_single_mode():
    o = oracle.Oracle(16)
    for kx, ky in [(1, 0), (3, -2), (0, 5), (7, 7)]:
        f = _mode_field(o, kx, ky)
        dx, dy = o.field(), o.field()
        o.dudx(f, dx)
        o.dudy(f, dy)
        assert np.allclose(o.cview(dx)[ky % o.nyG, kx], 1j * kx * (1 + 0.5j), atol=0, rtol=1e-15)
        assert np.allclose(o.cview(dy)[ky % o.nyG, kx], 1j * ky * (1 + 0.5j), atol=0, rtol=1e-15)
    # Nyquist masks (Spectral.F90:690-692, 735-737)
    f = _mode_field(o, o.cISize - 1, 3)
    dx = o.field()
    o.dudx(f, dx)
    assert np.all(dx == 0)
    f = _mode_field(o, 2, o.nyG // 2)
    dy = o.field()
    o.dudy(f, dy)
    assert np.all(dy == 0)


def test_psi_solve_inverts_laplacian():
    o = oracle.Oracle(24)
    rng = np.random.default_rng(9)
    c = rng.standard_normal((o.nyG, o.cISize)) + 1j * rng.standard_normal((o.nyG, o.cISize))
    w = o.from_complex(c)
    p = o.field()
    o.compute_psi(w, p)
    kx, ky = np.meshgrid(o.kx.astype(float), o.ky.astype(float))
    k2 = kx ** 2 + ky ** 2
    pc = o.cview(p)
    assert pc[0, 0] == 0
    m = k2 > 0
    assert np.max(np.abs(pc[m] * k2[m] - c[m]) / np.abs(c[m])) < 3e-16


def test_velocity_divergence_free_and_vorticity_roundtrip():
    o = oracle.Oracle(24)
    rng = np.random.default_rng(4)
    c = rng.standard_normal((o.nyG, o.cISize)) + 1j * rng.standard_normal((o.nyG, o.cISize))
    psi = o.from_complex(c)
    o.truncate(psi)
    u, v = o.field(), o.field()
    o.compute_velocity(u, v, psi)
    kx, ky = np.meshgrid(o.kx.astype(float), o.ky.astype(float))
    div = 1j * kx * o.cview(u) + 1j * ky * o.cview(v)
    assert np.max(np.abs(div)) < 1e-12
    # w = dv/dx - du/dy = |k|^2 psi
    w = o.field()
    o.compute_vorticity(u, v, w)
    assert np.max(np.abs(o.cview(w) - (kx ** 2 + ky ** 2) * o.cview(psi))) < 1e-10


def _bandlimited(o, kcut, seed):
    """Random Hermitian field with modes strictly inside |k| < kcut."""
    rng = np.random.default_rng(seed)
    r = rng.standard_normal((o.nyG, o.nxG))
    buf = o.from_real(r)
    o.transform_r2c(buf)
    kx, ky = np.meshgrid(o.kx.astype(float), o.ky.astype(float))
    c = o.cview(buf)
    c[np.sqrt(kx ** 2 + ky ** 2) >= kcut] = 0
    c[0, 0] = 0
    return buf


def _full_plane(o, buf):
    """dict (kx, ky) -> coefficient over the full plane (Hermitian extension)."""
    c = o.cview(buf)
    d = {}
    for j, ky in enumerate(o.ky):
        for i, kx in enumerate(o.kx):
            if c[j, i] != 0:
                d[(int(kx), int(ky))] = c[j, i]
                if kx > 0:
                    d[(-int(kx), -int(ky))] = np.conj(c[j, i])
    return d


def test_three_halves_rule_is_the_truncated_convolution():
    """dW(nu=0) = -(i k . (u w)^) where (u w)^ is the spectral convolution
    with wrap-around modulo the padded grid, truncated to the disk: a
    brute-force sum over mode pairs, independent of any FFT."""
    N = 8
    o = oracle.Oracle(N)
    w = _bandlimited(o, N / 2 + 0.5, 11)  # includes the boundary modes
    psi = o.field()
    o.compute_psi(w, psi)
    u, v = o.field(), o.field()
    o.compute_velocity(u, v, psi)
    W, U, V = _full_plane(o, w), _full_plane(o, u), _full_plane(o, v)
    Np = o.nxG

    def wrap(k):
        k = k % Np
        return k - Np if k > Np // 2 else k

    conv_u, conv_v = {}, {}
    for (p, up), (q, wq) in itertools.product(U.items(), W.items()):
        k = (wrap(p[0] + q[0]), wrap(p[1] + q[1]))
        conv_u[k] = conv_u.get(k, 0) + up * wq
    for (p, vp), (q, wq) in itertools.product(V.items(), W.items()):
        k = (wrap(p[0] + q[0]), wrap(p[1] + q[1]))
        conv_v[k] = conv_v.get(k, 0) + vp * wq
    ww = w.copy()
    dW = o.field()
    o.compute_rhs(0.0, o.field(), o.field(), ww, psi, dW)
    got = o.cview(dW)
    keep = o.kept_mask()
    err = 0.0
    for j, ky in enumerate(o.ky):
        for i, kx in enumerate(o.kx):
            k = (int(kx), int(ky))
            ref = -(1j * kx * conv_u.get(k, 0) + 1j * ky * conv_v.get(k, 0)) if keep[j, i] else 0
            err = max(err, abs(got[j, i] - ref))
    assert err < 1e-13


@pytest.mark.parametrize("N", [16, 32])
def test_inviscid_conservation(N):
    """Galerkin-truncated nonlinear term conserves energy and enstrophy for
    fields free of the boundary alias (|k| < N/2 - 1)."""
    o = oracle.Oracle(N)
    w = _bandlimited(o, N / 2 - 1, 5 + N)
    psi = o.field()
    o.compute_psi(w, psi)
    wc, pc = o.cview(w).copy(), o.cview(psi).copy()
    dW = o.field()
    o.compute_rhs(0.0, o.field(), o.field(), w.copy(), psi, dW)
    n = o.cview(dW)
    wt = np.where(o.kx[None, :] == 0, 1.0, 2.0)
    e_rate = np.sum(wt * np.real(np.conj(pc) * n))
    z_rate = np.sum(wt * np.real(np.conj(wc) * n))
    scale = np.sum(wt * np.abs(n) * (np.abs(wc) + np.abs(pc)))
    assert abs(e_rate) < 1e-13 * scale
    assert abs(z_rate) < 1e-13 * scale


def test_rhs_translation_equivariance():
    o = oracle.Oracle(16)
    w = _bandlimited(o, 8.5, 21)
    psi = o.field()
    o.compute_psi(w, psi)
    dW = o.field()
    o.compute_rhs(1e-2, o.field(), o.field(), w.copy(), psi, dW)
    # shift by (3, 5) padded grid points: multiply by exp(-i k . d)
    kx, ky = np.meshgrid(o.kx.astype(float), o.ky.astype(float))
    ph = np.exp(-2j * np.pi * (3 * kx + 5 * ky) / o.nxG)
    ws = o.from_complex(o.cview(w) * ph)
    ps = o.from_complex(o.cview(psi) * ph)
    dWs = o.field()
    o.compute_rhs(1e-2, o.field(), o.field(), ws, ps, dWs)
    assert np.max(np.abs(o.cview(dWs) - o.cview(dW) * ph)) < 1e-12 * np.max(np.abs(o.cview(dW)))


# ------------------------------------------------------------ stepping

def test_rk3_amplification_single_shell():
    """Taylor-Green has a zero nonlinear term, so one step multiplies every
    mode by R(z) = 1 + z + z^2/2 + z^3/6, z = -nu |k|^2 dt."""
    o = oracle.Oracle(16)
    st = o.new_state()
    o.set_taylor_green(st["w"], st["psi"])
    w0 = o.cview(st["w"]).copy()
    nu, dt = 0.05, 0.01
    o.step(st, nu, dt)
    z = -nu * 2 * dt
    R = 1 + z + z * z / 2 + z ** 3 / 6
    m = np.abs(w0) > 1e-3
    assert np.max(np.abs(o.cview(st["w"])[m] / w0[m] - R)) < 1e-14


def test_taylor_green_analytic():
    o = oracle.Oracle(32)
    st = o.new_state()
    o.set_taylor_green(st["w"], st["psi"])
    w0 = o.cview(st["w"]).copy()
    nu, dt, n = 0.01, 1e-3, 300
    for _ in range(n):
        o.step(st, nu, dt)
    exact = w0 * np.exp(-2 * nu * n * dt)
    assert np.linalg.norm(o.cview(st["w"]) - exact) / np.linalg.norm(exact) < 1e-12
    # the sampling offset x_i = i dx (SetIC.F90:91-94): array element p holds
    # x = (p+1) dx, so 2 sin x sin y has w_hat(1,1) = -1/2 exp(i (dx + dy))
    dx = 2 * np.pi / o.nxG
    assert abs(w0[1, 1] - (-0.5 * np.exp(2j * dx))) < 1e-15


def test_third_order_convergence():
    N, nu, T = 16, 1e-2, 0.2
    o = oracle.Oracle(N)
    w, psi = o.field(), o.field()
    o.set_turbulence(w, psi, seed=2, k0=3.0, ke=0.5)

    def run(nsteps):
        st = o.new_state()
        st["w"][:] = w
        st["psi"][:] = psi
        for _ in range(nsteps):
            o.step(st, nu, T / nsteps)
        return o.cview(st["w"]).copy()

    ref = run(256)
    e = [np.linalg.norm(run(n) - ref) for n in (16, 32, 64)]
    r1, r2 = e[0] / e[1], e[1] / e[2]
    assert 6.5 < r1 < 9.5 and 6.5 < r2 < 9.5, (r1, r2)


# ------------------------------------------------- IC and diagnostics

def test_turbulence_ic_recipe():
    N = 64
    o = oracle.Oracle(N)
    w, psi = o.field(), o.field()
    o.set_turbulence(w, psi, seed=7, k0=6.0, ke=0.5)
    wc = o.cview(w)
    # truncated, zero mean, Hermitian on kx = 0
    assert np.all(wc[~o.kept_mask()] == 0) and wc[0, 0] == 0
    for j in range(1, o.nyG):
        assert wc[j, 0] == np.conj(wc[(o.nyG - j) % o.nyG, 0])
    # TKE = 0.5 through the reference's ComputeSpectrum weights
    u, v = o.field(), o.field()
    o.compute_velocity(u, v, psi)
    ek, cnt, tke, eps, ed = o.spectrum(u, v, nu=1e-4)
    assert abs(tke - 0.5) < 1e-14
    # Parseval: 1/2 <u^2 + v^2> in physical space equals the spectral TKE
    o.transform_c2r(u)
    o.transform_c2r(v)
    ur, vr = o.rview(u)[:, : o.nxG], o.rview(v)[:, : o.nxG]
    assert abs(0.5 * np.mean(ur ** 2 + vr ** 2) - 0.5) < 1e-13
    # spectrum shape ~ k^4 exp(-(k/k0)^2) on shell averages
    k = np.arange(len(ek))
    ratio = ek[2:20] / (k[2:20] ** 4 * np.exp(-(k[2:20] / 6.0) ** 2))
    assert np.std(ratio) / np.mean(ratio) < 0.2


def test_uniform_hash_range_and_determinism():
    vals = np.array([oracle.uniform(1, kx, ky) for kx in range(20) for ky in range(-20, 20)])
    assert vals.min() >= 0 and vals.max() < 1
    assert abs(vals.mean() - 0.5) < 0.05
    assert oracle.uniform(3, 4, -5) == oracle.uniform(3, 4, -5)
    assert oracle.uniform(3, 4, -5) != oracle.uniform(4, 4, -5)


def test_spectrum_bins_closed_form():
    """Bin counts: shells [0,.5), [.5,1.5), ... with weight 1 at kx in
    {0, kxMG} and 2 otherwise, counted by brute force."""
    o = oracle.Oracle(16)
    u = o.field()
    ek, cnt, *_ = o.spectrum(u, u)
    nb = oracle.spectrum_bins(o.g.kxMG, o.g.kyMG)
    assert len(cnt) == nb == int(np.sqrt(o.g.kxMG ** 2 + o.g.kyMG ** 2) + 1) + 1
    ref = np.zeros(nb, int)
    for ky in o.ky:
        for kx in o.kx:
            km = np.hypot(kx, ky)
            b = 0 if km < 0.5 else min(int(np.floor(km + 0.5)), nb - 1)
            ref[b] += 1 if kx in (0, o.g.kxMG) else 2
    assert np.array_equal(cnt, ref)


def test_cfl_dt_definition():
    o = oracle.Oracle(32)
    w, psi = o.field(), o.field()
    o.set_turbulence(w, psi, seed=3, k0=4.0, ke=0.5)
    dt = o.cfl_dt(psi, 0.5)
    u, v = o.field(), o.field()
    o.compute_velocity(u, v, psi, trans=True)
    m = np.max(np.abs(o.rview(u)[:, : o.nxG]) + np.abs(o.rview(v)[:, : o.nxG]))
    assert dt == 0.5 * (2 * np.pi / 32) / m


def test_shell_index_integer_form_matches_reference_edges():
    """The device spectrum bins mode m = kx^2 + ky^2 into shell
    floor(sqrt(m) + 1/2) by integer bounds (b^2 - b + 1 <= m <= b^2 + b);
    the reference searches edges 0, 0.5, 1.5, ..., kIntMax - 0.5,
    kIntMax + 0.1 (src/Analysis.F90:358-362, :475-513). Equal for every m
    of a 3072^2 padded grid."""
    kxMG = kyMG = 1536
    nb = oracle.spectrum_bins(kxMG, kyMG)
    kint = nb - 1
    edges = np.concatenate([[0.0], 0.5 + np.arange(kint, dtype=np.float64), [kint + 0.1]])
    edges = edges[: nb + 1]
    m = np.arange(0, kxMG * kxMG + kyMG * kyMG + 1, dtype=np.int64)
    ref = np.searchsorted(edges, np.sqrt(m.astype(np.float64)), side="right") - 1
    b = np.floor(np.sqrt(m.astype(np.float64))).astype(np.int64)
    b -= (b * b > m)
    b += (m > b * b + b)
    assert np.array_equal(b, ref)
