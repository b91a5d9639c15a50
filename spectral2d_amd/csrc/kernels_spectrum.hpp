// kernels_spectrum.hpp -- the shell energy spectrum of ComputeSpectrum
// (src/Analysis.F90:267-460) on the device, in FP64.
//
// Shell b holds the modes with floor(|k| + 1/2) = b (the reference's edges
// 0, 0.5, 1.5, ..., with a binary search; |k| is never a half-integer for
// integer wavenumbers, so the two agree): kx^2 + ky^2 = 0 for b = 0 and
// b^2 - b + 1 <= kx^2 + ky^2 <= b^2 + b for b >= 1. One block per shell
// walks its annulus in a fixed order and reduces in a fixed tree, so the
// result does not depend on scheduling.
#pragma once

#include "fft_block.hpp"

namespace s2d {

// integer square root: the largest r with r*r <= m (m >= 0)
__device__ __forceinline__ long long isqrt_floor(long long m)
{
    long long r = (long long)sqrt((double)m);
    while (r * r > m) --r;
    while ((r + 1) * (r + 1) <= m) ++r;
    return r;
}

// Energy density Ehat = |u_hat|^2 + |v_hat|^2 of one mode from psi_hat,
// with u = DuDy(psi) (ky = kyMG row masked) and v = -DuDx(psi) (kx = kxMG
// column masked), src/Spectral.F90:576-610, the products of
// src/Analysis.F90:380-386.
template <typename T>
__device__ __forceinline__ double mode_energy(cpx<T> p, int kx, int ky, bool mx, bool my)
{
    const double a = (double)p.re, b = (double)p.im;
    const double fy = my ? 0.0 : (double)ky, fx = mx ? 0.0 : (double)kx;
    const double ur = -(fy * b), ui = fy * a;
    const double vr = -1.0 * (-(fx * b)), vi = -1.0 * (fx * a);
    return (ur * ur + ui * ui) + (vr * vr + vi * vi);
}

// Ehat of one mode from the velocity spectra themselves, as
// ComputeSpectrum forms it from its uC, vC arguments (src/Analysis.F90:384).
template <typename T>
__device__ __forceinline__ double mode_energy_uv(cpx<T> u, cpx<T> v)
{
    const double ur = (double)u.re, ui = (double)u.im, vr = (double)v.re, vi = (double)v.im;
    return (ur * ur + ui * ui) + (vr * vr + vi * vi);
}

// ek[b] for b in [0, nbins), in the reference layout (cISize, nyG) with
// FFTW-order rows: from psi (UV false; v unused) or from the velocity
// spectra u, v (UV true). kxMG / kyMG: the Nyquist wavenumbers (masked and
// weight 1/2 at kx = 0 and kx = kxMG, src/Analysis.F90:389-395).
template <typename T, bool UV = false>
__global__ __launch_bounds__(256) void k_spectrum(const cpx<T> *__restrict__ psi,
                                                  const cpx<T> *__restrict__ v, int cISize,
                                                  int nyG, int kxMG, int kyMG, int maskX,
                                                  int maskY, double *__restrict__ ek)
{
    __shared__ double red[256];
    const long long b = blockIdx.x;
    const long long lo = b == 0 ? 0 : b * b - b + 1, hi = b == 0 ? 0 : b * b + b;
    double acc = 0.0;
    for (int kx = threadIdx.x; kx < cISize && (long long)kx * kx <= hi; kx += blockDim.x) {
        const long long kx2 = (long long)kx * kx;
        // |ky| in [ky_lo, ky_hi]
        const long long r_hi = isqrt_floor(hi - kx2);
        const long long need = lo - kx2; // ky^2 >= need
        long long r_lo = need <= 0 ? 0 : isqrt_floor(need - 1) + 1;
        const double wt = (kx == 0 || kx == kxMG) ? 0.5 : 1.0;
        const bool mx = maskX && kx == kxMG;
        for (long long a = r_lo; a <= r_hi; ++a) {
            // +a and, for a > 0, -a (FFTW rows: ky >= 0 at j = ky, ky < 0 at nyG + ky)
            for (int sgn = 0; sgn < (a > 0 ? 2 : 1); ++sgn) {
                const long long ky = sgn == 0 ? a : -a;
                if (ky > nyG / 2 || ky < -(nyG / 2) + 1) continue;
                const int j = (int)(ky >= 0 ? ky : nyG + ky);
                const bool my = maskY && ky == kyMG;
                const size_t o = (size_t)kx + (size_t)j * cISize;
                if constexpr (UV)
                    acc += wt * mode_energy_uv(psi[o], v[o]);
                else
                    acc += wt * mode_energy(psi[o], kx, (int)ky, mx, my);
            }
        }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) ek[b] = red[0];
}

} // namespace s2d
