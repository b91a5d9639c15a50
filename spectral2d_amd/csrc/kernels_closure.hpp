// kernels_closure.hpp -- spectral operations between grids for the filtered
// DNS data and the closure coupling: sharp-cutoff restriction of one
// context's reference-layout field onto another grid, and a scaled sum.
#pragma once

#include "fft_block.hpp"
#include "kernels_ref.hpp"

namespace s2d {

// dst(kx, ky) = src(kx, ky) for kx^2 + ky^2 <= kcut2, else 0, for fields in
// the reference layout (cI, ny), rows in FFTW order. The coefficients are
// the normalised Fourier coefficients of the same periodic field, so they
// carry over unchanged, except at a Nyquist index that one grid has and
// the other resolves as two modes: a destination Nyquist row / column
// gathers both source modes (+k and -k fold onto one sample), and a source
// Nyquist coefficient is split in halves over the destination's +k and -k.
// blockIdx.y selects a pair of planes ss / sd elements further on.
template <typename Ts, typename Td>
__global__ void k_restrict(const cpx<Ts> *__restrict__ src, int s_cI, int s_ny,
                           cpx<Td> *__restrict__ dst, int d_cI, int d_ny, long long kcut2,
                           long long ss, long long sd)
{
#pragma clang fp contract(off)
    src += (size_t)blockIdx.y * ss;
    dst += (size_t)blockIdx.y * sd;
    S2D_GRID_LOOP(idx, (size_t)d_cI * d_ny)
    {
        const int i = (int)(idx % d_cI), jd = (int)(idx / d_cI);
        const int ky = (jd <= d_ny / 2) ? jd : jd - d_ny;
        const long long k2 = (long long)i * i + (long long)ky * ky;
        double re = 0.0, im = 0.0;
        if (i < s_cI && k2 <= kcut2) {
            auto row_of = [&](int k) -> int { // source row of wavenumber k, or -1
                if (k > s_ny / 2 || k < -(s_ny / 2) + 1) return -1;
                return k >= 0 ? k : s_ny + k;
            };
            const bool d_nyq_row = (jd == d_ny / 2) && (d_ny / 2 < s_ny / 2);
            if (d_nyq_row) { // fold +ky and -ky
                for (int sg = 0; sg < 2; ++sg) {
                    const int js = row_of(sg == 0 ? ky : -ky);
                    if (js >= 0) {
                        const cpx<Ts> a = src[(size_t)i + (size_t)js * s_cI];
                        re += (double)a.re;
                        im += (double)a.im;
                    }
                }
            } else {
                int js = row_of(ky);
                double f = 1.0;
                // -s_ny/2 is not a source row: the source Nyquist row holds
                // both signs
                if (js < 0 && ky == -(s_ny / 2)) js = s_ny / 2;
                if (js == s_ny / 2 && d_ny > s_ny) f = 0.5;
                if (js >= 0) {
                    const cpx<Ts> a = src[(size_t)i + (size_t)js * s_cI];
                    re = f * (double)a.re;
                    im = f * (double)a.im;
                }
            }
            // x: the half-complex storage counts interior kx twice and the
            // Nyquist column once
            if (i == d_cI - 1 && i < s_cI - 1) {
                re *= 2.0;
                im *= 2.0;
            } else if (i == s_cI - 1 && i < d_cI - 1) {
                re *= 0.5;
                im *= 0.5;
            }
        }
        dst[idx] = mk<Td>((Td)re, (Td)im);
    }
}

// y = a x + b y over n complex elements.
template <typename T>
__global__ void k_axpby(const cpx<T> *__restrict__ x, cpx<T> *__restrict__ y, double a, double b,
                        size_t n)
{
#pragma clang fp contract(off)
    S2D_GRID_LOOP(k, n)
    {
        const cpx<T> u = x[k], v = y[k];
        y[k] = mk<T>((T)(a * (double)u.re + b * (double)v.re),
                     (T)(a * (double)u.im + b * (double)v.im));
    }
}

// Shell guard of a closure forcing F (compact layout, column kx at
// kx * KP, row c: ky = c for c <= N/2, ky = c - N - 1 above): the term -F
// changes the enstrophy of shell b (kx^2 + ky^2 in [b^2 - b + 1, b^2 + b])
// at the rate -2 sum w_kx Re(conj(W) F) (w = 1 at kx = 0, 2 elsewhere);
// shells where that rate is positive (backscatter) get F = 0. One block per
// shell (blockIdx.x) and run (blockIdx.y, s elements apart), fixed reduction
// order.
template <typename T>
__global__ __launch_bounds__(256) void k_shell_guard(cpx<T> *__restrict__ Fc,
                                                     const cpx<T> *__restrict__ Wc, int N,
                                                     int K, int KP, long long s)
{
    __shared__ double red[256];
    Fc += (size_t)blockIdx.y * s;
    Wc += (size_t)blockIdx.y * s;
    const long long b = blockIdx.x;
    const long long lo = b == 0 ? 0 : b * b - b + 1, hi = b == 0 ? 0 : b * b + b;
    const int h = N / 2;
    auto isqrt = [](long long m) { // largest r with r * r <= m
        long long r = (long long)sqrt((double)m);
        while (r * r > m) --r;
        while ((r + 1) * (r + 1) <= m) ++r;
        return r;
    };
    auto visit = [&](auto &&f) {
        for (int kx = threadIdx.x; kx < K && (long long)kx * kx <= hi; kx += blockDim.x) {
            const long long kx2 = (long long)kx * kx;
            const long long a_hi = isqrt(hi - kx2);
            const long long need = lo - kx2;
            const long long a_lo = need <= 0 ? 0 : isqrt(need - 1) + 1;
            for (long long a = a_lo; a <= a_hi && a <= h; ++a)
                for (int sg = 0; sg < (a > 0 ? 2 : 1); ++sg) {
                    const int ky = (int)(sg == 0 ? a : -a);
                    const int c = ky >= 0 ? ky : ky + N + 1;
                    f((size_t)kx * KP + c, kx);
                }
        }
    };
    double acc = 0.0;
    visit([&](size_t i, int kx) {
        const cpx<T> w = Wc[i], f = Fc[i];
        acc += (kx == 0 ? 1.0 : 2.0) * ((double)w.re * (double)f.re + (double)w.im * (double)f.im);
    });
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (red[0] < 0.0) visit([&](size_t i, int) { Fc[i] = mk<T>((T)0, (T)0); });
}

} // namespace s2d
