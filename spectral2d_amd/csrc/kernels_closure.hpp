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
template <typename Ts, typename Td>
__global__ void k_restrict(const cpx<Ts> *__restrict__ src, int s_cI, int s_ny,
                           cpx<Td> *__restrict__ dst, int d_cI, int d_ny, long long kcut2)
{
#pragma clang fp contract(off)
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

} // namespace s2d
