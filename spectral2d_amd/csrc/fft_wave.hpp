// fft_wave.hpp -- compile-time Stockham FFT on a lane group of one wave.
//
// A length n = E*TG transform (n = 2^a 3^b) is carried by TG lanes of one
// wave (TG a power of two <= 64; 64/TG transforms share a wave). Lane t of
// its group holds x[t + TG*s], s < E. The early passes are radix-R
// Stockham passes (R | E, the radices multiply to TG) that exchange through
// the group's LDS slice; the final pass is one radix-E butterfly per lane,
// entirely in registers, whose outputs X[t + TG*m] are again the lane's own
// elements. So a transform costs (number of early passes) LDS exchanges:
// two for n = 1536 (8 x 8 x 24) or 3072 (8 x 8 x 48).
//
// The LDS slice is padded with one slot after every 128 bytes, which keeps
// both the strided Stockham writes and the contiguous reads free of bank
// conflicts. Only lanes of one wave touch a slice, so ordering needs
// wave-scope fences, not workgroup barriers.
//
// Twiddles come from the FP64-generated table tw[e] = exp(-2 pi i e / n)
// of the transform length (rounded to T); the in-register radix-E DFT uses
// the same table at stride n/E.
#pragma once

#include "fft_block.hpp"

namespace s2d {

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// elements per 128-byte LDS row
template <typename T>
struct LdsPad {
    static constexpr int shift = sizeof(T) == 8 ? 3 : 4;
    static __device__ __forceinline__ int idx(int i) { return i + (i >> shift); }
    static constexpr int size(int n) { return n + (n >> shift) + 1; }
};

// In-register DFT of compile-time size N on a[0..N-1] (natural order in and
// out). Base radices 2, 3, 4, 8; composite N = N1*N2 by Cooley-Tukey with
// N1 the largest base radix dividing N. Internal twiddles W_N^e are read
// from the length-n table at stride n/N (uniform across lanes).
template <typename T, int N, int D>
__device__ __forceinline__ void dft_ct(cpx<T> *a, const cpx<T> *__restrict__ tw, int n)
{
    if constexpr (N == 2 || N == 3 || N == 4 || N == 8) {
        dft<T, N, D>(a);
    } else {
        constexpr int N1 = (N % 8 == 0) ? 8 : (N % 4 == 0) ? 4 : (N % 3 == 0) ? 3 : 2;
        constexpr int N2 = N / N1;
        // step 1: for each n2, DFT_N1 over n1 of a[N2*n1 + n2]
        cpx<T> y[N];
#pragma unroll
        for (int n2 = 0; n2 < N2; ++n2) {
            cpx<T> b[N1];
#pragma unroll
            for (int n1 = 0; n1 < N1; ++n1) b[n1] = a[N2 * n1 + n2];
            dft<T, N1, D>(b);
#pragma unroll
            for (int k1 = 0; k1 < N1; ++k1) {
                // step 2: twiddle W_N^{n2 k1}
                if (n2 > 0 && k1 > 0) b[k1] = mul(b[k1], twiddle<T, D>(tw, (n / N) * n2 * k1));
                y[n2 * N1 + k1] = b[k1];
            }
        }
        // step 3: for each k1, DFT_N2 over n2 -> X[k1 + N1*k2]
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1) {
            cpx<T> c[N2];
#pragma unroll
            for (int n2 = 0; n2 < N2; ++n2) c[n2] = y[n2 * N1 + k1];
            dft_ct<T, N2, D>(c, tw, n);
#pragma unroll
            for (int k2 = 0; k2 < N2; ++k2) a[k1 + N1 * k2] = c[k2];
        }
    }
}

// Early radices of a lane group of TG lanes (their product is TG).
template <int TG>
struct GroupRadices;
template <> struct GroupRadices<2> { static constexpr int np = 1; static constexpr int r[2] = {2, 1}; };
template <> struct GroupRadices<4> { static constexpr int np = 1; static constexpr int r[2] = {4, 1}; };
template <> struct GroupRadices<8> { static constexpr int np = 1; static constexpr int r[2] = {8, 1}; };
template <> struct GroupRadices<16> { static constexpr int np = 2; static constexpr int r[2] = {4, 4}; };
template <> struct GroupRadices<32> { static constexpr int np = 2; static constexpr int r[2] = {8, 4}; };
template <> struct GroupRadices<64> { static constexpr int np = 2; static constexpr int r[2] = {8, 8}; };

// One early radix-R pass with Ls = product of the previous radices (which
// divides TG, so k = j % Ls = t % Ls for every butterfly of the lane).
template <typename T, int E, int TG, int R, int Ls, int D>
__device__ __forceinline__ void wave_pass(cpx<T> (&v)[E], cpx<T> *lds, int t, int n,
                                          const cpx<T> *__restrict__ tw)
{
    constexpr int Q = E / R;
    const int k = t & (Ls - 1);
    cpx<T> w[R];
    if constexpr (Ls > 1) {
        const int ts = n / (Ls * R);
#pragma unroll
        for (int m = 1; m < R; ++m) w[m] = twiddle<T, D>(tw, m * k * ts);
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int j = t + TG * q;
        const int base = (j - k) * R + k;
        cpx<T> a[R];
#pragma unroll
        for (int m = 0; m < R; ++m) a[m] = v[q + m * Q];
        if constexpr (Ls > 1) {
#pragma unroll
            for (int m = 1; m < R; ++m) a[m] = mul(a[m], w[m]);
        }
        dft<T, R, D>(a);
#pragma unroll
        for (int m = 0; m < R; ++m) lds[LdsPad<T>::idx(base + m * Ls)] = a[m];
    }
}

template <typename T, int E, int TG>
__device__ __forceinline__ void wave_gather(cpx<T> (&v)[E], const cpx<T> *lds, int t)
{
#pragma unroll
    for (int s = 0; s < E; ++s) v[s] = lds[LdsPad<T>::idx(t + TG * s)];
}

// Final radix-E pass of lane t (Ls = TG, k = t): X[t + TG*k] = sum_m
// v[m] W_n^{m t} W_E^{m k}. With E = 8*N2 and m = N2*n1 + n2 the twiddle
// factors as W_n^{N2 n1 t} * W_n^{n2 t}: the first goes into a radix-8 DFT
// over n1, the second (constant in n1) is applied after it together with
// the Cooley-Tukey twiddle W_E^{n2 k1}, which is uniform across lanes.
// That needs 7 + (N2 - 1) per-lane twiddles instead of E - 1.
template <typename T, int E, int TG, int D>
__device__ __forceinline__ void final_pass(cpx<T> (&v)[E], int t, const cpx<T> *__restrict__ tw)
{
    constexpr int n = E * TG;
    if constexpr (E % 8 != 0) {
#pragma unroll
        for (int m = 1; m < E; ++m) v[m] = mul(v[m], twiddle<T, D>(tw, m * t));
        dft_ct<T, E, D>(v, tw, n);
    } else {
        constexpr int N1 = 8, N2 = E / 8;
        cpx<T> r[N1];
#pragma unroll
        for (int n1 = 1; n1 < N1; ++n1) r[n1] = twiddle<T, D>(tw, N2 * n1 * t);
        cpx<T> y[E];
#pragma unroll
        for (int n2 = 0; n2 < N2; ++n2) {
            cpx<T> b[N1];
            b[0] = v[n2];
#pragma unroll
            for (int n1 = 1; n1 < N1; ++n1) b[n1] = mul(v[N2 * n1 + n2], r[n1]);
            dft<T, N1, D>(b);
            if (n2 > 0) {
                const cpx<T> s = twiddle<T, D>(tw, n2 * t);
#pragma unroll
                for (int k1 = 0; k1 < N1; ++k1) {
                    b[k1] = mul(b[k1], s);
                    if (k1 > 0) b[k1] = mul(b[k1], twiddle<T, D>(tw, (n / E) * n2 * k1));
                }
            }
#pragma unroll
            for (int k1 = 0; k1 < N1; ++k1) y[n2 * N1 + k1] = b[k1];
        }
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1) {
            cpx<T> c[N2];
#pragma unroll
            for (int n2 = 0; n2 < N2; ++n2) c[n2] = y[n2 * N1 + k1];
            dft_ct<T, N2, D>(c, tw, n);
#pragma unroll
            for (int k2 = 0; k2 < N2; ++k2) v[k1 + N1 * k2] = c[k2];
        }
    }
}

// Full transform of the group: on entry v[s] = x[t + TG*s], on exit
// v[s] = X[t + TG*s]. lds is the group's slice (LdsPad<T>::size(n)).
template <typename T, int E, int TG, int D>
__device__ __forceinline__ void wave_fft(cpx<T> (&v)[E], cpx<T> *lds, int t,
                                         const cpx<T> *__restrict__ tw)
{
    constexpr int n = E * TG;
    using GR = GroupRadices<TG>;
    static_assert(E % GR::r[0] == 0 && (GR::np < 2 || E % GR::r[1] == 0), "radix must divide E");
    wave_pass<T, E, TG, GR::r[0], 1, D>(v, lds, t, n, tw);
    wave_sync();
    wave_gather<T, E, TG>(v, lds, t);
    wave_sync();
    if constexpr (GR::np == 2) {
        wave_pass<T, E, TG, GR::r[1], GR::r[0], D>(v, lds, t, n, tw);
        wave_sync();
        wave_gather<T, E, TG>(v, lds, t);
        wave_sync();
    }
    final_pass<T, E, TG, D>(v, t, tw);
}

} // namespace s2d
