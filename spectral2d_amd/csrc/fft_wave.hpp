// fft_wave.hpp -- compile-time Stockham FFT on a lane group.
//
// A length n = E*TG transform (n = 2^a 3^b) is carried by a group of TG
// lanes (TG a power of two, 2..256). Groups of up to 64 lanes live inside
// one wave (64/TG groups per wave); larger groups span TG/64 waves of one
// workgroup. Lane t of the group holds x[t + TG*s], s < E. The early passes
// are radix-R Stockham passes (R in {8, 4, 2}, R | E, the radices multiply
// to TG) that exchange through the group's LDS slice; the final pass is one
// radix-E butterfly per lane, in registers, whose outputs X[t + TG*m] are
// again the lane's own elements.
//
// LDS slots are permuted inside aligned blocks of 16 elements (LdsPad), which
// spreads the strided Stockham writes and the contiguous reads over the
// banks without padding. Groups inside one wave synchronise with wave-scope fences;
// multi-wave groups use workgroup barriers, which every thread of the
// workgroup must reach the same number of times.
//
// Twiddles come from the FP64-generated table tw[e] = exp(-2 pi i e / n)
// of the transform length (rounded to T); the in-register DFTs use the same
// table at stride n/E, read at lane-uniform addresses.
#pragma once

#include "fft_block.hpp"

namespace s2d {

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int TG>
__device__ __forceinline__ void group_sync()
{
    if constexpr (TG > 64)
        __syncthreads();
    else
        wave_sync();
}

// LDS slot of element i: bits 0-3 XORed with bits 3-6, a permutation inside
// each aligned block of 16 elements (n is a multiple of 16 for every size
// used). It keeps the strided Stockham writes and the contiguous reads
// within 1.25-1.5x of conflict-free for both 32- and 64-bank LDS, with no
// padding.
template <typename T>
struct LdsPad {
    static __device__ __forceinline__ int idx(int i) { return i ^ ((i >> 3) & 15); }
    static constexpr int size(int n) { return n; }
};

// In-register DFT of compile-time size N on a[0..N-1] (natural order in and
// out). Base radices 2, 3, 4, 8; composite N = N1*N2 by Cooley-Tukey with
// N1 the largest base radix dividing N. Internal twiddles W_N^e are read
// from the length-n table at stride n/N.
template <typename T, int N, int D>
__device__ __forceinline__ void dft_ct(cpx<T> *a, const cpx<T> *__restrict__ tw, int n)
{
    if constexpr (N == 2 || N == 3 || N == 4 || N == 8) {
        dft<T, N, D>(a);
    } else {
        constexpr int N1 = (N % 8 == 0) ? 8 : (N % 4 == 0) ? 4 : (N % 3 == 0) ? 3 : 2;
        constexpr int N2 = N / N1;
        cpx<T> y[N];
#pragma unroll
        for (int n2 = 0; n2 < N2; ++n2) {
            cpx<T> b[N1];
#pragma unroll
            for (int n1 = 0; n1 < N1; ++n1) b[n1] = a[N2 * n1 + n2];
            dft<T, N1, D>(b);
#pragma unroll
            for (int k1 = 0; k1 < N1; ++k1) {
                if (n2 > 0 && k1 > 0) b[k1] = mul(b[k1], twiddle<T, D>(tw, (n / N) * n2 * k1));
                y[n2 * N1 + k1] = b[k1];
            }
        }
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

// Largest early radix in {8, 4, 2} dividing both E and the remaining
// product rem of the group size.
constexpr int early_radix(int E, int rem)
{
    return (E % 8 == 0 && rem % 8 == 0) ? 8 : (E % 4 == 0 && rem % 4 == 0) ? 4 : 2;
}

// One early radix-R pass with Ls = product of the previous radices (Ls
// divides TG, so k = j % Ls = t % Ls for every butterfly of the lane).
template <typename T, int E, int TG, int R, int Ls, int D>
__device__ __forceinline__ void group_pass(cpx<T> (&v)[E], cpx<T> *lds, int t, int n,
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

template <typename T, int E, int TG, int D, int Ls>
__device__ __forceinline__ void early_passes(cpx<T> (&v)[E], cpx<T> *lds, int t,
                                             const cpx<T> *__restrict__ tw)
{
    if constexpr (Ls < TG) {
        constexpr int R = early_radix(E, TG / Ls);
        static_assert(E % R == 0 && (TG / Ls) % R == 0, "no early radix");
        group_pass<T, E, TG, R, Ls, D>(v, lds, t, E * TG, tw);
        group_sync<TG>();
#pragma unroll
        for (int s = 0; s < E; ++s) v[s] = lds[LdsPad<T>::idx(t + TG * s)];
        group_sync<TG>();
        early_passes<T, E, TG, D, Ls * R>(v, lds, t, tw);
    }
}

// Final radix-E pass of lane t (Ls = TG, k = t): X[t + TG*k] = sum_m
// v[m] W_n^{m t} W_E^{m k}. With E = N1*N2 (N1 in {8, 4}) and m = N2*n1 +
// n2 the twiddle factors as W_n^{N2 n1 t} * W_n^{n2 t}: the first goes into
// a radix-N1 DFT over n1, the second (constant in n1) is applied after it
// together with the Cooley-Tukey twiddle W_E^{n2 k1}, which is uniform
// across lanes. That needs N1 - 1 + N2 - 1 per-lane twiddles, not E - 1.
template <typename T, int E, int TG, int D>
__device__ __forceinline__ void final_pass(cpx<T> (&v)[E], int t, const cpx<T> *__restrict__ tw)
{
    constexpr int n = E * TG;
    constexpr int N1 = (E % 8 == 0) ? 8 : (E % 4 == 0) ? 4 : 0;
    if constexpr (N1 == 0 || E == N1) {
#pragma unroll
        for (int m = 1; m < E; ++m) v[m] = mul(v[m], twiddle<T, D>(tw, m * t));
        dft_ct<T, E, D>(v, tw, n);
    } else {
        constexpr int N2 = E / N1;
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
__device__ __forceinline__ void group_fft(cpx<T> (&v)[E], cpx<T> *lds, int t,
                                          const cpx<T> *__restrict__ tw)
{
    early_passes<T, E, TG, D, 1>(v, lds, t, tw);
    final_pass<T, E, TG, D>(v, t, tw);
}

} // namespace s2d
