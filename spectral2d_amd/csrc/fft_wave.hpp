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
// Per-lane twiddles come from the FP64-generated table tw[e] =
// exp(-2 pi i e / n) of the transform length (rounded to T); the twiddles
// inside the in-register DFTs are compile-time constants.
#pragma once

#include <type_traits>

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
//
// Every exchange addresses slot b + c of a lane-dependent b and a constant c
// that share no set bit. idx distributes over such sums as idx(b) ^ idx(c),
// and as it only changes bits 0-3 that is (idx(b) ^ (idx(c) & 15)) +
// (c & ~15): with idx(b) computed once per pass, an access costs one XOR
// with a constant, and the addition folds into the LDS instruction's offset.
template <typename T>
struct LdsPad {
    static __host__ __device__ constexpr int idx(int i) { return i ^ ((i >> 3) & 15); }
    // idx(b + c) from ib = idx(b), for b & c == 0
    static __host__ __device__ constexpr int idx_sum(int ib, int c)
    {
        return (ib ^ (idx(c) & 15)) + (c & ~15);
    }
    static constexpr int size(int n) { return n; }
};

// idx_sum against idx over every pair of disjoint 8-bit values (idx reads
// bits 0-6 only, so higher bits just add).
constexpr bool lds_idx_sum_exact()
{
    for (int b = 0; b < 256; ++b)
        for (int c = 0; c < 256; ++c)
            if ((b & c) == 0 && LdsPad<float>::idx_sum(LdsPad<float>::idx(b), c) !=
                                    LdsPad<float>::idx(b + c))
                return false;
    return true;
}
static_assert(lds_idx_sum_exact(), "idx_sum differs from idx");

// Compile-time loop: f(std::integral_constant<int, I>) for I in [B, N).
template <int B, int N, typename F>
__device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (B < N) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, N>(f);
    }
}

// cos(15 a degrees), a in 0..6: the constant twiddles of sizes dividing 24.
constexpr double kCos15[7] = {1.0,
                              0.96592582628906828674974319972890,
                              0.86602540378443864676372317075294,
                              0.70710678118654752440084436210485,
                              0.5,
                              0.25881904510252076234889883762405,
                              0.0};
constexpr double cos15(int a)
{
    a = ((a % 24) + 24) % 24;
    return a <= 6 ? kCos15[a] : a <= 12 ? -kCos15[12 - a] : a <= 18 ? -kCos15[a - 12] : kCos15[24 - a];
}
constexpr double sin15(int a) { return cos15(a - 6); }

// x * exp(D i 2 pi e / N) for compile-time e and N | 24 (D = -1 forward,
// +1 inverse): multiples of a quarter turn are exact swaps and negations,
// the rest one complex product with constants.
template <typename T, int D, int N, int e>
__device__ __forceinline__ cpx<T> twiddle_const(cpx<T> x)
{
    static_assert(24 % N == 0, "constant twiddles cover sizes dividing 24");
    constexpr int a = (((24 / N) * e) % 24 + 24) % 24; // the angle in 15-degree units
    if constexpr (a == 0) {
        return x;
    } else if constexpr (a == 12) {
        return mk<T>(-x.re, -x.im);
    } else if constexpr (a == 6 || a == 18) { // W = (0, D) or (0, -D)
        constexpr int sg = a == 6 ? D : -D;
        return sg > 0 ? mk<T>(-x.im, x.re) : mk<T>(x.im, -x.re);
    } else {
        constexpr T c = (T)cos15(a), sn = (T)(D * sin15(a));
        return mul(x, mk<T>(c, sn));
    }
}

// In-register DFT of compile-time size N on a[0..N-1] (natural order in and
// out). Base radices 2, 3, 4, 8; composite N = N1*N2 by Cooley-Tukey with
// N1 the largest base radix dividing N, and compile-time twiddles.
template <typename T, int N, int D>
__device__ __forceinline__ void dft_ct(cpx<T> *a)
{
    if constexpr (N == 2 || N == 3 || N == 4 || N == 8) {
        dft<T, N, D>(a);
    } else {
        constexpr int N1 = (N % 8 == 0) ? 8 : (N % 4 == 0) ? 4 : (N % 3 == 0) ? 3 : 2;
        constexpr int N2 = N / N1;
        cpx<T> y[N];
        static_for<0, N2>([&](auto n2I) {
            constexpr int n2 = decltype(n2I)::value;
            cpx<T> b[N1];
#pragma unroll
            for (int n1 = 0; n1 < N1; ++n1) b[n1] = a[N2 * n1 + n2];
            dft<T, N1, D>(b);
            static_for<0, N1>([&](auto k1I) {
                constexpr int k1 = decltype(k1I)::value;
                y[n2 * N1 + k1] = twiddle_const<T, D, N, n2 * k1>(b[k1]);
            });
        });
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1) {
            cpx<T> c[N2];
#pragma unroll
            for (int n2 = 0; n2 < N2; ++n2) c[n2] = y[n2 * N1 + k1];
            dft_ct<T, N2, D>(c);
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

// Number of per-lane twiddles of the early passes before the one at Ls
// (the first pass, Ls = 1, needs none).
template <int E, int TG, int Ls, int L = 1>
constexpr int early_tw_offset()
{
    if constexpr (L >= Ls || L >= TG) {
        return 0;
    } else {
        constexpr int R = early_radix(E, TG / L);
        return (L > 1 ? R - 1 : 0) + early_tw_offset<E, TG, Ls, L * R>();
    }
}

// The per-lane twiddles of one lane-group transform: R - 1 per early pass
// and E - 1 for the final pass, as forward values (an inverse transform
// multiplies by their conjugates). PRE selects which are loaded in one
// batch ahead of the transform's LDS exchanges (bit 0: the early passes',
// bit 1: the final pass's); the others are loaded by their pass. A load
// issued after an exchange's fence cannot be moved above it, so per-pass
// loads wait once per pass; batched ones hold registers for the whole
// transform.
constexpr int TW_EARLY = 1, TW_FINAL = 2, TW_ALL = 3, TW_FINAL_POW = 4;

template <typename T, int E, int TG, int PRE = TW_ALL>
struct GroupTw {
    static constexpr int n = E * TG;
    static constexpr int NE = early_tw_offset<E, TG, TG>();
    static constexpr bool PE = (PRE & TW_EARLY) != 0, PF = (PRE & TW_FINAL) != 0;
    // bit 2: the final pass's twiddles W^(m t) as powers of W^t, one complex
    // multiplication each instead of a load (at most about E ulps apart
    // from the table's values)
    static constexpr bool PP = (PRE & TW_FINAL_POW) != 0;
    cpx<T> e[PE && NE > 0 ? NE : 1];
    cpx<T> f[PF ? E : 1];
    cpx<T> w1; // W^t (PP)
    const cpx<T> *tw;
    int t;

    __device__ __forceinline__ void load(const cpx<T> *__restrict__ tw_, int t_)
    {
        tw = tw_;
        t = t_;
        if constexpr (PE) load_early<1>();
        if constexpr (PP) w1 = tw[t];
        if constexpr (PF) {
#pragma unroll
            for (int m = 1; m < E; ++m) f[m] = tw[m * t];
        }
    }
    // twiddle m (1 <= m < R) of the radix-R early pass at Ls
    template <int Ls, int R>
    __device__ __forceinline__ cpx<T> early(int m) const
    {
        if constexpr (PE) return e[early_tw_offset<E, TG, Ls>() + m - 1];
        else return tw[m * (t & (Ls - 1)) * (n / (Ls * R))];
    }
    // twiddle m (1 <= m < E) of the final pass
    __device__ __forceinline__ cpx<T> final(int m) const
    {
        if constexpr (PF) return f[m];
        else return tw[m * t];
    }

  private:
    template <int Ls>
    __device__ __forceinline__ void load_early()
    {
        if constexpr (Ls < TG) {
            constexpr int R = early_radix(E, TG / Ls);
            if constexpr (Ls > 1) {
                constexpr int o = early_tw_offset<E, TG, Ls>(), ts = n / (Ls * R);
                const int k = t & (Ls - 1);
#pragma unroll
                for (int m = 1; m < R; ++m) e[o + m - 1] = tw[m * k * ts];
            }
            load_early<Ls * R>();
        }
    }
};

// w for the forward direction, conj(w) for the inverse.
template <typename T, int D>
__device__ __forceinline__ cpx<T> dir(cpx<T> w)
{
    if constexpr (D > 0) w.im = -w.im;
    return w;
}

// One early radix-R pass with Ls = product of the previous radices (Ls
// divides TG, so k = j % Ls = t % Ls for every butterfly of the lane).
template <typename T, int E, int TG, int R, int Ls, int D, class G>
__device__ __forceinline__ void group_pass(cpx<T> (&v)[E], cpx<T> *lds, int t, const G &g)
{
    constexpr int Q = E / R;
    const int k = t & (Ls - 1);
    // slot of output m of butterfly q: (t - k) R + k, the lane's part, plus
    // TG q R + m Ls, which shares no bit with it
    const int ib = LdsPad<T>::idx((t - k) * R + k);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        cpx<T> a[R];
#pragma unroll
        for (int m = 0; m < R; ++m) a[m] = v[q + m * Q];
        if constexpr (Ls > 1) {
#pragma unroll
            for (int m = 1; m < R; ++m) a[m] = mul(a[m], dir<T, D>(g.template early<Ls, R>(m)));
        }
        dft<T, R, D>(a);
#pragma unroll
        for (int m = 0; m < R; ++m) lds[LdsPad<T>::idx_sum(ib, TG * q * R + m * Ls)] = a[m];
    }
}

template <typename T, int E, int TG, int D, int Ls, class G>
__device__ __forceinline__ void early_passes(cpx<T> (&v)[E], cpx<T> *lds, int t, const G &g)
{
    if constexpr (Ls < TG) {
        constexpr int R = early_radix(E, TG / Ls);
        static_assert(E % R == 0 && (TG / Ls) % R == 0, "no early radix");
        group_pass<T, E, TG, R, Ls, D, G>(v, lds, t, g);
        group_sync<TG>();
        const int it = LdsPad<T>::idx(t);
#pragma unroll
        for (int s = 0; s < E; ++s) v[s] = lds[LdsPad<T>::idx_sum(it, TG * s)];
        group_sync<TG>();
        early_passes<T, E, TG, D, Ls * R, G>(v, lds, t, g);
    }
}

// Final radix-E pass of lane t (Ls = TG, k = t): X[t + TG*k] = sum_m
// v[m] W_n^{m t} W_E^{m k}: the E - 1 per-lane twiddles W_n^{m t}, then an
// in-register E-point DFT whose own twiddles are compile-time constants.
template <typename T, int E, int TG, int D, class G>
__device__ __forceinline__ void final_pass(cpx<T> (&v)[E], const G &g)
{
    if constexpr (G::PP) {
        cpx<T> p = g.w1;
        v[1] = mul(v[1], dir<T, D>(p));
#pragma unroll
        for (int m = 2; m < E; ++m) {
            p = mul(p, g.w1);
            v[m] = mul(v[m], dir<T, D>(p));
        }
    } else {
#pragma unroll
        for (int m = 1; m < E; ++m) v[m] = mul(v[m], dir<T, D>(g.final(m)));
    }
    dft_ct<T, E, D>(v);
}

// Full transform of the group: on entry v[s] = x[t + TG*s], on exit
// v[s] = X[t + TG*s]. lds is the group's slice (LdsPad<T>::size(n)); g holds
// the lane's twiddles (GroupTw::load with the table of length n).
template <typename T, int E, int TG, int D, int PRE>
__device__ __forceinline__ void group_fft(cpx<T> (&v)[E], cpx<T> *lds, int t,
                                          const GroupTw<T, E, TG, PRE> &g)
{
    early_passes<T, E, TG, D, 1>(v, lds, t, g);
    final_pass<T, E, TG, D>(v, g);
}

} // namespace s2d
