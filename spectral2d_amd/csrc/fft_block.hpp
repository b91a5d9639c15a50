// fft_block.hpp -- workgroup-cooperative Stockham FFT for gfx950 (wave64).
//
// A length-n complex transform (n = 2^a 3^b) is carried by T = n/E threads.
// Thread t always holds the E elements x[t + T*s], s = 0..E-1, in registers:
// a radix-r Stockham pass over butterflies j = t + T*q (q < E/r) reads
// x[j + m*n/r] = x[t + T*(q + m*E/r)], which are exactly those registers.
// Each pass writes its outputs to LDS at the Stockham positions
// (j/Ls)*Ls*r + j%Ls + m*Ls and reads back x[t + T*s] after a barrier, so
// a transform costs (passes - 1) LDS exchanges and its first load and last
// store are coalesced along t. Radices must divide E; the radix sequence
// is a runtime plan, E and the radix set are compile-time so register
// indexing stays static. Twiddles come from a table tw[e] =
// exp(-2 pi i e / n) generated on the host in FP64 or better.
#pragma once

#include <hip/hip_runtime.h>

namespace s2d {

template <typename T>
struct __attribute__((aligned(2 * sizeof(T)))) cpx {
    T re, im;
};

template <typename T>
__device__ __forceinline__ cpx<T> mk(T r, T i)
{
    cpx<T> c;
    c.re = r;
    c.im = i;
    return c;
}

template <typename T>
__device__ __forceinline__ cpx<T> add(cpx<T> a, cpx<T> b) { return mk(a.re + b.re, a.im + b.im); }
template <typename T>
__device__ __forceinline__ cpx<T> sub(cpx<T> a, cpx<T> b) { return mk(a.re - b.re, a.im - b.im); }
template <typename T>
__device__ __forceinline__ cpx<T> mul(cpx<T> a, cpx<T> b)
{
    return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
// a * conj(b)
template <typename T>
__device__ __forceinline__ cpx<T> mulc(cpx<T> a, cpx<T> b)
{
    return mk(a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im);
}

constexpr int kMaxPasses = 24;

// Runtime transform plan: length, threads per transform and radix list.
struct FftPlan {
    int n;
    int T;
    int npass;
    int radix[kMaxPasses];
};

// DFT of size R in registers, direction D (-1 forward, +1 inverse).
template <typename T, int R, int D>
__device__ __forceinline__ void dft(cpx<T> *a)
{
    if constexpr (R == 2) {
        cpx<T> t0 = add(a[0], a[1]);
        cpx<T> t1 = sub(a[0], a[1]);
        a[0] = t0;
        a[1] = t1;
    } else if constexpr (R == 3) {
        const T h = (T)0.86602540378443864676372317075294;
        const T s = (D < 0) ? h : -h;
        cpx<T> t1 = add(a[1], a[2]);
        cpx<T> t2 = mk(a[0].re - (T)0.5 * t1.re, a[0].im - (T)0.5 * t1.im);
        cpx<T> t3 = mk(s * (a[1].re - a[2].re), s * (a[1].im - a[2].im));
        a[0] = add(a[0], t1);
        a[1] = mk(t2.re + t3.im, t2.im - t3.re);
        a[2] = mk(t2.re - t3.im, t2.im + t3.re);
    } else if constexpr (R == 4) {
        const T s = (D < 0) ? (T)1 : (T)-1;
        cpx<T> t0 = add(a[0], a[2]);
        cpx<T> t1 = sub(a[0], a[2]);
        cpx<T> t2 = add(a[1], a[3]);
        cpx<T> t3 = sub(a[1], a[3]);
        a[0] = add(t0, t2);
        a[2] = sub(t0, t2);
        a[1] = mk(t1.re + s * t3.im, t1.im - s * t3.re);
        a[3] = mk(t1.re - s * t3.im, t1.im + s * t3.re);
    } else if constexpr (R == 8) {
        // 8 = 2 x 4: two radix-4 on even/odd, twiddle by W8^m, radix-2.
        const T r2 = (T)0.70710678118654752440084436210485;
        cpx<T> e[4] = {a[0], a[2], a[4], a[6]};
        cpx<T> o[4] = {a[1], a[3], a[5], a[7]};
        dft<T, 4, D>(e);
        dft<T, 4, D>(o);
        // W8^1 = (r2, D*r2), W8^2 = (0, D), W8^3 = (-r2, D*r2)
        const T sd = (D < 0) ? (T)-1 : (T)1;
        cpx<T> o1 = mk(r2 * (o[1].re - sd * o[1].im), r2 * (o[1].im + sd * o[1].re));
        cpx<T> o2 = mk(-sd * o[2].im, sd * o[2].re);
        cpx<T> o3 = mk(-r2 * (o[3].re + sd * o[3].im), r2 * (sd * o[3].re - o[3].im));
        a[0] = add(e[0], o[0]);
        a[4] = sub(e[0], o[0]);
        a[1] = add(e[1], o1);
        a[5] = sub(e[1], o1);
        a[2] = add(e[2], o2);
        a[6] = sub(e[2], o2);
        a[3] = add(e[3], o3);
        a[7] = sub(e[3], o3);
    }
}

// Read-only twiddle fetch, conjugated for the inverse direction.
template <typename T, int D>
__device__ __forceinline__ cpx<T> twiddle(const cpx<T> *__restrict__ tw, int e)
{
    cpx<T> w = tw[e];
    if constexpr (D > 0) w.im = -w.im;
    return w;
}

// One radix-R pass of thread t: butterflies j = t + T*q on registers v,
// outputs to lds at Stockham positions.
template <typename T, int E, int R, int D>
__device__ __forceinline__ void pass_to_lds(cpx<T> (&v)[E], cpx<T> *lds, int n,
                                            int Tn, int t, int Ls,
                                            const cpx<T> *__restrict__ tw)
{
    constexpr int Q = E / R;
    const int tstride = n / (Ls * R);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int j = t + Tn * q;
        const int k = j % Ls;
        const int base = (j - k) * R + k; // (j/Ls)*Ls*R + k
        cpx<T> a[R];
#pragma unroll
        for (int m = 0; m < R; ++m) a[m] = v[q + m * Q];
        if (Ls > 1) {
#pragma unroll
            for (int m = 1; m < R; ++m) a[m] = mul(a[m], twiddle<T, D>(tw, m * k * tstride));
        }
        dft<T, R, D>(a);
#pragma unroll
        for (int m = 0; m < R; ++m) lds[base + m * Ls] = a[m];
    }
}

// Same pass but results stay in registers at the positions the last pass
// produces (Ls * R == n, so j/Ls == 0 and output j + m*Ls = t + T*(q+m*Q)).
template <typename T, int E, int R, int D>
__device__ __forceinline__ void last_pass_regs(cpx<T> (&v)[E], int n, int Tn, int t,
                                               int Ls, const cpx<T> *__restrict__ tw)
{
    constexpr int Q = E / R;
    const int tstride = n / (Ls * R);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int j = t + Tn * q;
        const int k = j; // j < n/R == Ls
        cpx<T> a[R];
#pragma unroll
        for (int m = 0; m < R; ++m) a[m] = v[q + m * Q];
        if (Ls > 1) {
#pragma unroll
            for (int m = 1; m < R; ++m) a[m] = mul(a[m], twiddle<T, D>(tw, m * k * tstride));
        }
        dft<T, R, D>(a);
#pragma unroll
        for (int m = 0; m < R; ++m) v[q + m * Q] = a[m];
    }
}

template <typename T, int E, int R>
constexpr bool radix_ok() { return (E % R) == 0; }

template <typename T, int E, int D>
__device__ __forceinline__ void do_pass_lds(int r, cpx<T> (&v)[E], cpx<T> *lds, int n,
                                            int Tn, int t, int Ls,
                                            const cpx<T> *__restrict__ tw)
{
    if constexpr (E % 8 == 0) if (r == 8) { pass_to_lds<T, E, 8, D>(v, lds, n, Tn, t, Ls, tw); return; }
    if constexpr (E % 4 == 0) if (r == 4) { pass_to_lds<T, E, 4, D>(v, lds, n, Tn, t, Ls, tw); return; }
    if constexpr (E % 3 == 0) if (r == 3) { pass_to_lds<T, E, 3, D>(v, lds, n, Tn, t, Ls, tw); return; }
    if constexpr (E % 2 == 0) if (r == 2) { pass_to_lds<T, E, 2, D>(v, lds, n, Tn, t, Ls, tw); return; }
}

template <typename T, int E, int D>
__device__ __forceinline__ void do_last_pass(int r, cpx<T> (&v)[E], int n, int Tn, int t,
                                             int Ls, const cpx<T> *__restrict__ tw)
{
    if constexpr (E % 8 == 0) if (r == 8) { last_pass_regs<T, E, 8, D>(v, n, Tn, t, Ls, tw); return; }
    if constexpr (E % 4 == 0) if (r == 4) { last_pass_regs<T, E, 4, D>(v, n, Tn, t, Ls, tw); return; }
    if constexpr (E % 3 == 0) if (r == 3) { last_pass_regs<T, E, 3, D>(v, n, Tn, t, Ls, tw); return; }
    if constexpr (E % 2 == 0) if (r == 2) { last_pass_regs<T, E, 2, D>(v, n, Tn, t, Ls, tw); return; }
}

// Full transform in registers: on entry v[s] = x[t + T*s]; on exit
// v[s] = X[t + T*s]. lds points to this transform's n-element scratch.
// Every thread of the workgroup must call this the same number of times
// (it contains __syncthreads); threads with active == false still take
// part in the barriers but their data is ignored.
template <typename T, int E, int D>
__device__ __forceinline__ void block_fft(cpx<T> (&v)[E], cpx<T> *lds, const FftPlan &p,
                                          int t, bool active,
                                          const cpx<T> *__restrict__ tw)
{
    int Ls = 1;
    const int n = p.n, Tn = p.T;
    for (int ps = 0; ps < p.npass - 1; ++ps) {
        const int r = p.radix[ps];
        if (active) do_pass_lds<T, E, D>(r, v, lds, n, Tn, t, Ls, tw);
        __syncthreads();
        if (active) {
#pragma unroll
            for (int s = 0; s < E; ++s) v[s] = lds[t + Tn * s];
        }
        __syncthreads();
        Ls *= r;
    }
    if (active) do_last_pass<T, E, D>(p.radix[p.npass - 1], v, n, Tn, t, Ls, tw);
}

} // namespace s2d
