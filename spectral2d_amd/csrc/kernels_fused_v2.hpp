// kernels_fused_v2.hpp -- the fused RK3 stage kernels on lane-group FFTs
// (fft_wave.hpp), compiled for the padded sizes Np in {96, 192, 384, 768,
// 1536, 3072}. Same mathematics and data layout as kernels_fused.hpp
// (compact omega, mixed (kx, y) fields G/H); other sizes use those kernels.
//
// Elements per lane: 12 in both precisions (a 1536-point transform spans
// two waves with four LDS exchanges; 24 per lane would fit it in one wave
// with two exchanges, but at about twice the registers per wave, which
// spills pass B and halves the waves in flight).
//
//  A  one lane group per (field, kx): u_hat, v_hat or w_hat of a column,
//     inverse transform along y, stored into the rows of G (G[y*KG + kx]).
//  B  one lane group per row y (contiguous loads of its G rows): w from its
//     Hermitian spectrum, u + i v
//     (two-for-one), the products (u w + i v w) in registers, one forward
//     transform, the uw / vw spectra separated from P_k and conj P_{Np-k}.
//     Rows are assigned XCD-major so that the rows in flight on one XCD are
//     adjacent and share the 128-byte lines of the (kx, y) fields in its L2.
//  C  two halves per column group: one transforms uw, the other vw; they
//     swap halves of their results through LDS and finish the RHS,
//     truncation and Shu-Osher stage update on the compact state.
#pragma once

#include "fft_wave.hpp"
#include "kernels_fused.hpp"

namespace s2d {

template <typename T, int NP>
struct WaveCfg {
    static constexpr int E = 12;
    static constexpr int TG = NP / E;
    static constexpr int GPB = TG >= 64 ? 1 : 64 / TG; // groups per block (A, B)
    static constexpr int BLK = TG >= 64 ? TG : 64;      // threads per block (A, B)
    static constexpr int SLICE = LdsPad<T>::size(NP);
    static constexpr int WPE = sizeof(T) == 8 ? 3 : 4; // register budget, waves per SIMD
    // pass B holds w and u + i v; a multi-wave group needs more registers
    static constexpr int WPE_B = (sizeof(T) == 8 || TG > 64) ? 2 : 3;
    static_assert(E * TG == NP && TG >= 2 && TG <= 256 && (TG & (TG - 1)) == 0,
                  "unsupported group FFT size");
};

// XCD-major block order: the hardware deals blocks round-robin to 8 XCDs,
// so block b runs on XCD b % 8; give each XCD a contiguous range of work.
__device__ __forceinline__ int xcd_major(int b, int nblocks)
{
    if (nblocks % 8 != 0) return b; // the permutation below needs equal shares
    const int per = nblocks / 8;
    return (b % 8) * per + b / 8;
}

// A batch of independent runs of one padded size: run r's arrays start at
// base + r * stride; nu and the device scalars are per run.
struct Batch {
    int nruns;
    long long sW; // compact omega (Wn, Ws), elements per run
    long long sG; // each G field, elements per run
    long long sH; // each H field, elements per run
    const double *nu; // per-run viscosity, or null: StageArgs::nu for all
};

// The kernels are written as bodies over (block index blk of nblk blocks,
// the block's LDS); the __global__ wrappers below give them one size, the
// ensemble kernels (kernels_ensemble.hpp) select a size per block.

// ------------------------------------------------------------- pass A
template <typename T, int NP>
__device__ __forceinline__ void passA_body(int blk, int nblk, cpx<T> *lds,
                                           const cpx<T> *__restrict__ Wsrc,
                                           cpx<T> *__restrict__ Gu, cpx<T> *__restrict__ Gv,
                                           cpx<T> *__restrict__ Gw, const FusedDims &d,
                                           const cpx<T> *__restrict__ tw, DevScalars *sc,
                                           int reset_max, const Batch &bt)
{
#pragma clang fp contract(off)
    using C = WaveCfg<T, NP>;
    constexpr int E = C::E, TG = C::TG;
    const int g = threadIdx.x / TG, t = threadIdx.x % TG;
    // field-major jobs, XCD-major blocks: the blocks in flight on one XCD
    // hold adjacent kx of one field and fill whole lines of the G rows
    const int job = xcd_major(blk, nblk) * C::GPB + g;
    const bool active = job < 3 * d.K * bt.nruns;
    const int run = active ? job / (3 * d.K) : 0, rem = job - run * 3 * d.K;
    const int f = active ? rem / d.K : 0, kx = active ? rem % d.K : 0;
    if (reset_max && active && f == 0 && kx == 0 && t == 0) {
        sc[run].vmax_bits = 0ull;
        sc[run].bad = 0u;
    }
    const cpx<T> *src = Wsrc + run * bt.sW + (size_t)kx * d.KP;
    const T qx = (T)kx;
    cpx<T> v[E];
#pragma unroll
    for (int s = 0; s < E; ++s) {
        const int j = t + TG * s;
        const int c = compact_row(j, d);
        cpx<T> w = (active && c >= 0) ? src[c] : mk<T>((T)0, (T)0);
        const int kyi = (j <= NP / 2) ? j : j - NP;
        const T qy = (T)kyi;
        if (f == 2) {
            v[s] = w;
        } else {
            cpx<T> ps = mk<T>((T)0, (T)0);
            if (!(kx == 0 && kyi == 0) && c >= 0) {
                const T dd = (kx == 0) ? qy * qy : (kyi == 0) ? qx * qx : qx * qx + qy * qy;
                const T inv = (T)1 / dd; // one reciprocal per mode
                ps = mk<T>(w.re * inv, w.im * inv);
            }
            if (f == 0) {
                v[s] = mk<T>(-(qy * ps.im), qy * ps.re); // II*ky*psi
            } else {
                cpx<T> r = mk<T>(-(qx * ps.im), qx * ps.re); // II*kx*psi
                v[s] = mk<T>((T)-1 * r.re, (T)-1 * r.im);
            }
        }
    }
    group_fft<T, E, TG, 1>(v, lds + g * C::SLICE, t, tw);
    if (active) {
        cpx<T> *dst = (f == 0 ? Gu : f == 1 ? Gv : Gw) + run * bt.sG + kx;
#pragma unroll
        for (int s = 0; s < E; ++s) dst[(size_t)(t + TG * s) * d.KG] = v[s];
    }
}

template <typename T, int NP>
__global__ __launch_bounds__((WaveCfg<T, NP>::BLK))
__attribute__((amdgpu_waves_per_eu(WaveCfg<T, NP>::WPE))) void
k2_passA(const cpx<T> *__restrict__ Wsrc, cpx<T> *__restrict__ Gu, cpx<T> *__restrict__ Gv,
         cpx<T> *__restrict__ Gw, FusedDims d, const cpx<T> *__restrict__ tw, DevScalars *sc,
         int reset_max, Batch bt)
{
    using C = WaveCfg<T, NP>;
    __shared__ cpx<T> lds[C::GPB * C::SLICE];
    passA_body<T, NP>(blockIdx.x, gridDim.x, lds, Wsrc, Gu, Gv, Gw, d, tw, sc, reset_max, bt);
}

// ------------------------------------------------------------- pass B

// Hermitian extension of one row's x spectrum at frequency k of a length-NP
// transform (only Re of k = 0 is read).
template <typename T>
__device__ __forceinline__ cpx<T> hermitian_at(const cpx<T> *__restrict__ G, int k, int y,
                                               const FusedDims &d)
{
    const int h = d.N / 2;
    if (k <= h) {
        cpx<T> a = G[(size_t)y * d.KG + k];
        if (k == 0) a.im = (T)0;
        return a;
    }
    if (k >= d.Np - h) {
        cpx<T> a = G[(size_t)y * d.KG + (d.Np - k)];
        a.im = -a.im;
        return a;
    }
    return mk<T>((T)0, (T)0);
}

template <typename T, int NP>
__device__ __forceinline__ void passB_body(int blk, int nblk, cpx<T> *lds_all,
                                           const cpx<T> *__restrict__ Gu,
                                           const cpx<T> *__restrict__ Gv,
                                           const cpx<T> *__restrict__ Gw,
                                           cpx<T> *__restrict__ Huw, cpx<T> *__restrict__ Hvw,
                                           const FusedDims &d, const cpx<T> *__restrict__ tw,
                                           DevScalars *sc, int forward, int track_max,
                                           const Batch &bt)
{
#pragma clang fp contract(off)
    using C = WaveCfg<T, NP>;
    constexpr int E = C::E, TG = C::TG;
    const int g = threadIdx.x / TG, t = threadIdx.x % TG;
    cpx<T> *lds = lds_all + g * C::SLICE;
    const int job = xcd_major(blk, nblk) * C::GPB + g;
    const bool active = job < NP * bt.nruns;
    const int run = active ? job / NP : 0, y = job - run * NP;
    const int yy = active ? y : 0;
    const int h = d.N / 2;
    Gu += run * bt.sG;
    Gv += run * bt.sG;
    Gw += run * bt.sG;
    Huw += run * bt.sH;
    Hvw += run * bt.sH;

    // w of this row (the imaginary part of the result is the rounding of a
    // real transform and is dropped)
    cpx<T> v[E];
#pragma unroll
    for (int s = 0; s < E; ++s) v[s] = hermitian_at(Gw, t + TG * s, yy, d);
    group_fft<T, E, TG, 1>(v, lds, t, tw);
    T wr[E];
#pragma unroll
    for (int s = 0; s < E; ++s) wr[s] = v[s].re;

    // u + i v
#pragma unroll
    for (int s = 0; s < E; ++s) {
        const int k = t + TG * s;
        const cpx<T> a = hermitian_at(Gu, k, yy, d), b = hermitian_at(Gv, k, yy, d);
        v[s] = mk<T>(a.re - b.im, a.im + b.re);
    }
    group_fft<T, E, TG, 1>(v, lds, t, tw);
    T vmax = (T)0;
#pragma unroll
    for (int s = 0; s < E; ++s) {
        const T m = absv(v[s].re) + absv(v[s].im);
        vmax = (m > vmax || m != m) ? m : vmax; // NaN sticks
        v[s] = mk<T>(v[s].re * wr[s], v[s].im * wr[s]); // uw + i vw
    }
    if (forward) {
        group_fft<T, E, TG, -1>(v, lds, t, tw);
#pragma unroll
        for (int s = 0; s < E; ++s) lds[LdsPad<T>::idx(t + TG * s)] = v[s];
        group_sync<TG>();
        if (active) {
#pragma unroll
            for (int s = 0; s < E; ++s) {
                const int k = t + TG * s;
                if (k <= h) {
                    const cpx<T> P = lds[LdsPad<T>::idx(k)];
                    cpx<T> Pc = lds[LdsPad<T>::idx((NP - k) % NP)];
                    Pc.im = -Pc.im;
                    const cpx<T> su = add(P, Pc), df = sub(P, Pc);
                    Huw[(size_t)k * NP + y] = mk<T>((T)0.5 * su.re, (T)0.5 * su.im);
                    Hvw[(size_t)k * NP + y] = mk<T>((T)0.5 * df.im, (T)-0.5 * df.re);
                }
            }
        }
    }
    if (!track_max) return;
    double dm = active ? (double)vmax : 0.0;
    if (dm != dm) dm = __longlong_as_double(0x7ff8000000000000ll);
    for (int off = 32; off > 0; off >>= 1) {
        double o = __shfl_xor(dm, off);
        dm = (o > dm || o != o) ? o : dm;
    }
    // every group of a block belongs to the same run (GPB divides NP)
    if ((threadIdx.x & 63) == 0)
        atomicMax(&sc[run].vmax_bits, (unsigned long long)__double_as_longlong(dm));
}

template <typename T, int NP>
__global__ __launch_bounds__((WaveCfg<T, NP>::BLK))
__attribute__((amdgpu_waves_per_eu(WaveCfg<T, NP>::WPE_B))) void
k2_passB(const cpx<T> *__restrict__ Gu, const cpx<T> *__restrict__ Gv,
         const cpx<T> *__restrict__ Gw, cpx<T> *__restrict__ Huw, cpx<T> *__restrict__ Hvw,
         FusedDims d, const cpx<T> *__restrict__ tw, DevScalars *sc, int forward, int track_max,
         Batch bt)
{
    using C = WaveCfg<T, NP>;
    __shared__ cpx<T> lds_all[C::GPB * C::SLICE];
    passB_body<T, NP>(blockIdx.x, gridDim.x, lds_all, Gu, Gv, Gw, Huw, Hvw, d, tw, sc, forward,
                      track_max, bt);
}

// ------------------------------------------------------------- pass C
template <typename T, int NP>
__device__ __forceinline__ void passC_body(int blk, cpx<T> *lds_all,
                                           const cpx<T> *__restrict__ Huw,
                                           const cpx<T> *__restrict__ Hvw, const cpx<T> *Wn,
                                           const cpx<T> *Wst, cpx<T> *Wout, const FusedDims &d,
                                           const cpx<T> *__restrict__ tw,
                                           const int *__restrict__ trunc_i1, DevScalars *sc,
                                           const StageArgs &a, const Batch &bt)
{
#pragma clang fp contract(off)
    using C = WaveCfg<T, NP>;
    constexpr int E = C::E, TG = C::TG;
    const int half = threadIdx.x / C::BLK, lane = threadIdx.x % C::BLK; // half 0: uw, 1: vw
    const int g = lane / TG, t = lane % TG;
    cpx<T> *mine = lds_all + (half * C::GPB + g) * C::SLICE;
    cpx<T> *other = lds_all + ((1 - half) * C::GPB + g) * C::SLICE;
    const int job = blk * C::GPB + g;
    const bool active = job < d.K * bt.nruns;
    const int run = active ? job / d.K : 0, kx = active ? job - run * d.K : 0;
    const cpx<T> *src = (half == 0 ? Huw : Hvw) + run * bt.sH + (size_t)kx * NP;
    cpx<T> v[E];
#pragma unroll
    for (int s = 0; s < E; ++s) v[s] = src[t + TG * s];
    group_fft<T, E, TG, -1>(v, mine, t, tw);
    // half 0 keeps the even s and hands over the odd s, half 1 the reverse
#pragma unroll
    for (int s = 0; s < E; ++s)
        if ((s & 1) != half) mine[LdsPad<T>::idx(t + TG * s)] = v[s];
    __syncthreads();

    const double dtd = stage_dt(a, sc + run);
    const T dt = (T)dtd;
    const T qx = (T)kx;
    const double nu = bt.nu ? bt.nu[run] : a.nu;
    const T inv_norm = (T)(1.0 / a.norm); // multiplications instead of the
                                          // reference's divisions: <= 1 ulp
    const size_t col = run * bt.sW + (size_t)kx * d.KP;
#pragma unroll
    for (int s = 0; s < E; ++s) {
        if ((s & 1) != half || !active) continue;
        const int j = t + TG * s;
        const int c = compact_row(j, d);
        if (c < 0) continue;
        const int kyi = (j <= NP / 2) ? j : j - NP;
        const T qy = (T)kyi;
        cpx<T> out = mk<T>((T)0, (T)0);
        if (kx < trunc_i1[j]) {
            const cpx<T> o = other[LdsPad<T>::idx(j)];
            const cpx<T> Pr = half == 0 ? v[s] : o, Qr = half == 0 ? o : v[s];
            const cpx<T> P = mk<T>(Pr.re * inv_norm, Pr.im * inv_norm);
            const cpx<T> Q = mk<T>(Qr.re * inv_norm, Qr.im * inv_norm);
            const cpx<T> du = mk<T>(-(qx * P.im), qx * P.re); // DuDx of (uw)^
            const cpx<T> dv = mk<T>(-(qy * Q.im), qy * Q.re); // DuDy of (vw)^
            const T sn = (T)nu * (qx * qx + qy * qy);
            const cpx<T> ws = Wst[col + c];
            const cpx<T> dW = mk<T>((-du.re - dv.re) - sn * ws.re, (-du.im - dv.im) - sn * ws.im);
            const cpx<T> w0 = Wn[col + c];
            if (a.stage == 1) {
                out = mk<T>(w0.re + dt * dW.re, w0.im + dt * dW.im);
            } else if (a.stage == 2) {
                const T cc = (T)0.25 * dt;
                out = mk<T>(((T)0.75 * w0.re + (T)0.25 * ws.re) + cc * dW.re,
                            ((T)0.75 * w0.im + (T)0.25 * ws.im) + cc * dW.im);
            } else {
                const T cc = dt * (T)2;
                // (w + 2 w2 + 2 dt dW)/3 with one correctly rounded division:
                // a rounded 1/3 multiplier would bias |w| by its rounding
                // error every step
                out = mk<T>(((w0.re + (T)2 * ws.re) + cc * dW.re) / (T)3,
                            ((w0.im + (T)2 * ws.im) + cc * dW.im) / (T)3);
            }
        }
        Wout[col + c] = out;
    }
    if (a.stage == 3 && active && kx == 0 && half == 0 && t == 0) {
        sc[run].dt = dtd;
        sc[run].time = sc[run].time + dtd;
        sc[run].nadv = sc[run].nadv + 1;
    }
}

template <typename T, int NP>
__global__ __launch_bounds__((2 * WaveCfg<T, NP>::BLK))
__attribute__((amdgpu_waves_per_eu(WaveCfg<T, NP>::WPE))) void
k2_passC(const cpx<T> *__restrict__ Huw, const cpx<T> *__restrict__ Hvw, const cpx<T> *Wn,
         const cpx<T> *Wst, cpx<T> *Wout, FusedDims d, const cpx<T> *__restrict__ tw,
         const int *__restrict__ trunc_i1, DevScalars *sc, StageArgs a, Batch bt)
{
    using C = WaveCfg<T, NP>;
    __shared__ cpx<T> lds_all[2 * C::GPB * C::SLICE];
    passC_body<T, NP>(blockIdx.x, lds_all, Huw, Hvw, Wn, Wst, Wout, d, tw, trunc_i1, sc, a, bt);
}

} // namespace s2d
