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
//     inverse transform along y. A block holds LPB adjacent columns of one
//     field; it stages them in LDS and stores the G rows (G[y*KG + kx]) as
//     LPB-element contiguous segments.
//  B  one lane group per row y, or per two rows from Np = 1536 up
//     (contiguous loads of the G rows): w from its Hermitian spectrum (two
//     rows' w from one transform of w_a + i w_b), u + i v (two-for-one), the
//     products (u w + i v w) in registers, one forward transform, the uw /
//     vw spectra separated from P_k and conj P_{Np-k}. A block holds LPB
//     adjacent rows and stores
//     the H columns (H[kx*Np + y]) as LPB-element segments through LDS.
//     Blocks are assigned XCD-major so that the tiles in flight on one XCD
//     are adjacent and share the lines of the (kx, y) fields in its L2.
//  C  two halves per column group: one transforms uw, the other vw; they
//     swap halves of their results through LDS and finish the RHS,
//     truncation and Shu-Osher stage update on the compact state (all its
//     loads and stores are contiguous).
#pragma once

#include <type_traits>

#include "fft_wave.hpp"
#include "kernels_fused.hpp"

namespace s2d {

// Schedule variants of the pass kernels, a bit mask V (0: the default
// schedule below). A tuner (s2d_tune) times the compiled variants of each
// pass and picks one per context; every variant computes the same
// mathematics (the twiddle bits change rounding at the ulp level, the
// others only the order of work).
//   VAR_TW_A / _B / _C  the pass's twiddles the other way: all loaded in one
//                       batch (TW_ALL) where the default forms the final
//                       pass's as powers, or the reverse
//   VAR_PAIR            pass B with one lane group per two rows where the
//                       default has one per row, or the reverse
//   VAR_E24             24 elements per lane (radix-8 early passes and a
//                       radix-24 final pass) instead of 12
//   VAR_FUSE_A          pass A with one block per column tile for all three
//                       fields (the column loaded and psi formed once, three
//                       transforms in turn) instead of one block per field
constexpr int VAR_TW_A = 1, VAR_TW_B = 2, VAR_TW_C = 4, VAR_PAIR = 8, VAR_E24 = 16,
              VAR_FUSE_A = 32;

constexpr int tw_flip(int tw) { return tw == TW_ALL ? (TW_EARLY | TW_FINAL_POW) : TW_ALL; }

template <typename T, int NP, int V = 0>
struct WaveCfg {
    static constexpr int E = (V & VAR_E24) ? 24 : 12;
    static constexpr int TG = NP / E;
    // lines (columns / rows) per block of passes A and B, LPB * TG threads
    static constexpr int LPB_A = TG >= 256 ? (sizeof(T) == 4 ? 2 : 1) : 256 / TG;
    // pass B: two rows per lane group from Np = 1536 up (passB_pair)
    static constexpr bool PAIR_B = (NP >= 1536) != ((V & VAR_PAIR) != 0);
    static constexpr int LPB_B = PAIR_B ? 2 : TG >= 128 && sizeof(T) == 8 ? 1 : TG >= 256 ? 2 : 256 / TG;
    static constexpr int THR_B = PAIR_B ? LPB_B * TG / 2 : LPB_B * TG; // lanes per block
    // pass C: column groups per half, two halves per block
    static constexpr int GPB = TG >= 64 ? 1 : 64 / TG;
    static constexpr int BLK = TG >= 64 ? TG : 64; // threads per half
    static constexpr int SLICE = LdsPad<T>::size(NP);
    // register budgets, waves per SIMD, of passes A and C (FP64 pass C: 256
    // VGPRs hold a transform's data, its twiddles and the loads issued ahead
    // of it without spills; FP64 pass A fits three waves in 168 by loading
    // only the early-pass twiddles in one batch; 24 elements per lane hold
    // twice the data: two waves)
    static constexpr int WPE_A = (V & VAR_FUSE_A) ? (sizeof(T) == 8 ? 2 : 3)
                                 : E == 24 ? 2 : sizeof(T) == 8 ? 3 : 4;
    static constexpr int WPE_C = E == 24 ? 2 : sizeof(T) == 8 ? 2 : TG <= 64 ? 5 : 4;
    // pass C: the compact-state operands loaded ahead of the transform, and
    // the twiddles in one batch; FP64 and the ensemble's FP32 sizes form the
    // final pass's as powers of W^t (FP64: 168 VGPRs, three waves per SIMD,
    // 7 % less time at Np = 1536; FP32 Np <= 768: five waves, 7 % less
    // time for the mixed ensemble)
    static constexpr int TW_PRE_C0 =
        sizeof(T) == 8 || TG <= 64 ? (TW_EARLY | TW_FINAL_POW) : TW_ALL;
    static constexpr int TW_PRE_C = (V & VAR_TW_C) ? tw_flip(TW_PRE_C0) : TW_PRE_C0;
    static constexpr bool OPS_EARLY_C = true;
    // pass A: FP64 loads its early-pass twiddles in one batch (1.3 % less
    // time at Np = 1536 than all in one batch at two waves) and forms the
    // final pass's as powers of W^t (6 % less time than loading them per
    // use); FP32 loads all in one batch for the two-wave groups
    // (Np >= 1536: 5 % less time), and for the ensemble's sizes loads the
    // early-pass ones per pass and forms the final pass's as powers (4 %
    // less time for the mixed ensemble than all in one batch, then 2.6 %
    // less than loading the final ones too)
    static constexpr int TW_PRE_A0 =
        sizeof(T) == 8 ? (TW_EARLY | TW_FINAL_POW) : TG >= 128 ? TW_ALL : TW_FINAL_POW;
    static constexpr int TW_PRE_A = (V & VAR_TW_A) ? tw_flip(TW_PRE_A0) : TW_PRE_A0;
    // pass B: the lane's twiddles stay in registers for all of its row
    // transforms; FP64 has room next to the data (256 VGPRs) for the
    // early-pass ones only (measured: 9 % less time at Np = 1536 than
    // loading every twiddle per pass), and forms the final pass's as powers
    // of W^t (5 % less time than loading them)
    static constexpr int TW_PRE_B0 = sizeof(T) == 4 ? TW_ALL : (TW_EARLY | TW_FINAL_POW);
    static constexpr int TW_PRE_B = (V & VAR_TW_B) ? tw_flip(TW_PRE_B0) : TW_PRE_B0;
    // (one row per lane group: every twiddle in one batch by default)
    static constexpr int TW_PRE_B1 = (V & VAR_TW_B) ? tw_flip(TW_ALL) : TW_ALL;
    // pass B holds w and u + i v; a multi-wave group needs more registers
    static constexpr int WPE_B = (sizeof(T) == 8 || TG > 64 || E == 24) ? 2 : 3;
    static_assert(E * TG == NP && TG >= 2 && TG <= 256 && (TG & (TG - 1)) == 0,
                  "unsupported group FFT size");
};

// Tiles of passes A and B: LPB lines of NP elements (one lane group per
// line, or per two rows in passB_pair).
// The LDS slices are NP + 16 elements apart, so that the transposed reads
// (LPB slices at one index) spread evenly over the banks.
template <typename T, int NP, int LPB, int V = 0>
struct TileCfg {
    static constexpr int TG = WaveCfg<T, NP, V>::TG;
    static constexpr int BLK = LPB * TG;
    static constexpr int SL = NP + 16;
    static constexpr int LDS = LPB * SL;
    static_assert(LPB >= 1 && (LPB & (LPB - 1)) == 0 && BLK <= 1024, "bad tile");
};

// Component-wise select (a select of the whole struct can be lowered
// through a private-memory array indexed by the condition).
template <typename T>
__device__ __forceinline__ cpx<T> sel(bool c, const cpx<T> &a, const cpx<T> &b)
{
    return mk<T>(c ? a.re : b.re, c ? a.im : b.im);
}

// 1/x from the hardware reciprocal and Newton steps with FMA (one in FP32,
// two in FP64): within one ulp of the division, at a third of its cost.
__device__ __forceinline__ float recip(float x)
{
    const float r = __builtin_amdgcn_rcpf(x);
    return __builtin_fmaf(r, __builtin_fmaf(-x, r, 1.0f), r);
}
__device__ __forceinline__ double recip(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    r = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
    return __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
}

// x / 3, correctly rounded, without a division: q = RN(x * RN(1/3)) is
// within one ulp of x/3, the remainder x - 3q is exact under FMA, and one
// correction q + r * RN(1/3) rounds to RN(x/3) (Markstein's theorem).
template <typename T>
__device__ __forceinline__ T div3(T x)
{
    constexpr T y = (T)(1.0L / 3.0L);
    const T q = x * y;
    return fma(fma((T)-3, q, x), y, q);
}

// Where the padded index k = t + TG*S (0 <= t < TG) of a length-NP
// transform falls under the 3/2 rule (h = N/2 = NP/3), for the whole lane
// range of slot S: low (k <= h, kept ky >= 0 / kept kx), high
// (k >= NP - h, kept ky < 0 / the conjugate side), zero (the padding), or
// mixed (decided per lane). The loads of each slot are then unconditional,
// so a lane's loads are all in flight together.
enum PadRegion { PAD_LOW = 0, PAD_HIGH = 1, PAD_ZERO = 2, PAD_MIXED = 3 };

template <int NP, int TG, int S>
constexpr int pad_region()
{
    constexpr int h = NP / 3, lo = TG * S, hi = TG * S + TG - 1;
    if (hi <= h) return PAD_LOW;
    if (lo >= NP - h) return PAD_HIGH;
    if (lo > h && hi < NP - h) return PAD_ZERO;
    return PAD_MIXED;
}

// Compact row of padded index j (kept |ky| <= h), clamped to row 0 for
// padding rows; *kept tells which.
template <int NP, int TG, int S>
__device__ __forceinline__ int compact_row_of(int j, bool *kept)
{
    constexpr int h = NP / 3, N = 2 * NP / 3;
    constexpr int R = pad_region<NP, TG, S>();
    if constexpr (R == PAD_LOW) {
        *kept = true;
        return j;
    } else if constexpr (R == PAD_HIGH) {
        *kept = true;
        return j - NP + N + 1;
    } else if constexpr (R == PAD_ZERO) {
        *kept = false;
        return 0;
    } else {
        const bool lo = j <= h, hi = j >= NP - h;
        *kept = lo || hi;
        return lo ? j : hi ? j - NP + N + 1 : 0;
    }
}

// Signed wavenumber of padded (FFTW-order) index j.
template <int NP, int TG, int S>
__device__ __forceinline__ int signed_k(int j)
{
    constexpr int R = pad_region<NP, TG, S>();
    if constexpr (R == PAD_LOW) return j;
    else if constexpr (R == PAD_HIGH) return j - NP;
    else return (j <= NP / 2) ? j : j - NP;
}

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
    const void *forcing; // compact-layout RHS forcing F (stride sW), or null:
                         // dW = -N - nu k^2 w - F in every stage
    // null: pass B's CFL maximum goes to DevScalars::vmax_bits with one
    // atomic per wave (the ensemble's runs spread those over their own
    // addresses). Else run r's VSLOTS slots at vslot + (r VSLOTS + i)
    // VSTRIDE: the blocks of a run spread their atomics over them (all on
    // one address they serialize: at 1024^2 FP64, 768 blocks' atomics made
    // stage 1's pass B 5 us slower), pass A of stage 1 clears them, and
    // pass C takes their maximum (writing it to vmax_bits at stage 3).
    unsigned long long *vslot;
};
constexpr int VSLOTS = 8, VSTRIDE = 16; // 128 bytes apart

// Run r's CFL maximum (double bits): DevScalars::vmax_bits, or the maximum
// of the run's slots.
__device__ __forceinline__ unsigned long long vmax_of(const DevScalars *sc,
                                                      const unsigned long long *vs)
{
    if (!vs) return sc->vmax_bits;
    unsigned long long m = vs[0];
#pragma unroll
    for (int i = 1; i < VSLOTS; ++i) m = vs[i * VSTRIDE] > m ? vs[i * VSTRIDE] : m;
    return m;
}
// The slots are used from Np = 1536 up (the sizes whose pass B runs in one
// round of blocks); the kernels of smaller sizes are compiled without them
// (their registers are tight) and the host passes none.
constexpr int VSLOT_MIN_NP = 1536;
template <int NP>
__device__ __forceinline__ unsigned long long *vslot_of(const Batch &bt, int run)
{
    if constexpr (NP < VSLOT_MIN_NP) return nullptr;
    else return bt.vslot ? bt.vslot + (size_t)run * VSLOTS * VSTRIDE : nullptr;
}

// The kernels are written as bodies over (block index blk of nblk blocks,
// the block's LDS); the __global__ wrappers below give them one size, the
// ensemble kernels (kernels_ensemble.hpp) select a size per block.

// ------------------------------------------------------------- pass A
// Jobs: (run, field, tile of LPB adjacent kx) or (run, tile, field); tiles
// past K hold inactive columns (zero input, nothing stored).
//
// The last column kx = K - 1 = N/2 holds a single kept mode, ky = 0 (the
// compact state is truncated to the disk |k| <= N/2), so its three fields
// are constant along y (the inverse transform of a lone ky = 0 mode is
// that mode). When LPB divides K - 1 the blocks of the last full tile write
// it as such, instead of a tile of its own that would transform one active
// column.
__host__ __device__ constexpr bool passA_tail(int K, int LPB)
{
    return K > 1 && (K - 1) % LPB == 0;
}

__host__ __device__ constexpr int passA_tiles(int K, int LPB)
{
    return passA_tail(K, LPB) ? (K - 1) / LPB : (K + LPB - 1) / LPB;
}

template <typename T, int NP, int LPB, int V = 0>
__device__ __forceinline__ void passA_body(int blk, int nblk, cpx<T> *lds,
                                           const cpx<T> *__restrict__ Wsrc,
                                           cpx<T> *__restrict__ Gu, cpx<T> *__restrict__ Gv,
                                           cpx<T> *__restrict__ Gw, const FusedDims &d,
                                           const cpx<T> *__restrict__ tw, DevScalars *sc,
                                           int reset_max, const Batch &bt)
{
#pragma clang fp contract(off)
    using C = WaveCfg<T, NP, V>;
    using Tl = TileCfg<T, NP, LPB, V>;
    constexpr int E = C::E, TG = C::TG, BLK = Tl::BLK, SL = Tl::SL;
    const int g = threadIdx.x / TG, t = threadIdx.x % TG;
    const int tiles = passA_tiles(d.K, LPB);
    // XCD-major blocks: the jobs in flight on one XCD are adjacent
    const int job = xcd_major(blk, nblk);
    if (job >= 3 * tiles * bt.nruns) return; // the whole block
    // Large transforms (multi-wave groups) order the jobs field fastest:
    // the three blocks that read one compact column then run together on
    // one XCD and two of them find it in L2 (measured: 8% less time at
    // Np = 1536). Small ones keep field-major order, whose adjacent blocks
    // fill the same lines of one G field (measured: 7% less time for the
    // mixed ensemble).
    constexpr bool FIELD_FAST = TG >= 128;
    const int run = job / (3 * tiles), rem = job - run * 3 * tiles;
    const int f = FIELD_FAST ? rem % 3 : rem / tiles;
    const int kx0 = (FIELD_FAST ? rem / 3 : rem - f * tiles) * LPB;
    const int kx = kx0 + g;
    const bool active = kx < d.K;
    const cpx<T> *src = Wsrc + run * bt.sW + (size_t)(active ? kx : 0) * d.KP;
    const T qx = (T)kx;
    // every load of the column first (unconditional, in bounds), then the
    // twiddles, so that all of them are in flight together
    cpx<T> x[E];
    bool kept[E];
    static_for<0, E>([&](auto sI) {
        constexpr int s = decltype(sI)::value;
        bool kp;
        const int c = compact_row_of<NP, TG, s>(t + TG * s, &kp);
        kept[s] = kp && active;
        if constexpr (pad_region<NP, TG, s>() != PAD_ZERO) x[s] = src[c];
    });
    GroupTw<T, E, TG, C::TW_PRE_A> gt;
    gt.load(tw, t);
    // the stage's scalar reset after the loads (vmcnt counts loads and
    // stores in order: stores ahead of the loads would delay their waits)
    if (reset_max && f == 0 && kx0 == 0 && threadIdx.x == 0) {
        sc[run].vmax_bits = 0ull;
        sc[run].t_start = sc[run].time;
        sc[run].bad = 0u;
        if (unsigned long long *vs = vslot_of<NP>(bt, run))
            for (int i = 0; i < VSLOTS; ++i) vs[i * VSTRIDE] = 0ull;
    }
    if (passA_tail(d.K, LPB) && kx0 + LPB == d.K - 1) { // block-uniform
        // column K - 1 of field f: its ky = 0 mode (compact row 0) taken
        // through the same spectrum arithmetic as a transformed column;
        // written here, ahead of the transform (its load then waits for no
        // store)
        const int kt = d.K - 1;
        const cpx<T> w0 = Wsrc[run * bt.sW + (size_t)kt * d.KP];
        const T qt = (T)kt;
        const T inv = recip(qt * qt);
        const cpx<T> ps = mk<T>(w0.re * inv, w0.im * inv);
        const T a = (f == 0) ? (T)0 : -qt;
        const cpx<T> val = f == 2 ? w0 : mk<T>(-(a * ps.im), a * ps.re);
        cpx<T> *col = (f == 0 ? Gu : f == 1 ? Gv : Gw) + run * bt.sG + kt;
        for (int y = threadIdx.x; y < NP; y += BLK) col[(size_t)y * d.KG] = val;
    }
    cpx<T> v[E];
    if (f == 2) { // block-uniform
        static_for<0, E>([&](auto sI) {
            constexpr int s = decltype(sI)::value;
            v[s] = mk<T>((T)0, (T)0);
            if constexpr (pad_region<NP, TG, s>() != PAD_ZERO) v[s] = sel(kept[s], x[s], v[s]);
        });
    } else {
        // u_hat = II*ky*psi (f = 0), v_hat = -II*kx*psi (f = 1): a*II*psi
        static_for<0, E>([&](auto sI) {
            constexpr int s = decltype(sI)::value;
            cpx<T> ps = mk<T>((T)0, (T)0);
            const int kyi = signed_k<NP, TG, s>(t + TG * s);
            const T qy = (T)kyi;
            if constexpr (pad_region<NP, TG, s>() != PAD_ZERO) {
                const cpx<T> w = sel(kept[s], x[s], ps);
                const T dd = (kx == 0) ? qy * qy : (kyi == 0) ? qx * qx : qx * qx + qy * qy;
                const T inv = recip(dd == (T)0 ? (T)1 : dd); // one reciprocal per mode
                ps = sel(kx == 0 && kyi == 0, ps, mk<T>(w.re * inv, w.im * inv));
            }
            const T a = (f == 0) ? qy : -qx;
            v[s] = mk<T>(-(a * ps.im), a * ps.re);
        });
    }
    cpx<T> *slice = lds + g * SL;
    group_fft<T, E, TG, 1>(v, slice, t, gt);
    // transpose through LDS: the block's columns become LPB-element row segments
    const int it = LdsPad<T>::idx(t);
#pragma unroll
    for (int s = 0; s < E; ++s) slice[LdsPad<T>::idx_sum(it, TG * s)] = v[s];
    __syncthreads();
    // LPB * NP = E * BLK elements: element e = threadIdx.x + BLK m is row
    // e / LPB, column e % LPB of the tile (BLK = LPB * TG)
    const int gi = threadIdx.x % LPB, yb = threadIdx.x / LPB;
    if (kx0 + gi < d.K) {
        cpx<T> *col = (f == 0 ? Gu : f == 1 ? Gv : Gw) + run * bt.sG + kx0 + gi;
        const cpx<T> *sl = lds + gi * SL;
        // (the full index here: with idx_sum the mixed ensemble's pass A took
        // 12 % longer)
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int y = yb + TG * m;
            col[(size_t)y * d.KG] = sl[LdsPad<T>::idx(y)];
        }
    }
}

// Pass A with the three fields of a column tile in one block (VAR_FUSE_A):
// the column's loads, kept flags and psi = w / |k|^2 are formed once, then
// u_hat, v_hat and w_hat are transformed and stored in turn through the same
// LDS tile (a barrier between the fields). Per field the arithmetic is
// passA_body's, so the G fields are bitwise the same.
template <typename T, int NP, int LPB, int V>
__device__ __forceinline__ void passA_fused_body(int blk, int nblk, cpx<T> *lds,
                                                 const cpx<T> *__restrict__ Wsrc,
                                                 cpx<T> *__restrict__ Gu, cpx<T> *__restrict__ Gv,
                                                 cpx<T> *__restrict__ Gw, const FusedDims &d,
                                                 const cpx<T> *__restrict__ tw, DevScalars *sc,
                                                 int reset_max, const Batch &bt)
{
#pragma clang fp contract(off)
    using C = WaveCfg<T, NP, V>;
    using Tl = TileCfg<T, NP, LPB, V>;
    constexpr int E = C::E, TG = C::TG, BLK = Tl::BLK, SL = Tl::SL;
    const int g = threadIdx.x / TG, t = threadIdx.x % TG;
    const int tiles = passA_tiles(d.K, LPB);
    const int job = xcd_major(blk, nblk);
    if (job >= tiles * bt.nruns) return; // the whole block
    const int run = job / tiles, kx0 = (job - run * tiles) * LPB;
    const int kx = kx0 + g;
    const bool active = kx < d.K;
    const cpx<T> *src = Wsrc + run * bt.sW + (size_t)(active ? kx : 0) * d.KP;
    const T qx = (T)kx;
    cpx<T> x[E];
    bool kept[E];
    static_for<0, E>([&](auto sI) {
        constexpr int s = decltype(sI)::value;
        bool kp;
        const int c = compact_row_of<NP, TG, s>(t + TG * s, &kp);
        kept[s] = kp && active;
        if constexpr (pad_region<NP, TG, s>() != PAD_ZERO) x[s] = src[c];
    });
    GroupTw<T, E, TG, C::TW_PRE_A> gt;
    gt.load(tw, t);
    if (reset_max && kx0 == 0 && threadIdx.x == 0) { // (after the loads, as passA_body)
        sc[run].vmax_bits = 0ull;
        sc[run].t_start = sc[run].time;
        sc[run].bad = 0u;
        if (unsigned long long *vs = vslot_of<NP>(bt, run))
            for (int i = 0; i < VSLOTS; ++i) vs[i * VSTRIDE] = 0ull;
    }
    cpx<T> ps[E], v[E];
    static_for<0, E>([&](auto sI) {
        constexpr int s = decltype(sI)::value;
        ps[s] = mk<T>((T)0, (T)0);
        v[s] = mk<T>((T)0, (T)0);
        const int kyi = signed_k<NP, TG, s>(t + TG * s);
        const T qy = (T)kyi;
        if constexpr (pad_region<NP, TG, s>() != PAD_ZERO) {
            const cpx<T> w = sel(kept[s], x[s], ps[s]);
            const T dd = (kx == 0) ? qy * qy : (kyi == 0) ? qx * qx : qx * qx + qy * qy;
            const T inv = recip(dd == (T)0 ? (T)1 : dd); // one reciprocal per mode
            ps[s] = sel(kx == 0 && kyi == 0, ps[s], mk<T>(w.re * inv, w.im * inv));
            v[s] = sel(kept[s], x[s], v[s]); // w_hat (x and kept end here)
        }
    });
    cpx<T> *slice = lds + g * SL;
    const int it = LdsPad<T>::idx(t);
    const int gi = threadIdx.x % LPB, yb = threadIdx.x / LPB;
    // fields in the order w, u, v: w_hat first frees the column's registers
#pragma unroll 1
    for (int i = 0; i < 3; ++i) {
        const int f = i == 0 ? 2 : i - 1;
        if (i > 0) {
            static_for<0, E>([&](auto sI) {
                constexpr int s = decltype(sI)::value;
                const T qy = (T)signed_k<NP, TG, s>(t + TG * s);
                const T a = (f == 0) ? qy : -qx;
                v[s] = mk<T>(-(a * ps[s].im), a * ps[s].re);
            });
            __syncthreads(); // the previous field's tile has been read
        }
        group_fft<T, E, TG, 1>(v, slice, t, gt);
#pragma unroll
        for (int s = 0; s < E; ++s) slice[LdsPad<T>::idx_sum(it, TG * s)] = v[s];
        __syncthreads();
        if (kx0 + gi < d.K) {
            cpx<T> *col = (f == 0 ? Gu : f == 1 ? Gv : Gw) + run * bt.sG + kx0 + gi;
            const cpx<T> *sl = lds + gi * SL;
#pragma unroll
            for (int m = 0; m < E; ++m) {
                const int y = yb + TG * m;
                col[(size_t)y * d.KG] = sl[LdsPad<T>::idx(y)];
            }
        }
        if (passA_tail(d.K, LPB) && kx0 + LPB == d.K - 1) { // block-uniform
            const int kt = d.K - 1;
            const cpx<T> w0 = Wsrc[run * bt.sW + (size_t)kt * d.KP];
            const T qt = (T)kt;
            const T inv = recip(qt * qt);
            const cpx<T> pt = mk<T>(w0.re * inv, w0.im * inv);
            const T a = (f == 0) ? (T)0 : -qt;
            const cpx<T> val = f == 2 ? w0 : mk<T>(-(a * pt.im), a * pt.re);
            cpx<T> *col = (f == 0 ? Gu : f == 1 ? Gv : Gw) + run * bt.sG + kt;
            for (int y = threadIdx.x; y < NP; y += BLK) col[(size_t)y * d.KG] = val;
        }
    }
}

// Blocks of pass A for nruns runs (V: the schedule variant; VAR_FUSE_A takes
// the three fields of a tile in one block).
template <typename T, int NP, int LPB, int V = 0>
__host__ __device__ constexpr int passA_blocks(int K, int nruns)
{
    return ((V & VAR_FUSE_A) ? 1 : 3) * passA_tiles(K, LPB) * nruns;
}

template <typename T, int NP, int V = 0>
__global__ __launch_bounds__((TileCfg<T, NP, WaveCfg<T, NP, V>::LPB_A, V>::BLK))
__attribute__((amdgpu_waves_per_eu(WaveCfg<T, NP, V>::WPE_A))) void
k2_passA(const cpx<T> *__restrict__ Wsrc, cpx<T> *__restrict__ Gu, cpx<T> *__restrict__ Gv,
         cpx<T> *__restrict__ Gw, FusedDims d, const cpx<T> *__restrict__ tw, DevScalars *sc,
         int reset_max, Batch bt)
{
    constexpr int LPB = WaveCfg<T, NP, V>::LPB_A;
    __shared__ cpx<T> lds[TileCfg<T, NP, LPB, V>::LDS];
    if constexpr ((V & VAR_FUSE_A) != 0)
        passA_fused_body<T, NP, LPB, V>(blockIdx.x, gridDim.x, lds, Wsrc, Gu, Gv, Gw, d, tw, sc,
                                        reset_max, bt);
    else
        passA_body<T, NP, LPB, V>(blockIdx.x, gridDim.x, lds, Wsrc, Gu, Gv, Gw, d, tw, sc,
                                  reset_max, bt);
}

// ------------------------------------------------------------- pass B

// Hermitian extension of one G row at frequency k = t + TG*S of a
// length-NP transform: G[k] for k <= h, conj G[NP - k] for k >= NP - h,
// zero between (only Re of k = 0 is read).
template <typename T, int NP, int TG, int S>
__device__ __forceinline__ cpx<T> herm_load(const cpx<T> *__restrict__ row, int t)
{
    constexpr int h = NP / 3;
    constexpr int R = pad_region<NP, TG, S>();
    const int k = t + TG * S;
    if constexpr (R == PAD_ZERO) {
        return mk<T>((T)0, (T)0);
    } else if constexpr (R == PAD_LOW) {
        cpx<T> a = row[k];
        if constexpr (S == 0) a.im = (t == 0) ? (T)0 : a.im;
        return a;
    } else if constexpr (R == PAD_HIGH) {
        cpx<T> a = row[NP - k];
        a.im = -a.im;
        return a;
    } else {
        const bool lo = k <= h, hi = k >= NP - h;
        const cpx<T> x = row[lo ? k : hi ? NP - k : 0]; // in bounds for every lane
        const T re = (lo || hi) ? x.re : (T)0;
        const T im = lo ? x.im : hi ? -x.im : (T)0;
        return mk<T>(re, k == 0 ? (T)0 : im);
    }
}

// Jobs: (run, tile of LPB adjacent rows). Two bodies: passB_single (a lane
// group per row, LPB * TG lanes) and passB_pair (a lane group per two rows,
// LPB * TG / 2 lanes, five transforms per two rows instead of six). The
// single-row tiles keep more waves in flight where blocks are many and
// short (the mixed ensemble); the pairs win where the transform is long
// (Np >= 1536).
template <typename T, int NP, int LPB, int V = 0>
__device__ __forceinline__ void passB_single(int blk, int nblk, cpx<T> *lds,
                                           const cpx<T> *__restrict__ Gu,
                                           const cpx<T> *__restrict__ Gv,
                                           const cpx<T> *__restrict__ Gw,
                                           cpx<T> *__restrict__ Huw, cpx<T> *__restrict__ Hvw,
                                           const FusedDims &d, const cpx<T> *__restrict__ tw,
                                           DevScalars *sc, int forward, int track_max,
                                           const Batch &bt)
{
#pragma clang fp contract(off)
    using C = WaveCfg<T, NP, V>;
    using Tl = TileCfg<T, NP, LPB, V>;
    constexpr int E = C::E, TG = C::TG, BLK = Tl::BLK, SL = Tl::SL; // BLK lanes
    constexpr int tiles = NP / LPB;
    static_assert(NP % LPB == 0, "row tiles must divide the rows");
    const int g = threadIdx.x / TG, t = threadIdx.x % TG;
    const int job = xcd_major(blk, nblk);
    if (job >= tiles * bt.nruns) return; // the whole block
    const int run = job / tiles, y0 = (job - run * tiles) * LPB, y = y0 + g;
    cpx<T> *slice = lds + g * SL;
    Gu += run * bt.sG;
    Gv += run * bt.sG;
    Gw += run * bt.sG;

    // w of this row (the imaginary part of the result is the rounding of a
    // real transform and is dropped)
    const cpx<T> *rw = Gw + (size_t)y * d.KG, *ru = Gu + (size_t)y * d.KG,
                 *rv = Gv + (size_t)y * d.KG;
    // u + i v of this row (FP32 has the registers to issue its loads ahead
    // of the w transform: measured 7 % less time for the mixed ensemble)
    auto load_uv = [&](cpx<T>(&z)[E]) {
        static_for<0, E>([&](auto sI) {
            constexpr int s = decltype(sI)::value;
            const cpx<T> a = herm_load<T, NP, TG, s>(ru, t), b = herm_load<T, NP, TG, s>(rv, t);
            z[s] = mk<T>(a.re - b.im, a.im + b.re);
        });
    };
    cpx<T> v[E], z[E];
    static_for<0, E>([&](auto sI) {
        constexpr int s = decltype(sI)::value;
        v[s] = herm_load<T, NP, TG, s>(rw, t);
    });
    if constexpr (sizeof(T) == 4) load_uv(z);
    GroupTw<T, E, TG, C::TW_PRE_B1> gt; // the row's three transforms share the lane's twiddles
    gt.load(tw, t);
    group_fft<T, E, TG, 1>(v, slice, t, gt);
    T wr[E];
#pragma unroll
    for (int s = 0; s < E; ++s) wr[s] = v[s].re;
    if constexpr (sizeof(T) == 4) {
#pragma unroll
        for (int s = 0; s < E; ++s) v[s] = z[s];
    } else {
        load_uv(v);
    }
    group_fft<T, E, TG, 1>(v, slice, t, gt);
    T vmax = (T)0;
    if (track_max) { // uniform
#pragma unroll
        for (int s = 0; s < E; ++s) {
            const T m = absv(v[s].re) + absv(v[s].im);
            vmax = (m > vmax || m != m) ? m : vmax; // NaN sticks
        }
    }
#pragma unroll
    for (int s = 0; s < E; ++s) v[s] = mk<T>(v[s].re * wr[s], v[s].im * wr[s]); // uw + i vw
    if (forward) {
        group_fft<T, E, TG, -1>(v, slice, t, gt);
        const int it = LdsPad<T>::idx(t);
#pragma unroll
        for (int s = 0; s < E; ++s) slice[LdsPad<T>::idx_sum(it, TG * s)] = v[s];
        __syncthreads();
        // k <= N/2 = NP/3 of the block's LPB rows, rows fastest: LPB-element
        // segments of the H columns
        constexpr int H = NP / 3, CNT = (H + 1) * LPB;
        Huw += run * bt.sH + y0;
        Hvw += run * bt.sH + y0;
#pragma unroll 1
        for (int m = 0; m < (CNT + BLK - 1) / BLK; ++m) {
            const int e = threadIdx.x + BLK * m;
            if (e < CNT) {
                const int k = e / LPB, r = e % LPB;
                const cpx<T> *sl = lds + r * SL;
                const cpx<T> P = sl[LdsPad<T>::idx(k)];
                cpx<T> Pc = sl[LdsPad<T>::idx((NP - k) % NP)];
                Pc.im = -Pc.im;
                const cpx<T> su = add(P, Pc), df = sub(P, Pc);
                Huw[(size_t)k * NP + r] = mk<T>((T)0.5 * su.re, (T)0.5 * su.im);
                Hvw[(size_t)k * NP + r] = mk<T>((T)0.5 * df.im, (T)-0.5 * df.re);
            }
        }
    }
    if (!track_max) return;
    double dm = (double)vmax;
    if (dm != dm) dm = __longlong_as_double(0x7ff8000000000000ll);
    for (int off = 32; off > 0; off >>= 1) {
        double o = __shfl_xor(dm, off);
        dm = (o > dm || o != o) ? o : dm;
    }
    // every row of a block belongs to the same run (LPB divides NP)
    if ((threadIdx.x & 63) == 0) {
        unsigned long long *vs = vslot_of<NP>(bt, run);
        atomicMax(vs ? vs + (blk % VSLOTS) * VSTRIDE : &sc[run].vmax_bits,
                  (unsigned long long)__double_as_longlong(dm));
    }
}

template <typename T, int NP, int LPB, int V = 0>
__device__ __forceinline__ void passB_pair(int blk, int nblk, cpx<T> *lds,
                                           const cpx<T> *__restrict__ Gu,
                                           const cpx<T> *__restrict__ Gv,
                                           const cpx<T> *__restrict__ Gw,
                                           cpx<T> *__restrict__ Huw, cpx<T> *__restrict__ Hvw,
                                           const FusedDims &d, const cpx<T> *__restrict__ tw,
                                           DevScalars *sc, int forward, int track_max,
                                           const Batch &bt)
{
#pragma clang fp contract(off)
    using C = WaveCfg<T, NP, V>;
    using Tl = TileCfg<T, NP, LPB, V>;
    constexpr int E = C::E, TG = C::TG, THR = Tl::BLK / 2, SL = Tl::SL;
    constexpr int tiles = NP / LPB;
    static_assert(NP % LPB == 0 && LPB % 2 == 0, "row tiles: even, dividing the rows");
    const int g = threadIdx.x / TG, t = threadIdx.x % TG;
    const int job = xcd_major(blk, nblk);
    if (job >= tiles * bt.nruns) return; // the whole block
    // lane group g takes the rows ya = y0 + 2g and ya + 1
    const int run = job / tiles, y0 = (job - run * tiles) * LPB, ya = y0 + 2 * g;
    cpx<T> *sla = lds + (2 * g) * SL, *slb = sla + SL;
    Gu += run * bt.sG;
    Gv += run * bt.sG;
    Gw += run * bt.sG;
    const size_t oa = (size_t)ya * d.KG, ob = oa + d.KG;

    // u + i v of a row from its two Hermitian spectra (two-for-one)
    auto load_uv = [&](cpx<T>(&z)[E], size_t o) {
        static_for<0, E>([&](auto sI) {
            constexpr int s = decltype(sI)::value;
            const cpx<T> a = herm_load<T, NP, TG, s>(Gu + o, t);
            const cpx<T> b = herm_load<T, NP, TG, s>(Gv + o, t);
            z[s] = mk<T>(a.re - b.im, a.im + b.re);
        });
    };
    // w of both rows from one transform: w_a + i w_b
    cpx<T> v[E], z[E];
    static_for<0, E>([&](auto sI) {
        constexpr int s = decltype(sI)::value;
        const cpx<T> a = herm_load<T, NP, TG, s>(Gw + oa, t);
        const cpx<T> b = herm_load<T, NP, TG, s>(Gw + ob, t);
        v[s] = mk<T>(a.re - b.im, a.im + b.re);
    });
    // FP32 has the registers to issue row a's u, v loads ahead of the w
    // transform
    if constexpr (sizeof(T) == 4) load_uv(z, oa);
    GroupTw<T, E, TG, C::TW_PRE_B> gt; // the lane's twiddles, shared by the five transforms
    gt.load(tw, t);
    group_fft<T, E, TG, 1>(v, sla, t, gt);
    // w_a stays in registers; w_b waits in row b's slice (unused until row
    // b), each lane reading back only what it wrote
    T wa[E];
    T *wb_lds = reinterpret_cast<T *>(slb);
#pragma unroll
    for (int s = 0; s < E; ++s) {
        wa[s] = v[s].re;
        wb_lds[t + TG * s] = v[s].im;
    }
    T vmax = (T)0;
    // one row: u + i v, the products (u w + i v w), their forward transform
    // into the row's LDS slice
    auto row = [&](cpx<T>(&x)[E], const T(&w)[E], cpx<T> *sl) {
        group_fft<T, E, TG, 1>(x, sl, t, gt);
        if (track_max) { // uniform
#pragma unroll
            for (int s = 0; s < E; ++s) {
                const T m = absv(x[s].re) + absv(x[s].im);
                vmax = (m > vmax || m != m) ? m : vmax; // NaN sticks
            }
        }
        if (forward) {
#pragma unroll
            for (int s = 0; s < E; ++s) x[s] = mk<T>(x[s].re * w[s], x[s].im * w[s]);
            group_fft<T, E, TG, -1>(x, sl, t, gt);
            const int it = LdsPad<T>::idx(t);
#pragma unroll
            for (int s = 0; s < E; ++s) sl[LdsPad<T>::idx_sum(it, TG * s)] = x[s];
        }
    };
    if constexpr (sizeof(T) != 4) load_uv(z, oa);
    row(z, wa, sla);
    load_uv(z, ob);
    T wb[E];
#pragma unroll
    for (int s = 0; s < E; ++s) wb[s] = wb_lds[t + TG * s];
    // row b's transform writes anywhere in the slice: in a multi-wave group
    // every wave must have read its w_b first
    group_sync<TG>();
    row(z, wb, slb);
    if (forward) {
        __syncthreads();
        // k <= N/2 = NP/3 of the block's LPB rows, rows fastest: LPB-element
        // segments of the H columns
        constexpr int H = NP / 3, CNT = (H + 1) * LPB;
        Huw += run * bt.sH + y0;
        Hvw += run * bt.sH + y0;
#pragma unroll // (unrolled: the LDS reads and stores of several
                // iterations in flight, 1 % less time at Np = 1536)
        for (int m = 0; m < (CNT + THR - 1) / THR; ++m) {
            const int e = threadIdx.x + THR * m;
            if (e < CNT) {
                const int k = e / LPB, r = e % LPB;
                const cpx<T> *sl = lds + r * SL;
                const cpx<T> P = sl[LdsPad<T>::idx(k)];
                cpx<T> Pc = sl[LdsPad<T>::idx((NP - k) % NP)];
                Pc.im = -Pc.im;
                const cpx<T> su = add(P, Pc), df = sub(P, Pc);
                Huw[(size_t)k * NP + r] = mk<T>((T)0.5 * su.re, (T)0.5 * su.im);
                Hvw[(size_t)k * NP + r] = mk<T>((T)0.5 * df.im, (T)-0.5 * df.re);
            }
        }
    }
    if (!track_max) return;
    double dm = (double)vmax;
    if (dm != dm) dm = __longlong_as_double(0x7ff8000000000000ll);
    for (int off = 32; off > 0; off >>= 1) {
        double o = __shfl_xor(dm, off);
        dm = (o > dm || o != o) ? o : dm;
    }
    // every row of a block belongs to the same run (LPB divides NP). The
    // block's waves combine their maxima through LDS, so that each block
    // issues one atomic, to one of the run's slots (Batch::vslot): atomics
    // on one address are serialized there (one per wave on one address:
    // stage 1's pass B, the one that tracks the maximum, took 38 us against
    // 26 for the other stages at 1024^2 FP64; one per block: 31 us)
    constexpr int NWB = THR / 64;
    unsigned long long *vs = vslot_of<NP>(bt, run);
    unsigned long long *vdst = vs ? vs + (blk % VSLOTS) * VSTRIDE : &sc[run].vmax_bits;
    if constexpr (NWB > 1) {
        __syncthreads(); // the slices' last reads are done
        double *red = reinterpret_cast<double *>(lds);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = dm;
        __syncthreads();
        if (threadIdx.x == 0) {
#pragma unroll
            for (int w = 1; w < NWB; ++w) {
                const double o = red[w];
                dm = (o > dm || o != o) ? o : dm;
            }
            atomicMax(vdst, (unsigned long long)__double_as_longlong(dm));
        }
    } else if ((threadIdx.x & 63) == 0) {
        atomicMax(vdst, (unsigned long long)__double_as_longlong(dm));
    }
}

template <typename T, int NP, int LPB>
__host__ __device__ constexpr int passB_blocks(int nruns)
{
    return NP / LPB * nruns;
}

template <typename T, int NP, int V = 0>
__global__ __launch_bounds__((WaveCfg<T, NP, V>::THR_B))
__attribute__((amdgpu_waves_per_eu(WaveCfg<T, NP, V>::WPE_B))) void
k2_passB(const cpx<T> *__restrict__ Gu, const cpx<T> *__restrict__ Gv,
         const cpx<T> *__restrict__ Gw, cpx<T> *__restrict__ Huw, cpx<T> *__restrict__ Hvw,
         FusedDims d, const cpx<T> *__restrict__ tw, DevScalars *sc, int forward, int track_max,
         Batch bt)
{
    constexpr int LPB = WaveCfg<T, NP, V>::LPB_B;
    __shared__ cpx<T> lds[TileCfg<T, NP, LPB, V>::LDS];
    if constexpr (WaveCfg<T, NP, V>::PAIR_B)
        passB_pair<T, NP, LPB, V>(blockIdx.x, gridDim.x, lds, Gu, Gv, Gw, Huw, Hvw, d, tw, sc,
                               forward, track_max, bt);
    else
        passB_single<T, NP, LPB, V>(blockIdx.x, gridDim.x, lds, Gu, Gv, Gw, Huw, Hvw, d, tw, sc,
                                 forward, track_max, bt);
}

// ------------------------------------------------------------- pass C

// The compact-state operands of half HALF's rows (s with s % 2 == HALF) of
// one column: the stage and step-start values and the forcing, loaded
// before the column's transform so that their latency overlaps it.
template <typename T, int NP, int HALF, int V = 0>
struct PassCOperands {
    static constexpr int E = WaveCfg<T, NP, V>::E, TG = WaveCfg<T, NP, V>::TG;
    cpx<T> ws[E / 2], w0[E / 2], fv[E / 2];
    int c[E / 2];
    bool row[E / 2];

    __device__ __forceinline__ void load(int t, const cpx<T> *Wn, const cpx<T> *Wst,
                                         const cpx<T> *Fc, size_t col)
    {
        static_for<0, E / 2>([&](auto iI) {
            constexpr int i = decltype(iI)::value, s = 2 * i + HALF;
            if constexpr (pad_region<NP, TG, s>() != PAD_ZERO) {
                c[i] = compact_row_of<NP, TG, s>(t + TG * s, &row[i]);
                ws[i] = Wst[col + c[i]];
                w0[i] = Wn[col + c[i]];
            }
        });
        if (Fc) { // uniform
            static_for<0, E / 2>([&](auto iI) {
                constexpr int i = decltype(iI)::value, s = 2 * i + HALF;
                if constexpr (pad_region<NP, TG, s>() != PAD_ZERO) fv[i] = Fc[col + c[i]];
            });
        }
    }
};

// The Shu-Osher stage update of one mode (src/TimeIntegration.F90:156-202)
// from the step-start value w0, the stage value ws and dW.
template <typename T>
__device__ __forceinline__ cpx<T> rk3_update(int stage, T dt, cpx<T> w0v, cpx<T> wsv, cpx<T> dW)
{
#pragma clang fp contract(off)
    if (stage == 1)
        return mk<T>(w0v.re + dt * dW.re, w0v.im + dt * dW.im);
    if (stage == 2) {
        const T cc = (T)0.25 * dt;
        return mk<T>(((T)0.75 * w0v.re + (T)0.25 * wsv.re) + cc * dW.re,
                     ((T)0.75 * w0v.im + (T)0.25 * wsv.im) + cc * dW.im);
    }
    const T cc = dt * (T)2;
    // (w + 2 w2 + 2 dt dW)/3 correctly rounded: a rounded 1/3 multiplier
    // alone would bias |w| by its rounding error every step
    return mk<T>(div3(((w0v.re + (T)2 * wsv.re) + cc * dW.re)),
                 div3(((w0v.im + (T)2 * wsv.im) + cc * dW.im)));
}

// The RHS, truncation and stage update of half HALF's rows of column kx;
// P / Q are the uw / vw column spectra (one of them in registers, the other
// read from the partner half's LDS slice).
template <typename T, int NP, int HALF, int V = 0>
__device__ __forceinline__ void passC_finish(const cpx<T> (&v)[WaveCfg<T, NP, V>::E],
                                             const cpx<T> *other, int t, int kx, bool active,
                                             const PassCOperands<T, NP, HALF, V> &op, bool forced,
                                             cpx<T> *Wout, size_t col, const StageArgs &a,
                                             double dtd, double nu)
{
#pragma clang fp contract(off)
    using C = WaveCfg<T, NP, V>;
    constexpr int E = C::E, TG = C::TG, H = NP / 3;
    const T dt = (T)dtd;
    const T qx = (T)kx;
    const T inv_norm = (T)a.inv_norm; // multiplications instead of the
                                      // reference's divisions: <= 1 ulp
    const int it = LdsPad<T>::idx(t);
    static_for<0, E / 2>([&](auto iI) {
        constexpr int i = decltype(iI)::value, s = 2 * i + HALF;
        if constexpr (pad_region<NP, TG, s>() != PAD_ZERO) {
            const int j = t + TG * s;
            const int kyi = signed_k<NP, TG, s>(j);
            const T qy = (T)kyi;
            // SetupTruncation's |k| <= kMax + 1e-10 (src/Spectral.F90:272-313)
            // is kx^2 + ky^2 <= (N/2)^2 on integer wavenumbers
            const bool keep = kx * kx + kyi * kyi <= H * H;
            const cpx<T> o = other[LdsPad<T>::idx_sum(it, TG * s)]; // j = t + TG s
            const cpx<T> Pr = HALF == 0 ? v[s] : o, Qr = HALF == 0 ? o : v[s];
            const cpx<T> P = mk<T>(Pr.re * inv_norm, Pr.im * inv_norm);
            const cpx<T> Q = mk<T>(Qr.re * inv_norm, Qr.im * inv_norm);
            const cpx<T> du = mk<T>(-(qx * P.im), qx * P.re); // DuDx of (uw)^
            const cpx<T> dv = mk<T>(-(qy * Q.im), qy * Q.re); // DuDy of (vw)^
            const T sn = (T)nu * (qx * qx + qy * qy);
            const cpx<T> wsv = op.ws[i], w0v = op.w0[i];
            cpx<T> dW = mk<T>((-du.re - dv.re) - sn * wsv.re, (-du.im - dv.im) - sn * wsv.im);
            if (forced) dW = mk<T>(dW.re - op.fv[i].re, dW.im - op.fv[i].im); // closure / forcing
            cpx<T> out = rk3_update<T>(a.stage, dt, w0v, wsv, dW);
            out = sel(keep, out, mk<T>((T)0, (T)0));
            if (active && op.row[i]) Wout[col + op.c[i]] = out;
        }
    });
}

// Column K - 1 written by passC_tail (one sum instead of a transformed
// column) for the sizes above the ensemble's (Np >= 1536): at Np = 1536 in
// FP64 two blocks fit a CU, and the 513th column ran as a second round
// (measured: pass C 13 % less time in FP64, 6 % in FP32); in the
// ensemble, whose sizes keep the plain columns, the per-run tails made
// pass C 8 % slower.
template <int NP>
constexpr bool passC_tail_on()
{
    return NP >= 1536;
}

// One half of a pass C block: half 0 transforms uw, half 1 vw.
template <typename T, int NP, int HALF, bool TAIL, int V = 0>
__device__ __forceinline__ void passC_half(int blk, cpx<T> *lds_all, int lane,
                                           const cpx<T> *__restrict__ Hsrc, const cpx<T> *Wn,
                                           const cpx<T> *Wst, cpx<T> *Wout, const FusedDims &d,
                                           const cpx<T> *__restrict__ tw, DevScalars *sc,
                                           const StageArgs &a, const Batch &bt,
                                           double &dt_out, double &nu_out)
{
    using C = WaveCfg<T, NP, V>;
    constexpr int E = C::E, TG = C::TG;
    const int g = lane / TG, t = lane % TG;
    cpx<T> *mine = lds_all + (HALF * C::GPB + g) * C::SLICE;
    const cpx<T> *other = lds_all + ((1 - HALF) * C::GPB + g) * C::SLICE;
    const int job = blk * C::GPB + g;
    const int Kt = TAIL ? d.K - 1 : d.K; // transformed columns
    const bool active = job < Kt * bt.nruns;
    const int run = active ? job / Kt : 0, kx = active ? job - run * Kt : 0;
    const cpx<T> *src = Hsrc + run * bt.sH + (size_t)kx * NP;
    cpx<T> v[E];
#pragma unroll
    for (int s = 0; s < E; ++s) v[s] = src[t + TG * s];
    const size_t col = run * bt.sW + (size_t)kx * d.KP;
    const cpx<T> *Fc = static_cast<const cpx<T> *>(bt.forcing);
    // the twiddles, then the operands the column's update needs after the
    // transform (vmcnt counts loads in issue order: the transform's wait for
    // its twiddles does not include the operands)
    GroupTw<T, E, TG, C::TW_PRE_C> gt;
    gt.load(tw, t);
    __builtin_amdgcn_sched_barrier(0); // (keeps that order)
    PassCOperands<T, NP, HALF, V> op;
    if constexpr (C::OPS_EARLY_C) op.load(t, Wn, Wst, Fc, col);
    const unsigned long long *vs = vslot_of<NP>(bt, run);
    unsigned long long vbits = 0ull;
    double dtd;
    if constexpr (NP >= VSLOT_MIN_NP) {
        vbits = vmax_of(sc + run, vs);
        dtd = stage_dt(a, sc + run, vbits);
    } else {
        dtd = stage_dt(a, sc + run);
    }
    const double nu = bt.nu ? bt.nu[run] : a.nu;
    dt_out = dtd;
    nu_out = nu;
    // the run's step-end scalars: one counter, vmcnt, tracks loads and
    // stores in order, so loads issued after the column's stores wait for
    // all of them. The large transforms (the sizes that run in one round,
    // with registers to spare) load them before the stores; the others
    // (five waves per SIMD in FP32) after.
    constexpr bool SC_EARLY = TG >= 128;
    const bool upd = HALF == 0 && a.stage == 3 && active && kx == 0 && t == 0;
    double t_old = 0.0, t_end = 0.0;
    long long n_old = 0;
    if (SC_EARLY && upd) {
        t_old = sc[run].time;
        t_end = sc[run].t_end;
        n_old = sc[run].nadv;
    }
    group_fft<T, E, TG, -1>(v, mine, t, gt);
    if constexpr (!C::OPS_EARLY_C) op.load(t, Wn, Wst, Fc, col);
    // half 0 keeps the even s and hands over the odd s, half 1 the reverse
    const int it = LdsPad<T>::idx(t);
#pragma unroll
    for (int s = 0; s < E; ++s)
        if ((s & 1) != HALF) mine[LdsPad<T>::idx_sum(it, TG * s)] = v[s];
    __syncthreads();
    // a run at its end time (dt = 0) stores nothing: its step-start state
    // stays in Wn (and Ws, which only this run's later stages read, is unused)
    passC_finish<T, NP, HALF, V>(v, other, t, kx, active && dtd != 0.0, op, Fc != nullptr, Wout,
                                 col, a, dtd, nu);
    if (upd) {
        if constexpr (!SC_EARLY) {
            t_old = sc[run].time;
            t_end = sc[run].t_end;
            n_old = sc[run].nadv;
        }
        sc[run].dt = dtd;
        sc[run].time = t_old + dtd;
        if (!(dtd == 0.0 && t_end > 0.0)) sc[run].nadv = n_old + 1;
        if constexpr (NP >= VSLOT_MIN_NP)
            if (vs) sc[run].vmax_bits = vbits; // the step's maximum, for the host
    }
}

// Column kx = K - 1 = N/2 of run `run`: its one kept mode is ky = 0 (the
// compact state is truncated to the disk |k| <= N/2), whose transform
// term is the plain sum over y of the H_uw column (the vw term enters as
// i ky = 0). The sum runs in a fixed order (strided partial sums, then the
// block's reduction), the RHS and stage update as passC_finish's, and the
// column's other rows are written as zeros. Every thread of the block
// takes part.
template <typename T, int NP, int V = 0>
__device__ __forceinline__ void passC_tail(int run, cpx<T> *lds_all, const cpx<T> *Huw,
                                           const cpx<T> *Wn, const cpx<T> *Wst, cpx<T> *Wout,
                                           const FusedDims &d, const StageArgs &a,
                                           const Batch &bt, double dtd, double nu_run)
{
#pragma clang fp contract(off)
    constexpr int NT = 2 * WaveCfg<T, NP, V>::BLK, NW = NT / 64;
    static_assert(WaveCfg<T, NP, V>::GPB == 1, "tail: one column per half (its run's column 0)");
    const int kx = d.K - 1;
    const cpx<T> *h = Huw + run * bt.sH + (size_t)kx * NP;
    const size_t col = run * bt.sW + (size_t)kx * d.KP;
    const cpx<T> *Fc = static_cast<const cpx<T> *>(bt.forcing);
    // every load of the tail issued together, after the column's stores,
    // which their one wait includes (vmcnt counts both); the run's dt and
    // viscosity come from the column's own pass (thread 0: half 0, the
    // run's column 0)
    cpx<T> wsv = mk<T>((T)0, (T)0), w0v = wsv, fcv = wsv;
    if (threadIdx.x == 0) {
        wsv = Wst[col];
        w0v = Wn[col];
        if (Fc) fcv = Fc[col];
    }
    const T nu = (T)nu_run;
    static_assert(NP % NT == 0, "tail: whole strides");
    cpx<T> hv[NP / NT];
#pragma unroll
    for (int k = 0; k < NP / NT; ++k) hv[k] = h[threadIdx.x + NT * k];
    cpx<T> p = mk<T>((T)0, (T)0);
#pragma unroll
    for (int k = 0; k < NP / NT; ++k) p = add(p, hv[k]); // y = threadIdx.x + NT k, in order
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        p = mk<T>(p.re + __shfl_xor(p.re, off), p.im + __shfl_xor(p.im, off));
    __syncthreads(); // the slices' last reads (passC_finish) are done
    if ((threadIdx.x & 63) == 0) lds_all[threadIdx.x / 64] = p;
    __syncthreads();
    if (threadIdx.x == 0) {
        cpx<T> P = lds_all[0];
#pragma unroll
        for (int w = 1; w < NW; ++w) P = add(P, lds_all[w]);
        const T inv_norm = (T)a.inv_norm;
        P = mk<T>(P.re * inv_norm, P.im * inv_norm);
        const T qx = (T)kx;
        const cpx<T> du = mk<T>(-(qx * P.im), qx * P.re); // DuDx of (uw)^
        const T sn = nu * (qx * qx);
        cpx<T> dW = mk<T>((-du.re) - sn * wsv.re, (-du.im) - sn * wsv.im);
        if (Fc) dW = mk<T>(dW.re - fcv.re, dW.im - fcv.im);
        if (dtd != 0.0) Wout[col] = rk3_update<T>(a.stage, (T)dtd, w0v, wsv, dW);
    }
    for (int c = 1 + threadIdx.x; c <= d.N; c += NT) Wout[col + c] = mk<T>((T)0, (T)0);
}

template <typename T, int NP, bool TAIL = false, int V = 0>
__device__ __forceinline__ void passC_body(int blk, cpx<T> *lds_all,
                                           const cpx<T> *__restrict__ Huw,
                                           const cpx<T> *__restrict__ Hvw, const cpx<T> *Wn,
                                           const cpx<T> *Wst, cpx<T> *Wout, const FusedDims &d,
                                           const cpx<T> *__restrict__ tw, DevScalars *sc,
                                           const StageArgs &a, const Batch &bt, int tid = -1)
{
    using C = WaveCfg<T, NP, V>;
    // tid: the thread's index within the body's 2 * BLK threads (a block
    // may hold two bodies: the ensemble's stage queue kernel)
    const int th = tid < 0 ? (int)threadIdx.x : tid;
    const int half = th / C::BLK, lane = th % C::BLK;
    // wave-uniform; both halves pass the same barriers
    double dtd = 0.0, nu = 0.0;
    if (half == 0)
        passC_half<T, NP, 0, TAIL, V>(blk, lds_all, lane, Huw, Wn, Wst, Wout, d, tw, sc, a, bt,
                                      dtd, nu);
    else
        passC_half<T, NP, 1, TAIL, V>(blk, lds_all, lane, Hvw, Wn, Wst, Wout, d, tw, sc, a, bt,
                                      dtd, nu);
    if constexpr (TAIL) {
        // the block holding column 0 of a run also writes the run's column K - 1
        const int Kt = d.K - 1, j0 = blk * C::GPB;
        const int tr = (j0 + Kt - 1) / Kt; // first run whose column 0 is at or after j0
        if (tr < bt.nruns && tr * Kt < j0 + C::GPB) // block-uniform
            passC_tail<T, NP, V>(tr, lds_all, Huw, Wn, Wst, Wout, d, a, bt, dtd, nu);
    }
}

template <typename T, int NP, int V = 0>
__global__ __launch_bounds__((2 * WaveCfg<T, NP, V>::BLK))
__attribute__((amdgpu_waves_per_eu(WaveCfg<T, NP, V>::WPE_C))) void
k2_passC(const cpx<T> *__restrict__ Huw, const cpx<T> *__restrict__ Hvw, const cpx<T> *Wn,
         const cpx<T> *Wst, cpx<T> *Wout, FusedDims d, const cpx<T> *__restrict__ tw,
         const int *__restrict__ trunc_i1, DevScalars *sc, StageArgs a, Batch bt)
{
    using C = WaveCfg<T, NP, V>;
    __shared__ cpx<T> lds_all[2 * C::GPB * C::SLICE];
    passC_body<T, NP, passC_tail_on<NP>(), V>(blockIdx.x, lds_all, Huw, Hvw, Wn, Wst, Wout, d, tw,
                                          sc, a, bt);
}

} // namespace s2d
