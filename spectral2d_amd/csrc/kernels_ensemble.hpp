// kernels_ensemble.hpp -- one launch per pass for an ensemble of runs of
// mixed sizes.
//
// Runs of one padded size form a size class (a batch, kernels_fused_v2.hpp
// Batch). A launch covers every class: the grid is the concatenation of the
// classes' block ranges (each a multiple of 8 blocks, so that a block's XCD,
// blockIdx % 8, is also its XCD within the class), and each block runs the
// pass body of its class's compiled size. The sizes served are those whose
// lane group fits one wave (Np / E <= 64): every class then uses the same
// blocks: BLK_A lanes for pass A (BLK_A / TG columns), BLK_B for pass B
// (BLK_B / TG rows, passB_single) and 128 for pass C, and every block
// carries the same number of points whatever its class.
#pragma once

#include "kernels_fused_v2.hpp"

namespace s2d {

constexpr int ENS_MAX_CLASSES = 6;

// One size class as the kernels see it.
template <typename T>
struct EnsClass {
    int np;          // padded size (selects the compiled body)
    int nblk[3];     // blocks of pass A, B, C
    FusedDims d;
    Batch bt;
    StageArgs sa;    // dx, norm, dt, cfl of the class (stage set per launch)
    cpx<T> *Wn, *Ws;
    cpx<T> *G0, *G1, *G2, *H0, *H1;
    const cpx<T> *tw_col, *tw_row;
    const int *trunc;
    DevScalars *sc;
    // the stage queue kernel: items of pass A, B, C, items per run of
    // passes A and B, tiles per run of passes A and B, tiles (pass C:
    // bodies) per item, the class's first run among all classes
    int nitem[3], per_run[2], ntile[2], tpi[3], run0;
};

template <typename T>
struct EnsArgs {
    int ncls;
    EnsClass<T> c[ENS_MAX_CLASSES];
    // the stage queue kernel (k3_ens_stage): work items of each pass
    // over all classes, and the runs over all classes
    int tot[3];
    int nruns_all;
};

template <typename T>
struct EnsCfg {
    static constexpr int E = WaveCfg<T, 192>::E;
    // lanes of a pass A / pass B block (pass A: 512, so that the 512^2
    // class writes its G rows in 64-byte segments; measured 1.5 % less time
    // per ensemble step than 256)
    static constexpr int BLK_A = 512, BLK_B = 256;
    // lines per tile of passes A and B for padded size np
    static constexpr int lpb_a(int np) { return BLK_A * E / np; }
    static constexpr int lpb_b(int np) { return BLK_B * E / np; }
    // the largest tiles of the served sizes (Np = 96)
    static constexpr int LDS_A = TileCfg<T, 96, BLK_A * E / 96>::LDS;
    static constexpr int LDS_B = TileCfg<T, 96, BLK_B * E / 96>::LDS;
    static constexpr int LDS_C = 2 * 64 * E; // 2 * GPB * SLICE of every served size
};

// The padded sizes an ensemble class may have.
template <typename T>
constexpr bool ens_size_ok(int np)
{
    return np == 96 || np == 192 || np == 384 || np == 768;
}

// Class of block b for pass p; b becomes the block index within the class.
template <typename T>
__device__ __forceinline__ int ens_class(const EnsArgs<T> &ea, int p, int &b)
{
    int k = 0;
    while (k < ea.ncls - 1 && b >= ea.c[k].nblk[p]) {
        b -= ea.c[k].nblk[p];
        ++k;
    }
    return k;
}

#define S2D_ENS_SIZES(T, np, CALL)                                       \
    switch (np) {                                                        \
    case 96: { constexpr int NP = 96; CALL; } break;                     \
    case 192: { constexpr int NP = 192; CALL; } break;                   \
    case 384: { constexpr int NP = 384; CALL; } break;                   \
    case 768: { constexpr int NP = 768; CALL; } break;                   \
    }

// V = 0: a block per (field, column tile); V = VAR_FUSE_A: a block per
// column tile for the three fields (passA_fused_body; the class block counts
// of EnsArgs follow the variant).
template <typename T, int V = 0>
__global__ __launch_bounds__(EnsCfg<T>::BLK_A)
__attribute__((amdgpu_waves_per_eu(WaveCfg<T, 192, V>::WPE_A))) void
k3_ens_passA(EnsArgs<T> ea, int stage)
{
    __shared__ cpx<T> lds[EnsCfg<T>::LDS_A];
    int b = blockIdx.x;
    const EnsClass<T> &c = ea.c[ens_class(ea, 0, b)];
    const cpx<T> *src = stage == 1 ? c.Wn : c.Ws;
    if constexpr ((V & VAR_FUSE_A) != 0)
        S2D_ENS_SIZES(T, c.np,
                      (passA_fused_body<T, NP, EnsCfg<T>::lpb_a(NP), V>(
                          b, c.nblk[0], lds, src, c.G0, c.G1, c.G2, c.d, c.tw_col, c.sc,
                          stage == 1, c.bt)))
    else
        S2D_ENS_SIZES(T, c.np,
                      (passA_body<T, NP, EnsCfg<T>::lpb_a(NP)>(b, c.nblk[0], lds, src, c.G0,
                                                              c.G1, c.G2, c.d, c.tw_col, c.sc,
                                                              stage == 1, c.bt)))
}

template <typename T>
__global__ __launch_bounds__(EnsCfg<T>::BLK_B)
__attribute__((amdgpu_waves_per_eu(WaveCfg<T, 192>::WPE_B))) void
k3_ens_passB(EnsArgs<T> ea, int stage)
{
    __shared__ cpx<T> lds[EnsCfg<T>::LDS_B];
    int b = blockIdx.x;
    const EnsClass<T> &c = ea.c[ens_class(ea, 1, b)];
    S2D_ENS_SIZES(T, c.np,
                  (passB_single<T, NP, EnsCfg<T>::lpb_b(NP)>(b, c.nblk[1], lds, c.G0, c.G1, c.G2,
                                                          c.H0, c.H1, c.d, c.tw_row, c.sc, 1,
                                                          stage == 1, c.bt)))
}

template <typename T>
__global__ __launch_bounds__(128)
__attribute__((amdgpu_waves_per_eu(WaveCfg<T, 192>::WPE_C))) void
k3_ens_passC(EnsArgs<T> ea, int stage)
{
    __shared__ cpx<T> lds[EnsCfg<T>::LDS_C];
    int b = blockIdx.x;
    const EnsClass<T> &c = ea.c[ens_class(ea, 2, b)];
    StageArgs a = c.sa;
    a.stage = stage;
    const cpx<T> *src = stage == 1 ? c.Wn : c.Ws;
    cpx<T> *dst = stage == 3 ? c.Wn : c.Ws;
    S2D_ENS_SIZES(T, c.np,
                  (passC_body<T, NP>(b, lds, c.H0, c.H1, c.Wn, src, dst, c.d, c.tw_col, c.sc, a,
                                     c.bt)))
}

// ---- one kernel per RK stage over a work queue (S2D_ENS_STAGE_QUEUE)
//
// A stage's work is one queue of items: every pass A item of every class,
// then every pass B item, then every pass C item. An item is up to
// ENS_SQ_TILES tiles of one run (pass A, B) or ENS_SQ_BODIES pass C bodies
// of one class, two at a time (one per half of the 256 threads). The launch has one
// block per item; each block takes the next ticket of the stage's counter
// (not its blockIdx) and runs that item. A pass B item of a run waits until
// every pass A item of that run has finished, a pass C item until every
// pass B item of its runs has: kernel boundaries become per-run
// dependencies, so the passes of different runs overlap. An item only waits
// on items with smaller tickets, which blocks that have started hold, so
// every wait ends whatever the dispatch order. A wait is still bounded
// (spin_limit polls); one that gives up sets the slot's error word, and the
// remaining items are skipped, so the kernel always exits and the host
// reports the failure (the error word is sticky: later waits give up at
// once, until the host reads and clears it). (One item per block rather than a loop over tickets
// in resident blocks: in a loop the compiler hoists the per-thread set-up
// of all twelve pass bodies out of it, and the kernel spills.)
//
// The queue buffer holds three slots (one per stage) of qstride words,
// [0] ticket counter, [64 + r] pass A items done of run r, [64 + R + r]
// pass B items done (R = runs over all classes), then the error word. The
// launch of stage s uses slot s - 1 and clears slot s % 3, which the next
// launch uses (kernel order keeps the two apart).
constexpr int ENS_PST_THREADS = 256;
constexpr int ENS_SQ_TILES = 8, ENS_SQ_BODIES = 8;
static_assert(ENS_SQ_BODIES % 2 == 0, "pass C bodies go two at a time");

template <typename T>
struct EnsPstCfg {
    static constexpr int LPB_A = ENS_PST_THREADS * EnsCfg<T>::E / 96; // at Np = 96
    static constexpr int LDS_A = TileCfg<T, 96, LPB_A>::LDS;
    static constexpr int LDS_B = EnsCfg<T>::LDS_B;
    static constexpr int LDS_C = 2 * EnsCfg<T>::LDS_C;
    static constexpr int LDS = LDS_A > LDS_B ? (LDS_A > LDS_C ? LDS_A : LDS_C)
                                             : (LDS_B > LDS_C ? LDS_B : LDS_C);
    static constexpr int lpb_a(int np) { return ENS_PST_THREADS * EnsCfg<T>::E / np; }
    static constexpr int WPE = sizeof(T) == 8 ? 2 : 3;
};

// Every served size's pass A and B tiles fill the 256 threads and fit the
// kernel's LDS, and its two pass C bodies fit it too.
template <typename T, int NP>
constexpr bool ens_stage_fits()
{
    using A = TileCfg<T, NP, EnsPstCfg<T>::lpb_a(NP)>;
    using B = TileCfg<T, NP, EnsCfg<T>::lpb_b(NP)>;
    using C = WaveCfg<T, NP>;
    return A::BLK == ENS_PST_THREADS && B::BLK == ENS_PST_THREADS &&
           A::LDS <= EnsPstCfg<T>::LDS && B::LDS <= EnsPstCfg<T>::LDS &&
           2 * C::BLK == 128 && 2 * C::GPB * C::SLICE <= EnsCfg<T>::LDS_C;
}
static_assert(ens_stage_fits<double, 96>() && ens_stage_fits<double, 192>() &&
                  ens_stage_fits<double, 384>() && ens_stage_fits<double, 768>() &&
                  ens_stage_fits<float, 96>() && ens_stage_fits<float, 192>() &&
                  ens_stage_fits<float, 384>() && ens_stage_fits<float, 768>(),
              "a served size does not fit the stage queue kernel");

// p, as a value the compiler cannot see through: the twiddle loads of a
// loop over tiles then stay in the loop instead of being hoisted out of it
// and held in registers across the tiles (which spills).
template <typename P>
__device__ __forceinline__ const P *opaque(const P *p)
{
    __asm__ volatile("" : "+s"(p));
    return p;
}

// Thread 0 waits until q[i] >= target (relaxed polls, then one acquire
// fence); returns false when it gave up or an earlier wait had.
__device__ __forceinline__ bool ens_wait(unsigned *q, int i, unsigned target, unsigned *err,
                                         long long spin_limit)
{
    long long n = 0;
    while (__hip_atomic_load(q + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
        if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u)
            return false;
        if (++n > spin_limit) {
            __hip_atomic_fetch_or(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return true;
}

template <typename T>
__global__ __launch_bounds__(ENS_PST_THREADS)
__attribute__((amdgpu_waves_per_eu(EnsPstCfg<T>::WPE))) void
k3_ens_stage(EnsArgs<T> ea, int stage, unsigned *queue, int qstride, unsigned *err,
             long long spin_limit)
{
    __shared__ cpx<T> lds[EnsPstCfg<T>::LDS];
    __shared__ int s_item, s_ok;
    unsigned *slot = queue + (size_t)(stage - 1) * qstride;
    unsigned *next = queue + (size_t)(stage % 3) * qstride;
    if (blockIdx.x == 0)
        for (int i = threadIdx.x; i < qstride; i += ENS_PST_THREADS) next[i] = 0u;
    const int R = ea.nruns_all;
    const int total = ea.tot[0] + ea.tot[1] + ea.tot[2];
    if (threadIdx.x == 0)
        s_item = (int)__hip_atomic_fetch_add(slot, 1u, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    // block-uniform: read as a scalar, so that the class's fields below
    // are scalar loads
    const int item = __builtin_amdgcn_readfirstlane(s_item);
    if (item >= total) return;
    int p = 0, b = item;
    if (b >= ea.tot[0]) {
        b -= ea.tot[0];
        p = 1;
        if (b >= ea.tot[1]) {
            b -= ea.tot[1];
            p = 2;
        }
    }
    int k = 0;
    while (k < ea.ncls - 1 && b >= ea.c[k].nitem[p]) {
        b -= ea.c[k].nitem[p];
        ++k;
    }
    const EnsClass<T> &c = ea.c[k];
    // the runs this item reads
    int r_lo = 0, r_hi = 0;
    if (p == 1) {
        r_lo = r_hi = b / c.per_run[1];
    } else if (p == 2) {
        const int gpb = 64 * EnsCfg<T>::E / c.np; // column groups per body
        const int j0 = b * c.tpi[2] * gpb, j1 = j0 + c.tpi[2] * gpb - 1;
        r_lo = j0 / c.d.K;
        r_hi = j1 / c.d.K < c.bt.nruns - 1 ? j1 / c.d.K : c.bt.nruns - 1;
    }
    if (threadIdx.x == 0) {
        bool ok = __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u;
        for (int r = r_lo; p > 0 && ok && r <= r_hi; ++r)
            ok = ens_wait(slot, 64 + (p == 1 ? 0 : R) + c.run0 + r,
                          (unsigned)c.per_run[p - 1], err, spin_limit);
        s_ok = ok ? 1 : 0;
    }
    __syncthreads();
    if (__builtin_amdgcn_readfirstlane(s_ok)) {
        const cpx<T> *src = stage == 1 ? c.Wn : c.Ws;
        if (p < 2) {
            // tiles [t0, t1) of run `run`
            const int run = b / c.per_run[p], t0 = (b - run * c.per_run[p]) * c.tpi[p];
            const int t1 = t0 + c.tpi[p] < c.ntile[p] ? t0 + c.tpi[p] : c.ntile[p];
            const int j0 = run * c.ntile[p] + t0, j1 = run * c.ntile[p] + t1;
            auto tiles_a = [&](auto npc) {
                constexpr int NP = decltype(npc)::value;
                for (int j = j0; j < j1; ++j) {
                    if (j > j0) __syncthreads(); // the last tile's LDS reads are done
                    passA_body<T, NP, EnsPstCfg<T>::lpb_a(NP)>(j, 1, lds, src, c.G0, c.G1,
                                                               c.G2, c.d, opaque(c.tw_col), c.sc,
                                                               stage == 1, c.bt);
                }
            };
            auto tiles_b = [&](auto npc) {
                constexpr int NP = decltype(npc)::value;
                for (int j = j0; j < j1; ++j) {
                    if (j > j0) __syncthreads();
                    passB_single<T, NP, EnsCfg<T>::lpb_b(NP)>(j, 1, lds, c.G0, c.G1, c.G2,
                                                              c.H0, c.H1, c.d, opaque(c.tw_row),
                                                              c.sc, 1, stage == 1, c.bt);
                }
            };
            if (p == 0)
                S2D_ENS_SIZES(T, c.np, (tiles_a(std::integral_constant<int, NP>{})))
            else
                S2D_ENS_SIZES(T, c.np, (tiles_b(std::integral_constant<int, NP>{})))
        } else {
            StageArgs a = c.sa;
            a.stage = stage;
            cpx<T> *dst = stage == 3 ? c.Wn : c.Ws;
            const int half = threadIdx.x / 128;
            cpx<T> *lh = lds + half * EnsCfg<T>::LDS_C;
            const int j0 = b * c.tpi[2], j1 = j0 + c.tpi[2];
            auto bodies_c = [&](auto npc) {
                constexpr int NP = decltype(npc)::value;
                for (int j = j0; j < j1; j += 2) {
                    if (j > j0) __syncthreads();
                    passC_body<T, NP>(j + half, lh, c.H0, c.H1, c.Wn, src, dst, c.d,
                                      opaque(c.tw_col), c.sc, a, c.bt, (int)threadIdx.x % 128);
                }
            };
            S2D_ENS_SIZES(T, c.np, (bodies_c(std::integral_constant<int, NP>{})))
        }
    }
    // The item counts as done once every wave's stores are visible to the
    // device: the barrier orders the other waves' stores before thread 0's
    // release (at agent scope a release writes back the XCD's L2, the cost
    // that matters here, so one release per item and none for pass C,
    // whose stores only the next launch reads).
    __syncthreads();
    if (threadIdx.x == 0 && s_ok && p < 2) {
        const int run = b / c.per_run[p];
        __hip_atomic_fetch_add(slot + 64 + (p == 0 ? 0 : R) + c.run0 + run, 1u,
                               __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}


} // namespace s2d
