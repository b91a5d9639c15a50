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
};

template <typename T>
struct EnsArgs {
    int ncls;
    EnsClass<T> c[ENS_MAX_CLASSES];
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

} // namespace s2d
