// cnn.hip -- the learned subgrid closure: a periodic 3x3 convolutional
// network (inputs omega, psi; output tau) on BF16 activations and weights
// with FP32 accumulation, its backward pass and an Adam step, behind the
// s2d_cnn_* entry points of include/s2d.h. No reference counterpart.
//
// Network: layer 0 conv 2 -> 64, layers 1..4 conv 64 -> 64, layer 5
// conv 64 -> 1; GELU (exact, erf) after layers 0..4. Activations are NHWC
// BF16 (64 channels = 128 bytes per pixel). Layers 0..4 run on the matrix
// cores (v_mfma_f32_16x16x32_bf16) as implicit GEMMs, forward, backward
// data and weight gradient; the thin last layer and the thin layers'
// weight gradients run on the vector ALUs.
//
// Rounding points (stated once here; the CPU oracle follows them): BF16 of
// every weight, of the scaled inputs, of each layer's pre-activation z
// (kept for the backward pass) and of gelu(z) (of the unrounded z); FP32
// output; BF16 of every dZ the backward pass stores.
//
// Parameter vector (FP32), in order: W0 [9][2][64] (tap, ci, co), b0 [64];
// for l = 1..4: Wl [9][64][64] (tap, co, ci), bl [64]; W5 [9][64] (tap,
// ci), b5 [1]. Tap t = 3 dy + dx is the offset (dy - 1, dx - 1), periodic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdarg>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "s2d.h"

namespace {

using u16 = uint16_t;
typedef float float4_t __attribute__((ext_vector_type(4)));

constexpr int CH = 64;       // channels of the hidden layers
constexpr int NMID = 4;      // 64 -> 64 layers
constexpr int WSEG = 128;    // W is a multiple of this (the layer-0 row segment)
constexpr int W64 = 9 * CH * CH;

constexpr long long OFF_W0 = 0, OFF_B0 = 9 * 2 * CH;
constexpr long long off_wl(int l) { return OFF_B0 + CH + (long long)(l - 1) * (W64 + CH); }
constexpr long long off_bl(int l) { return off_wl(l) + W64; }
constexpr long long OFF_W5 = off_wl(NMID + 1), OFF_B5 = OFF_W5 + 9 * CH;
constexpr long long NPARAM = OFF_B5 + 1;

__host__ __device__ inline u16 f2bf(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (u16)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (u16)(u >> 16);
}
__host__ __device__ inline float bf2f(u16 h)
{
    const uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
__device__ inline float bfr(float f) { return bf2f(f2bf(f)); }

// erf by Abramowitz & Stegun 7.1.26, erf(x) = 1 - (a1 t + ... + a5 t^5)
// exp(-x^2) with t = 1 / (1 + p |x|), odd in x; |error| <= 1.5e-7, far below
// the BF16 rounding of every result it feeds. *g = exp(-x^2), which the
// derivative reuses. (The hardware reciprocal and exp2 are within an ulp.)
__device__ inline float erf_as(float x, float *g)
{
    const float ax = fabsf(x);
    const float t = __builtin_amdgcn_rcpf(__builtin_fmaf(0.3275911f, ax, 1.0f));
    const float p =
        t * (0.254829592f +
             t * (-0.284496736f + t * (1.421413741f + t * (-1.453152027f + t * 1.061405429f))));
    *g = __builtin_amdgcn_exp2f(-1.44269504088896341f * (x * x));
    return copysignf(1.0f - p * *g, x);
}
__device__ inline float gelu(float z)
{
    float g;
    return 0.5f * z * (1.0f + erf_as(z * 0.70710678118654752f, &g));
}
// gelu'(z) = Phi(z) + z phi(z), phi(z) = exp(-z^2 / 2) / sqrt(2 pi)
__device__ inline float gelu_grad(float z)
{
    float g;
    const float e = erf_as(z * 0.70710678118654752f, &g);
    return 0.5f * (1.0f + e) + z * g * 0.39894228040143268f;
}

__device__ inline int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

#define S2D_GRID_STRIDE(i, n) \
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (n); \
         i += (long long)gridDim.x * blockDim.x)

// ------------------------------------------------- layers 1..4 (MFMA)
// f(std::integral_constant<int, i>) for i = I..N-1, unrolled at compile time.
template <int N, int I = 0, typename F>
__device__ __forceinline__ void for_each_c(F &&f)
{
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        for_each_c<N, I + 1>(f);
    }
}

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ bf16x8_t as_bf8(uint4 v) { return __builtin_bit_cast(bf16x8_t, v); }

// Two floats to packed BF16, round to nearest even (v_cvt_pk_bf16_f32: the
// rounding of f2bf for every non-NaN value).
__device__ __forceinline__ uint32_t pk_bf16(float lo, float hi)
{
    const bf16x2_t v = {(__bf16)lo, (__bf16)hi};
    return __builtin_bit_cast(uint32_t, v);
}

// LDS rows of 64 BF16 channels (eight 16-byte chunks): chunk c of row r is
// stored at chunk c ^ (r & 7) of the row, so that any eight consecutive rows
// an operand read takes (eight lanes, one chunk each) fall on distinct banks
// (measured: 3 % less time than c ^ ((r >> 1) & 7), which spreads sixteen).
__device__ __forceinline__ int sw_chunk(int r, int c) { return r * 8 + (c ^ (r & 7)); }

// ------------------------------------------------------------ layer 0
// z0[p][co] = b0[co] + sum_t sum_ci W0[t][ci][co] xs_ci[p + off_t], on the
// matrix cores: K = 9 taps x 2 input channels = 18, zero-padded to one
// 32-wide step of v_mfma_f32_16x16x32_bf16 (k = 2 t + ci). A = the weights
// [co][k] (held in registers), B = the scaled, BF16-rounded inputs [k][pixel]
// (staged in LDS as (ci 0, ci 1) pairs per pixel). Input planes: channel c
// of sample n at xc + n*sN, row pitch ld (floats).
// One block: a 128-pixel row segment; wave w computes pixels 32 w .. 32 w + 31
// (2 N tiles) x 64 channels (4 M tiles).
constexpr int C0_SEG = 128;
__global__ __launch_bounds__(256) void k_conv_in(const float *__restrict__ x0,
                                                 const float *__restrict__ x1, long long sN,
                                                 int ld, float s0, float s1,
                                                 const float *__restrict__ P, u16 *__restrict__ z,
                                                 u16 *__restrict__ a, int B, int H, int W)
{
    __shared__ uint32_t xs[3][C0_SEG + 2]; // (ci 0, ci 1) BF16 pairs, pixels x0 - 1 .. x0 + 128
    const int segs = W / C0_SEG;
    const int row = blockIdx.x / segs, seg = blockIdx.x % segs; // row = n H + y
    const int n = row / H, yy = row - n * H, xb = seg * C0_SEG;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int l16 = lane & 15, g4 = lane >> 4;
    for (int q = tid; q < 3 * (C0_SEG + 2); q += 256) {
        const int r = q / (C0_SEG + 2), j = q % (C0_SEG + 2);
        const int y2 = wrap(yy + r - 1, H), x2 = wrap(xb + j - 1, W);
        const long long o = n * sN + (long long)y2 * ld + x2;
        xs[r][j] = pk_bf16(s0 * x0[o], s1 * x1[o]);
    }
    // A: lane (co = 16 mt + l16, k = 8 g4 + e): W0[t][ci][co] for k = 2 t + ci < 18
    bf16x8_t av[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        uint32_t w[4];
#pragma unroll
        for (int e2 = 0; e2 < 4; ++e2) {
            const int k = 8 * g4 + 2 * e2; // k and k + 1: tap k / 2, ci 0 and 1
            const int co = 16 * mt + l16;
            w[e2] = k < 18 ? pk_bf16(P[OFF_W0 + (k / 2 * 2 + 0) * CH + co],
                                     P[OFF_W0 + (k / 2 * 2 + 1) * CH + co])
                           : 0u;
        }
        av[mt] = as_bf8(uint4{w[0], w[1], w[2], w[3]});
    }
    // the bias before any store (vmcnt counts stores too: a load issued
    // after the stores would wait for all of them)
    float4 bb[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
        bb[mt] = *reinterpret_cast<const float4 *>(P + OFF_B0 + 16 * mt + 4 * g4);
    __syncthreads();
    // B: lane (pixel 32 w + 16 nt + l16, k = 8 g4 + e): the pair of tap (8 g4 + e) / 2
    float4_t acc[4][2];
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        const int j = 32 * wv + 16 * nt + l16; // pixel xb + j
        uint32_t w[4];
#pragma unroll
        for (int e2 = 0; e2 < 4; ++e2) {
            const int t = 4 * g4 + e2; // the tap of k = 8 g4 + 2 e2
            w[e2] = t < 9 ? xs[t / 3][j + t % 3] : 0u;
        }
        const bf16x8_t bv = as_bf8(uint4{w[0], w[1], w[2], w[3]});
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
            acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                av[mt], bv, float4_t{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
    }
    // lane: channels co = 16 mt + 4 g4 + i of pixel xb + 32 w + 16 nt + l16
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        const size_t pix = ((size_t)row * W + xb + 32 * wv + 16 * nt + l16) * CH + 4 * g4;
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            const float v[4] = {acc[mt][nt][0] + bb[mt].x, acc[mt][nt][1] + bb[mt].y,
                                acc[mt][nt][2] + bb[mt].z, acc[mt][nt][3] + bb[mt].w};
            const size_t o = pix + 16 * mt;
            if (z) *reinterpret_cast<uint2 *>(z + o) = uint2{pk_bf16(v[0], v[1]), pk_bf16(v[2], v[3])};
            *reinterpret_cast<uint2 *>(a + o) =
                uint2{pk_bf16(gelu(v[0]), gelu(v[1])), pk_bf16(gelu(v[2]), gelu(v[3]))};
        }
    }
}

// One block: R rows (a multiple of 8) of a 64-pixel column segment of one
// sample, all 64 output channels. The nine taps' weights stay in LDS for
// the whole block (72 KiB); the input rows go through a ring of ten LDS
// rows (64 + 2 halo pixels each). A step makes eight output rows, one per
// wave: wave w's row is a 64 (co) x 64 (pixel) tile, 4 x 4 MFMA tiles of
// v_mfma_f32_16x16x32_bf16 (A = weights [co][ci], B = input [pixel][ci];
// K = 9 taps x 64 ci in chunks of 32). The eight input rows of the next
// step are loaded into registers during the step and written over the ring
// rows it no longer needs.
//   mode 0: out = acc + bias -> zo (BF16; skipped when zo is null),
//           gelu(acc + bias) -> ao (BF16)
//   mode 1: out = acc * gelu'(zp) -> zo (BF16)   (the backward data pass:
//           in = dZ_l, wt = the flipped transpose of W_l, zp = z_{l-1})
// wt: BF16 [9][64 co][64 ci]. Output stores: each lane writes four
// channels (8 bytes) of one pixel per tile.
constexpr int CV_SEG = 64, CV_PX = CV_SEG + 2, CV_RING = 10, CV_THREADS = 512;
constexpr int CV_ROWCH = CV_PX * 8; // 16-byte chunks of a ring row

__global__ __launch_bounds__(CV_THREADS) __attribute__((amdgpu_waves_per_eu(2))) void k_conv64(
    const u16 *__restrict__ in, const u16 *__restrict__ wt, const float *__restrict__ bias,
    const u16 *__restrict__ zp, u16 *__restrict__ zo, u16 *__restrict__ ao, int H, int W, int R,
    int mode)
{
    __shared__ uint4 sw[9 * CH * 8];             // 72 KiB
    __shared__ uint4 sr[CV_RING * CV_ROWCH];     // 82.5 KiB
    __shared__ __align__(16) float sb[CH];       // the bias (zeros without one)
    const int segs = W / CV_SEG, nyb = H / R;
    int b = blockIdx.x;
    const int seg = b % segs;
    b /= segs;
    const int yb = b % nyb, n = b / nyb;
    const int x0 = seg * CV_SEG, y0 = yb * R, S = R / 8;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int l16 = lane & 15, g4 = lane >> 4;
    const uint4 *gin = reinterpret_cast<const uint4 *>(in);
    // chunk q of ring row i (input row y0 - 1 + i)
    auto row_chunk = [&](int i, int q) {
        const int j = q >> 3, c = q & 7;
        const int yy = wrap(y0 - 1 + i, H), xx = wrap(x0 + j - 1, W);
        return gin[(((size_t)n * H + yy) * W + xx) * 8 + c];
    };
    {
        // the weights, then the first ten rows (each: the loads, then the stores)
        const uint4 *gw = reinterpret_cast<const uint4 *>(wt);
        constexpr int NWC = 9 * CH * 8 / CV_THREADS; // 9
        constexpr int NRC = (CV_RING * CV_ROWCH + CV_THREADS - 1) / CV_THREADS; // 11
        uint4 vw[NWC], vr[NRC];
#pragma unroll
        for (int k = 0; k < NWC; ++k) vw[k] = gw[tid + CV_THREADS * k];
#pragma unroll
        for (int k = 0; k < NWC; ++k) {
            const int q = tid + CV_THREADS * k;
            sw[sw_chunk(q >> 3, q & 7)] = vw[k];
        }
#pragma unroll
        for (int k = 0; k < NRC; ++k) {
            const int q = tid + CV_THREADS * k;
            if (q < CV_RING * CV_ROWCH) vr[k] = row_chunk(q / CV_ROWCH, q % CV_ROWCH);
        }
#pragma unroll
        for (int k = 0; k < NRC; ++k) {
            const int q = tid + CV_THREADS * k;
            if (q < CV_RING * CV_ROWCH) {
                const int i = q / CV_ROWCH, r = q % CV_ROWCH;
                sr[sw_chunk(i * CV_PX + (r >> 3), r & 7)] = vr[k];
            }
        }
        if (tid < CH) sb[tid] = bias ? bias[tid] : 0.f;
    }
    __syncthreads();
    // Waves 0-3 (group 0) and 4-7 (group 1) share the SIMDs one to one and run
    // in opposite phases, so that one wave's MFMAs overlap the other's
    // epilogue (VALU) on every SIMD. Step s, output rows y0 + 8 s + 4 grp + wr:
    //   phase 1: group 0 MFMAs; group 1 the loads of ring rows 8 s + 10 ..
    //            8 s + 13 and the epilogue of its rows of step s - 1
    //   phase 2: group 1 stores those rows, then its MFMAs; group 0 the
    //            loads of ring rows 8 s + 14 .. 8 s + 17 (stored at the start
    //            of step s + 1) and its epilogue.
    // Ring rows 8 s + 10 .. 13 take the slots of rows 8 s .. 8 s + 3 (last read
    // by group 0 in phase 1), rows 8 s + 14 .. 17 those of 8 s + 4 .. 8 s + 7
    // (last read by group 1 in phase 2); a barrier ends each phase.
    const int grp = wv >> 2, wr = wv & 3, gt = tid & 255;
    constexpr int NPF = (4 * CV_ROWCH + 255) / 256; // 9 chunks per thread of a group
    uint4 pf[NPF];
    auto load_rows = [&](int i0) { // ring rows i0 .. i0 + 3, the group's 256 threads
#pragma unroll
        for (int k = 0; k < NPF; ++k) {
            const int q = gt + 256 * k;
            if (q < 4 * CV_ROWCH) pf[k] = row_chunk(i0 + q / CV_ROWCH, q % CV_ROWCH);
        }
    };
    auto store_rows = [&](int i0) {
#pragma unroll
        for (int k = 0; k < NPF; ++k) {
            const int q = gt + 256 * k;
            if (q < 4 * CV_ROWCH) {
                const int i = i0 + q / CV_ROWCH, r = q % CV_ROWCH;
                sr[sw_chunk((i % CV_RING) * CV_PX + (r >> 3), r & 7)] = pf[k];
            }
        }
    };
    float4_t acc[4][4];
    // the output row of ring row i0 + 1 (input rows i0 .. i0 + 2): 18 K steps
    // (tap, 32-channel half); the fragments of step k + 1 are read before the
    // MFMAs of step k, and the schedule keeps that order (8 LDS reads, then
    // 16 MFMAs): 13 % less time than letting the compiler interleave them.
    auto mfma_row = [&](int i0) {
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) acc[mt][nt] = float4_t{0.f, 0.f, 0.f, 0.f};
        int bases[3];
#pragma unroll
        for (int dy = 0; dy < 3; ++dy) bases[dy] = ((i0 + dy) % CV_RING) * CV_PX + l16;
        bf16x8_t av[2][4], bv[2][4];
        auto load = [&](auto kI, int buf) {
            constexpr int kk = decltype(kI)::value, t = kk / 2, kc = kk % 2;
            constexpr int dy = t / 3, dx = t % 3;
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
                av[buf][mt] = as_bf8(sw[sw_chunk(t * CH + 16 * mt + l16, 4 * kc + g4)]);
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
                bv[buf][nt] = as_bf8(sr[sw_chunk(bases[dy] + dx + 16 * nt, 4 * kc + g4)]);
        };
        load(std::integral_constant<int, 0>{}, 0);
        for_each_c<18>([&](auto kI) {
            constexpr int kk = decltype(kI)::value, cur = kk & 1;
            if constexpr (kk + 1 < 18) load(std::integral_constant<int, kk + 1>{}, cur ^ 1);
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
#pragma unroll
                for (int nt = 0; nt < 4; ++nt)
                    acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                        av[cur][mt], bv[cur][nt], acc[mt][nt], 0, 0, 0);
            if constexpr (kk + 1 < 18) __builtin_amdgcn_sched_group_barrier(0x100, 8, 0); // DS reads
            __builtin_amdgcn_sched_group_barrier(0x008, 16, 0);                          // MFMAs
        });
    };
    // lane: channels co = 16 mt + 4 g4 + i of pixel x0 + 16 nt + l16 of image row y.
    // No global load follows a store here: vmcnt counts both, so a load
    // issued after the stores would wait for all of them (measured: 24 %
    // of the forward time, with a bias load per tile). The bias comes from
    // LDS; the backward pass loads its 16 z values before the first store.
    auto epilogue = [&](int y) __attribute__((always_inline)) {
        const size_t pix0 = (((size_t)n * H + y) * W + x0 + l16) * CH + 4 * g4;
        if (mode == 0) {
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
#pragma unroll
                for (int mt = 0; mt < 4; ++mt) {
                    const size_t o = pix0 + (size_t)16 * nt * CH + 16 * mt;
                    const float4 bb = *reinterpret_cast<const float4 *>(sb + 16 * mt + 4 * g4);
                    const float v0 = acc[mt][nt][0] + bb.x, v1 = acc[mt][nt][1] + bb.y,
                                v2 = acc[mt][nt][2] + bb.z, v3 = acc[mt][nt][3] + bb.w;
                    if (zo)
                        *reinterpret_cast<uint2 *>(zo + o) = uint2{pk_bf16(v0, v1), pk_bf16(v2, v3)};
                    *reinterpret_cast<uint2 *>(ao + o) =
                        uint2{pk_bf16(gelu(v0), gelu(v1)), pk_bf16(gelu(v2), gelu(v3))};
                }
            }
        } else {
            uint2 zz[16]; // tile q: nt = q / 4, mt = q % 4
#pragma unroll
            for (int q = 0; q < 16; ++q)
                zz[q] = *reinterpret_cast<const uint2 *>(zp + pix0 + (size_t)16 * (q >> 2) * CH +
                                                         16 * (q & 3));
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int nt = q >> 2, mt = q & 3;
                const float4_t a = acc[mt][nt];
                const float g0 = a[0] * gelu_grad(bf2f((u16)(zz[q].x & 0xffffu)));
                const float g1 = a[1] * gelu_grad(bf2f((u16)(zz[q].x >> 16)));
                const float g2 = a[2] * gelu_grad(bf2f((u16)(zz[q].y & 0xffffu)));
                const float g3 = a[3] * gelu_grad(bf2f((u16)(zz[q].y >> 16)));
                *reinterpret_cast<uint2 *>(zo + pix0 + (size_t)16 * nt * CH + 16 * mt) =
                    uint2{pk_bf16(g0, g1), pk_bf16(g2, g3)};
            }
        }
    };
    for (int s = 0; s < S; ++s) {
        const bool more = s + 1 < S;
        // phase 1
        if (grp == 0) {
            if (s > 0) store_rows(8 * s + 6); // rows 8 (s - 1) + 14 .. 17
            mfma_row(8 * s + wr);
        } else { // (the loads ahead of the stores: their wait leaves the stores in flight)
            if (more) load_rows(8 * s + 10);
            if (s > 0) epilogue(y0 + 8 * (s - 1) + 4 + wr);
        }
        __syncthreads();
        // phase 2
        if (grp == 1) {
            if (more) store_rows(8 * s + 10);
            mfma_row(8 * s + 4 + wr);
        } else {
            if (more) load_rows(8 * s + 14);
            epilogue(y0 + 8 * s + wr);
        }
        __syncthreads();
    }
    if (grp == 1) epilogue(y0 + 8 * (S - 1) + 4 + wr);
}

// Weight gradient of a 64 -> 64 layer on the matrix cores:
// dW[t][co][ci] = sum_p dZ[p][co] a[p + off_t][ci] = sum_q dZ[q - off_t][co] a[q][ci],
// db[co] = sum_p dZ[p][co]. K = pixels, so both operands are staged
// channel-major ([channel][pixel], 16-byte chunks of 8 pixels); the tap
// shift goes on dZ, staged once per row in three x-shifted copies, so every
// operand read is an aligned chunk.
// One block: a 32-pixel column strip of R rows of one sample (rows q of a).
// Row step q: the MFMA K is the strip's 32 pixels; wave w owns co tile w
// for the 9 taps x 4 ci tiles (36 accumulator tiles). dZ rows q - 1 .. q + 1
// sit in a ring of four (three x-shifts each), a's row q in one of two
// buffers; the next step's rows are loaded into registers during the step.
// Writes per-block partial sums (reduced by k_reduce).
constexpr int DW_SEG = 32, DW_THREADS = 256;

// chunk position of [row r][chunk c] in a channel-major LDS row of four
// chunks (eight consecutive rows, two per 128 bytes, on distinct banks)
__device__ __forceinline__ int dw_chunk(int r, int c) { return r * 4 + (c ^ ((r >> 1) & 3)); }

__global__ __launch_bounds__(DW_THREADS) __attribute__((amdgpu_waves_per_eu(2))) void k_dw64(
    const u16 *__restrict__ dz, const u16 *__restrict__ a, int H, int W, int R,
    float *__restrict__ part)
{
    __shared__ uint4 zt[4 * 3 * CH * 4]; // [ring slot][dx][co][4 chunks]: 48 KiB
    __shared__ uint4 at[2 * CH * 4];     // [buffer][ci][4 chunks]: 8 KiB
    __shared__ float dbs[4][CH];
    const int segs = W / DW_SEG, nyb = H / R;
    int b = blockIdx.x;
    const int seg = b % segs;
    b /= segs;
    const int yb = b % nyb, n = b / nyb;
    const int x0 = seg * DW_SEG, q0 = yb * R;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int l16 = lane & 15, g4 = lane >> 4;
    // staging roles: channel ch = lane, pixel group jg = wave (pixels 8 jg .. 8 jg + 7)
    const int ch = lane, jg = wv;
    // dZ row k of the ring (image row q0 - 1 + k): the ten pixels
    // x0 + 8 jg - 1 .. x0 + 8 jg + 8 of channel ch
    auto load_dz = [&](int k, u16 (&v)[10]) {
        const size_t rb = ((size_t)n * H + wrap(q0 - 1 + k, H)) * W;
#pragma unroll
        for (int m = 0; m < 10; ++m) v[m] = dz[(rb + wrap(x0 + 8 * jg - 1 + m, W)) * CH + ch];
    };
    // zt[dx][ch][8 jg + i] = dZ[x0 + 8 jg + i - dx + 1] = v[i - dx + 2]
    auto store_dz = [&](int k, const u16 (&v)[10]) {
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) {
            uint32_t w[4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
                w[i] = (uint32_t)v[2 * i + 2 - dx] | ((uint32_t)v[2 * i + 3 - dx] << 16);
            zt[((k & 3) * 3 + dx) * CH * 4 + dw_chunk(ch, jg)] = uint4{w[0], w[1], w[2], w[3]};
        }
    };
    auto load_a = [&](int q, u16 (&v)[8]) {
        const size_t rb = ((size_t)n * H + q) * W + x0 + 8 * jg;
#pragma unroll
        for (int m = 0; m < 8; ++m) v[m] = a[(rb + m) * CH + ch];
    };
    auto store_a = [&](int buf, const u16 (&v)[8]) {
        uint32_t w[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) w[i] = (uint32_t)v[2 * i] | ((uint32_t)v[2 * i + 1] << 16);
        at[buf * CH * 4 + dw_chunk(ch, jg)] = uint4{w[0], w[1], w[2], w[3]};
    };
    float dsum = 0.f; // db: the strip's own pixels of rows q0 .. q0 + R - 1 (ring rows 1 .. R)
    auto add_db = [&](int k, const u16 (&v)[10]) {
        if (k >= 1 && k <= R)
#pragma unroll
            for (int m = 1; m <= 8; ++m) dsum += bf2f(v[m]);
    };
    {
        u16 v0[10], v1[10], v2[10], va[8];
        load_dz(0, v0);
        load_dz(1, v1);
        load_dz(2, v2);
        load_a(q0, va);
        store_dz(0, v0);
        store_dz(1, v1);
        store_dz(2, v2);
        store_a(0, va);
        add_db(0, v0);
        add_db(1, v1);
        add_db(2, v2);
    }
    __syncthreads();
    float4_t acc[9][4];
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[t][j] = float4_t{0.f, 0.f, 0.f, 0.f};
    // Step i: the MFMAs of output row q0 + i; its end stores the rows of step
    // i + 1 (dZ ring row i + 3, a row q0 + i + 1), loaded a step earlier, and
    // its start issues the loads of those of step i + 2 into the other
    // register set (a whole step of MFMAs covers their latency).
    auto step = [&](int i, u16(&cz)[10], u16(&ca)[8], u16(&fz)[10], u16(&fa)[8])
                    __attribute__((always_inline)) {
        // (unconditional, the last steps' unused rows clamped in bounds: under
        // a branch the compiler packs the halves right after their loads, and
        // waits for them there; nor may the scheduler move code across)
        load_dz(i + 4, fz);
        load_a(i + 2 < R ? q0 + i + 2 : q0, fa);
        __builtin_amdgcn_sched_barrier(0);
        bf16x8_t bv[4];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
            bv[nt] = as_bf8(at[(i & 1) * CH * 4 + dw_chunk(16 * nt + l16, g4)]);
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            const int dy = t / 3, dx = t % 3; // dZ row q - dy + 1 = ring row i + 2 - dy
            const bf16x8_t av = as_bf8(
                zt[(((i + 2 - dy) & 3) * 3 + dx) * CH * 4 + dw_chunk(16 * wv + l16, g4)]);
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
                acc[t][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv[nt], acc[t][nt], 0, 0, 0);
        }
        __builtin_amdgcn_sched_barrier(0);
        // ring row i + 3 takes the slot of row i - 1 and buffer (i + 1) & 1 that
        // of a row q0 + i - 1: both were last read in the previous step
        if (i + 1 < R) {
            store_dz(i + 3, cz);
            store_a((i + 1) & 1, ca);
            add_db(i + 3, cz);
        }
        __syncthreads();
    };
    u16 za[10], aa[8], zb[10], ab[8];
    if (R > 1) {
        load_dz(3, za);
        load_a(q0 + 1, aa);
    }
    for (int i = 0; i < R; i += 2) { // (R is a power of two)
        step(i, za, aa, zb, ab);
        if (i + 1 < R) step(i + 1, zb, ab, za, aa);
    }
    // partial: D[co][ci] per tap; lane holds co = 16 w + 4 g4 + e, ci = 16 nt + l16
    float *pb = part + (size_t)blockIdx.x * (W64 + CH);
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int e = 0; e < 4; ++e)
                pb[((size_t)t * CH + 16 * wv + 4 * g4 + e) * CH + 16 * nt + l16] = acc[t][nt][e];
    dbs[jg][ch] = dsum;
    __syncthreads();
    if (tid < CH) pb[W64 + tid] = dbs[0][tid] + dbs[1][tid] + dbs[2][tid] + dbs[3][tid];
}

// ------------------------------------------------------------ layer 5
// y[p] = b5 + sum_t sum_ci W5[t][ci] a[p + off_t][ci]  (FP32 output, dense
// [B][H][W]).
// Grid (W / 32, B * H): 8 lanes per pixel, 8 channels each.
__global__ __launch_bounds__(256) void k_conv_out(const u16 *__restrict__ a,
                                                  const float *__restrict__ P,
                                                  float *__restrict__ y, int B, int H, int W)
{
    constexpr int NW = 9 * CH + 1, NWI = (NW + 255) / 256;
    __shared__ float w[NW];
    const int row = blockIdx.y, n = row / H, yy = row - n * H; // uniform
    const int xx = blockIdx.x * 32 + (threadIdx.x >> 3), c8 = 8 * (threadIdx.x & 7);
    // every load first (weights and the nine input chunks), then the arithmetic
    float wv[NWI];
    uint4 v[9];
#pragma unroll
    for (int i = 0; i < NWI; ++i) {
        const int j = threadIdx.x + 256 * i;
        if (j < NW) wv[i] = P[OFF_W5 + j]; // OFF_B5 = OFF_W5 + 9 CH
    }
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        const int y2 = wrap(yy + t / 3 - 1, H), x2 = wrap(xx + t % 3 - 1, W);
        v[t] = *reinterpret_cast<const uint4 *>(a + (((size_t)n * H + y2) * W + x2) * CH + c8);
    }
#pragma unroll
    for (int i = 0; i < NWI; ++i) {
        const int j = threadIdx.x + 256 * i;
        if (j < NW) w[j] = j < 9 * CH ? bfr(wv[i]) : wv[i];
    }
    __syncthreads();
    float s = 0.f;
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        const u16 *h = reinterpret_cast<const u16 *>(&v[t]);
        for (int k = 0; k < 8; ++k) s += bf2f(h[k]) * w[t * CH + c8 + k];
    }
    s += __shfl_xor(s, 1);
    s += __shfl_xor(s, 2);
    s += __shfl_xor(s, 4);
    if ((threadIdx.x & 7) == 0) y[(size_t)row * W + xx] = s + w[9 * CH];
}

// Scaled output planes: out[n][yy][xx] (row pitch ld, sample stride sN) =
// scale * y (the prediction in physical units). clip != 0 keeps only the
// enstrophy sink: tau where omega * tau > 0, zero elsewhere (the closure
// term -tau then never raises sum omega^2 pointwise).
__global__ void k_write_out(const float *__restrict__ y, float *__restrict__ out, long long sN,
                            int ld, float scale, int B, int H, int W,
                            const float *__restrict__ w, long long wN, int wld, int clip)
{
    // grid (W / 128, B * H), 128 pixels of one row per block
    const int row = blockIdx.y, n = row / H, yy = row - n * H; // uniform
    const int xx = blockIdx.x * 128 + threadIdx.x;
    float v = scale * y[(size_t)row * W + xx];
    if (clip && !(v * w[n * wN + (long long)yy * wld + xx] > 0.f)) v = 0.f;
    out[n * sN + (long long)yy * ld + xx] = v;
}

// Loss (mean square of y - target / out_scale) partial sums and dy.
__global__ __launch_bounds__(256) void k_loss(const float *__restrict__ y,
                                              const float *__restrict__ tg, long long sN, int ld,
                                              float inv_scale, float *__restrict__ dy,
                                              double *__restrict__ part, int B, int H, int W)
{
    __shared__ double red[256];
    const long long npx = (long long)B * H * W;
    double s = 0.0;
    S2D_GRID_STRIDE(p, npx)
    {
        const int xx = (int)(p % W), yy = (int)((p / W) % H), n = (int)(p / ((long long)W * H));
        const float r = y[p] - tg[n * sN + (long long)yy * ld + xx] * inv_scale;
        dy[p] = 2.0f * r / (float)npx;
        s += (double)r * (double)r;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Backward of layer 5: dZ4[q][ci] = (sum_t dy[q - off_t] W5[t][ci]) *
// gelu'(z4[q][ci]) (BF16). Eight lanes per pixel, eight channels each; a
// block takes OB_PX pixels, OB_PPT per lane, whose loads are all issued
// before the first store (a loop with a store between loads waits for the
// store at every load: vmcnt counts both). The rounded W5 sits in LDS.
constexpr int OB_PPT = 4, OB_PX = 256 / 8 * OB_PPT;
__global__ __launch_bounds__(256) void k_out_bwd_data(const float *__restrict__ dy,
                                                      const float *__restrict__ P,
                                                      const u16 *__restrict__ z4,
                                                      u16 *__restrict__ dz, int B, int H, int W)
{
    __shared__ __align__(16) float ws[9 * CH];
    for (int i = threadIdx.x; i < 9 * CH; i += 256) ws[i] = bfr(P[OFF_W5 + i]);
    const int c8 = 8 * (threadIdx.x & 7), npx = B * H * W; // pixels < 2^31 (s2d_cnn_create)
    int qs[OB_PPT];
    uint4 zz[OB_PPT];
    float d[OB_PPT][9];
#pragma unroll
    for (int j = 0; j < OB_PPT; ++j) {
        qs[j] = blockIdx.x * OB_PX + 32 * j + (int)(threadIdx.x >> 3);
        const int q = qs[j] < npx ? qs[j] : npx - 1; // (past the end: loaded, not stored)
        const int xx = q % W, yy = (q / W) % H, n = q / (W * H);
        zz[j] = *reinterpret_cast<const uint4 *>(z4 + (size_t)q * CH + c8);
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            const int y2 = wrap(yy - (t / 3 - 1), H), x2 = wrap(xx - (t % 3 - 1), W);
            d[j][t] = dy[((size_t)n * H + y2) * W + x2];
        }
    }
    __syncthreads();
    float w[9][8];
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        const float4 lo = *reinterpret_cast<const float4 *>(ws + t * CH + c8);
        const float4 hi = *reinterpret_cast<const float4 *>(ws + t * CH + c8 + 4);
        w[t][0] = lo.x, w[t][1] = lo.y, w[t][2] = lo.z, w[t][3] = lo.w;
        w[t][4] = hi.x, w[t][5] = hi.y, w[t][6] = hi.z, w[t][7] = hi.w;
    }
#pragma unroll
    for (int j = 0; j < OB_PPT; ++j) {
        const uint32_t zw[4] = {zz[j].x, zz[j].y, zz[j].z, zz[j].w};
        float o[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            float s = 0.f;
#pragma unroll
            for (int t = 0; t < 9; ++t) s += d[j][t] * w[t][k];
            const u16 zh = (u16)((k & 1) ? zw[k >> 1] >> 16 : zw[k >> 1] & 0xffffu);
            o[k] = s * gelu_grad(bf2f(zh));
        }
        if (qs[j] < npx)
            *reinterpret_cast<uint4 *>(dz + (size_t)qs[j] * CH + c8) =
                uint4{pk_bf16(o[0], o[1]), pk_bf16(o[2], o[3]), pk_bf16(o[4], o[5]),
                      pk_bf16(o[6], o[7])};
    }
}

// Weight gradients of the thin layers: one block per 64-pixel column
// segment of R rows of one sample; thread (channel = lane, pixel group =
// wave) walks 16 pixels of each row, accumulates in FP32, and the four
// waves' sums are added in LDS. Per-block partial sums, reduced by k_reduce.
// R: GR, doubled (dividing H) while at least `width` blocks remain.
constexpr int GR = 8, GSEG = 64, CNN_THIN_WIDTH = 512;
int thin_rows(int B, int H, int W, int width)
{
    int R = GR;
    while (H % (2 * R) == 0 && (long long)B * (H / (2 * R)) * (W / GSEG) >= width) R *= 2;
    return R;
}
int thin_blocks(int B, int H, int W, int width)
{
    return B * (H / thin_rows(B, H, W, width)) * (W / GSEG);
}

// Layer 5: dW5[t][ci] = sum_p dy[p] a4[p + off_t][ci], db5 = sum_p dy[p];
// partial layout [t][ci] then db5.
__global__ __launch_bounds__(256) void k_out_bwd_w(const float *__restrict__ dy,
                                                   const u16 *__restrict__ a, int H, int W, int R,
                                                   float *__restrict__ part)
{
    __shared__ float red[4][9 * CH + 1];
    const int segs = W / GSEG, rb = H / R;
    const int seg = blockIdx.x % segs, r = (blockIdx.x / segs) % rb, n = blockIdx.x / (segs * rb);
    const int ci = threadIdx.x & 63, jg = threadIdx.x >> 6;
    const int xs0 = seg * GSEG + 16 * jg;
    float s[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, sb = 0.f;
    for (int yy = r * R; yy < r * R + R; ++yy) {
        const size_t rows[3] = {((size_t)n * H + wrap(yy - 1, H)) * W, ((size_t)n * H + yy) * W,
                                ((size_t)n * H + wrap(yy + 1, H)) * W};
        // window: v[dy][dx] = a[yy + dy - 1][xx + dx - 1][ci]
        float v[3][3];
#pragma unroll
        for (int dy = 0; dy < 3; ++dy)
#pragma unroll
            for (int dx = 0; dx < 2; ++dx)
                v[dy][dx + 1] = bf2f(a[(rows[dy] + wrap(xs0 + dx - 1, W)) * CH + ci]);
        // (unrolled: the row's 48 + 16 loads in flight together instead of
        // four per pixel in turn)
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int xx = xs0 + j;
#pragma unroll
            for (int dy = 0; dy < 3; ++dy) {
                v[dy][0] = v[dy][1];
                v[dy][1] = v[dy][2];
                v[dy][2] = bf2f(a[(rows[dy] + wrap(xx + 1, W)) * CH + ci]);
            }
            const float d = dy[(size_t)n * H * W + (size_t)yy * W + xx];
            sb += d;
#pragma unroll
            for (int t = 0; t < 9; ++t) s[t] += d * v[t / 3][t % 3];
        }
    }
#pragma unroll
    for (int t = 0; t < 9; ++t) red[jg][t * CH + ci] = s[t];
    if (ci == 0) red[jg][9 * CH] = sb;
    __syncthreads();
    float *pb = part + (size_t)blockIdx.x * (9 * CH + 1);
    for (int i = threadIdx.x; i < 9 * CH + 1; i += 256)
        pb[i] = red[0][i] + red[1][i] + red[2][i] + red[3][i];
}

// Layer 0: dW0[t][ci][co] = sum_p xs_ci[p + off_t] dZ0[p][co] (xs the scaled,
// BF16-rounded input), db0[co] = sum_p dZ0[p][co]; partial layout
// [(t, ci)][co] then db0. The inputs go through LDS GR rows at a time.
__global__ __launch_bounds__(256) void k_in_bwd_w(const float *__restrict__ x0,
                                                  const float *__restrict__ x1, long long sN,
                                                  int ld, float s0, float s1,
                                                  const u16 *__restrict__ dz, int H, int W, int R,
                                                  float *__restrict__ part)
{
    constexpr int NO = 9 * 2 * CH + CH;
    __shared__ float xs[2][GR + 2][GSEG + 2];
    __shared__ float red[4][NO];
    const int segs = W / GSEG, rb = H / R;
    const int seg = blockIdx.x % segs, r = (blockIdx.x / segs) % rb, n = blockIdx.x / (segs * rb);
    const int co = threadIdx.x & 63, jg = threadIdx.x >> 6;
    const int xb = seg * GSEG;
    float s[9][2], sb = 0.f;
#pragma unroll
    for (int t = 0; t < 9; ++t) s[t][0] = s[t][1] = 0.f;
    for (int y0 = r * R; y0 < r * R + R; y0 += GR) {
        __syncthreads(); // the previous rows' reads are done
        for (int q = threadIdx.x; q < (GR + 2) * (GSEG + 2); q += 256) {
            const int i = q / (GSEG + 2), j = q % (GSEG + 2);
            const long long o = n * sN + (long long)wrap(y0 + i - 1, H) * ld + wrap(xb + j - 1, W);
            xs[0][i][j] = bfr(s0 * x0[o]);
            xs[1][i][j] = bfr(s1 * x1[o]);
        }
        __syncthreads();
        // (unrolled: the dZ loads of the GR rows in flight together instead
        // of one per pixel in turn)
#pragma unroll
        for (int i = 0; i < GR; ++i)
#pragma unroll
            for (int jj = 0; jj < 16; ++jj) {
                const int j = 16 * jg + jj;
                const float d = bf2f(dz[(((size_t)n * H + y0 + i) * W + xb + j) * CH + co]);
                sb += d;
#pragma unroll
                for (int t = 0; t < 9; ++t) {
                    s[t][0] += xs[0][i + t / 3][j + t % 3] * d;
                    s[t][1] += xs[1][i + t / 3][j + t % 3] * d;
                }
            }
    }
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        red[jg][(2 * t + 0) * CH + co] = s[t][0];
        red[jg][(2 * t + 1) * CH + co] = s[t][1];
    }
    red[jg][9 * 2 * CH + co] = sb;
    __syncthreads();
    float *pb = part + (size_t)blockIdx.x * NO;
    for (int i = threadIdx.x; i < NO; i += 256) pb[i] = red[0][i] + red[1][i] + red[2][i] + red[3][i];
}

// out[i] = sum_b part[b * len + i] for i < len, in FP64 and a fixed order:
// the partials are split into RQ contiguous ranges of blocks, each summed in
// block order by one quarter of the threads (the loads of 32 partials issued
// ahead of their in-order additions), and the RQ range sums are added in
// range order. 64 outputs per block.
constexpr int RQ = 4;
__global__ __launch_bounds__(64 * RQ) void k_reduce(const float *__restrict__ part, int nb,
                                                    int len, float *__restrict__ out)
{
    __shared__ double rs[RQ][64];
    const int il = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + il;
    const int per = (nb + RQ - 1) / RQ, b0 = q * per, b1 = b0 + per < nb ? b0 + per : nb;
    double s = 0.0;
    if (i < len) {
        int b = b0;
        for (; b + 32 <= b1; b += 32) {
            float v[32];
#pragma unroll
            for (int k = 0; k < 32; ++k) v[k] = part[(size_t)(b + k) * len + i];
#pragma unroll
            for (int k = 0; k < 32; ++k) s += (double)v[k];
        }
        for (; b < b1; ++b) s += (double)part[(size_t)b * len + i];
    }
    rs[q][il] = s;
    __syncthreads();
    if (q == 0 && i < len) {
        double t = rs[0][il];
#pragma unroll
        for (int k = 1; k < RQ; ++k) t += rs[k][il];
        out[i] = (float)t;
    }
}

// BF16 copies of the layer weights: wf[l] = W_l (tap, co, ci) and wt[l] =
// the flipped transpose (tap', ci, co) = W_l(8 - tap', co, ci), the weights
// of the backward data pass.
__global__ void k_refresh(const float *__restrict__ P, u16 *__restrict__ wf, u16 *__restrict__ wt)
{
    S2D_GRID_STRIDE(i, (long long)NMID * W64)
    {
        const int l = (int)(i / W64), r = (int)(i % W64);
        const int t = r / (CH * CH), co = (r / CH) % CH, ci = r % CH;
        const u16 v = f2bf(P[off_wl(l + 1) + r]);
        wf[(size_t)l * W64 + r] = v;
        wt[(size_t)l * W64 + ((size_t)(8 - t) * CH + ci) * CH + co] = v;
    }
}

// Adam (bias-corrected), FP32.
__global__ void k_adam(float *__restrict__ P, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, long long n, float lr, float b1, float b2,
                       float eps, float c1, float c2)
{
    S2D_GRID_STRIDE(i, n)
    {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        P[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// Gather samples idx[b] of a dataset of planes (H x W, dense) into a batch.
__global__ void k_gather(const float *__restrict__ src, const int *__restrict__ idx,
                         float *__restrict__ dst, int B, long long plane)
{
    S2D_GRID_STRIDE(i, (long long)B * plane)
    {
        const int b = (int)(i / plane);
        dst[i] = src[(long long)idx[b] * plane + (i - (long long)b * plane)];
    }
}

thread_local std::string g_cnn_err = "no error";

} // namespace

struct s2d_cnn {
    int H = 0, W = 0, Bmax = 0, device = 0;
    hipStream_t stream = nullptr;
    float in_scale[2] = {1.f, 1.f}, out_scale = 1.f;
    int clip = 0; // output mode of s2d_cnn_forward
    int width = 0; // launch width of every kernel (0: each kernel's default)
    long long step = 0;
    // device buffers
    float *P = nullptr, *G = nullptr, *M = nullptr, *V = nullptr;
    u16 *wf = nullptr, *wt = nullptr;
    u16 *z[NMID + 1] = {}, *a[NMID + 1] = {};
    u16 *dz[2] = {};
    float *y = nullptr, *dy = nullptr, *part = nullptr;
    double *lpart = nullptr;
    size_t part_floats = 0;
    int lblocks = 0;
    // training set (device): n samples of w, psi, tau planes; batch buffers
    long long ds_n = 0;
    float *ds[3] = {}, *bt[3] = {};
    int *d_idx = nullptr;
    std::string err;
};

namespace {

int cfail(s2d_cnn *c, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    g_cnn_err = buf;
    return code;
}

#define CNN_HIP(c, call)                                                                   \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) return cfail((c), S2D_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

int check_batch(s2d_cnn *c, int B)
{
    if (!c) return cfail(nullptr, S2D_ERR_INVALID_ARG, "null network");
    if (B < 1 || B > c->Bmax)
        return cfail(c, S2D_ERR_INVALID_ARG, "batch %d outside [1, %d]", B, c->Bmax);
    return S2D_OK;
}

int launched(s2d_cnn *c, const char *what)
{
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return cfail(c, S2D_ERR_HIP, "launch of %s failed: %s", what, hipGetErrorString(e));
    return S2D_OK;
}

#define CNN_TRY(expr)                \
    do {                             \
        int r_ = (expr);             \
        if (r_ != S2D_OK) return r_; \
    } while (0)

// Launch widths (blocks a row-tiled launch keeps at least; the most blocks
// of a grid-stride launch): each kernel's default, or the one
// s2d_cnn_set_launch_width set for every kernel.
int width_of(const s2d_cnn *c, int dflt) { return c->width > 0 ? c->width : dflt; }

unsigned grid_for(long long n, int blk)
{
    long long g = (n + blk - 1) / blk;
    if (g > 65535LL * 16) g = 65535LL * 16;
    return (unsigned)(g < 1 ? 1 : g);
}

// A 64 -> 64 layer: blocks of R rows x 64 pixels, R the largest multiple of
// 8 dividing H that still gives one block per CU (256 at 256^2 x 16).
void conv64(s2d_cnn *c, int B, const u16 *in, const u16 *wt, const float *bias, const u16 *zp,
            u16 *zo, u16 *ao, int mode)
{
    const int H = c->H, W = c->W, segs = W / CV_SEG, width = width_of(c, 256);
    int R = 8;
    while (H % (2 * R) == 0 && (long long)B * segs * (H / (2 * R)) >= width) R *= 2;
    hipLaunchKernelGGL(k_conv64, dim3(B * segs * (H / R)), dim3(CV_THREADS), 0, c->stream, in,
                       wt, bias, zp, zo, ao, H, W, R, mode);
}

// Rows per k_dw64 block: the largest power of two dividing H that keeps at
// least `width` blocks (default 512, two per CU), or 1.
int dw_rows(int B, int H, int W, int width)
{
    int R = 1;
    while (H % (2 * R) == 0 && (long long)B * (W / DW_SEG) * (H / (2 * R)) >= width) R *= 2;
    return R;
}
long long dw_blocks(int B, int H, int W, int width)
{
    return (long long)B * (W / DW_SEG) * (H / dw_rows(B, H, W, width));
}

// Forward of B samples into c->y (raw network output).
// keep_z: store the pre-activations z (the backward pass reads them).
int forward(s2d_cnn *c, int B, const float *x0, const float *x1, long long sN, int ld,
            bool keep_z)
{
    const int H = c->H, W = c->W;
    hipLaunchKernelGGL(k_conv_in, dim3(B * H * (W / C0_SEG)), dim3(256), 0, c->stream, x0, x1, sN,
                       ld, c->in_scale[0], c->in_scale[1], c->P, keep_z ? c->z[0] : nullptr,
                       c->a[0], B, H, W);
    CNN_TRY(launched(c, "k_conv_in"));
    for (int l = 1; l <= NMID; ++l) {
        conv64(c, B, c->a[l - 1], c->wf + (size_t)(l - 1) * W64, c->P + off_bl(l),
               (const u16 *)nullptr, keep_z ? c->z[l] : nullptr, c->a[l], 0);
        CNN_TRY(launched(c, "k_conv64"));
    }
    hipLaunchKernelGGL(k_conv_out, dim3(W / 32, B * H), dim3(256), 0, c->stream, c->a[NMID],
                       c->P, c->y, B, H, W);
    return launched(c, "k_conv_out");
}

// Loss and gradient of B samples (c->G).
int backward(s2d_cnn *c, int B, const float *x0, const float *x1, long long sN, int ld,
             const float *tg, long long tN, int tld, double *loss)
{
    const int H = c->H, W = c->W;
    const long long npx = (long long)B * H * W;
    hipLaunchKernelGGL(k_loss, dim3(c->lblocks), dim3(256), 0, c->stream, c->y, tg, tN, tld,
                       1.0f / c->out_scale, c->dy, c->lpart, B, H, W);
    CNN_TRY(launched(c, "k_loss"));
    // (thin layers: 512 blocks by default; 2048 gave layer 5 20 us less and
    // layer 0 5 us more, and their two reductions 50 us more, at 256^2 x 16)
    const int wth = width_of(c, CNN_THIN_WIDTH), nth = thin_blocks(B, H, W, wth);
    auto reduce = [&](int nb, int len, float *out) {
        hipLaunchKernelGGL(k_reduce, dim3((len + 63) / 64), dim3(64 * RQ), 0, c->stream, c->part,
                           nb, len, out);
        return launched(c, "k_reduce");
    };
    // layer 5
    const int rth = thin_rows(B, H, W, wth);
    hipLaunchKernelGGL(k_out_bwd_w, dim3(nth), dim3(256), 0, c->stream, c->dy, c->a[NMID], H, W,
                       rth, c->part);
    CNN_TRY(launched(c, "k_out_bwd_w"));
    CNN_TRY(reduce(nth, 9 * CH + 1, c->G + OFF_W5));
    hipLaunchKernelGGL(k_out_bwd_data, dim3((unsigned)((npx + OB_PX - 1) / OB_PX)), dim3(256), 0,
                       c->stream, c->dy, c->P, c->z[NMID], c->dz[0], B, H, W);
    CNN_TRY(launched(c, "k_out_bwd_data"));
    int cur = 0;
    const int rdw = dw_rows(B, H, W, width_of(c, 512)), ndw = (int)dw_blocks(B, H, W, width_of(c, 512));
    for (int l = NMID; l >= 1; --l) {
        hipLaunchKernelGGL(k_dw64, dim3(ndw), dim3(DW_THREADS), 0, c->stream, c->dz[cur],
                           c->a[l - 1], H, W, rdw, c->part);
        CNN_TRY(launched(c, "k_dw64"));
        CNN_TRY(reduce(ndw, W64 + CH, c->G + off_wl(l)));
        conv64(c, B, c->dz[cur], c->wt + (size_t)(l - 1) * W64, (const float *)nullptr,
               c->z[l - 1], c->dz[1 - cur], (u16 *)nullptr, 1);
        CNN_TRY(launched(c, "k_conv64 (backward data)"));
        cur = 1 - cur;
    }
    hipLaunchKernelGGL(k_in_bwd_w, dim3(nth), dim3(256), 0, c->stream, x0, x1, sN, ld,
                       c->in_scale[0], c->in_scale[1], c->dz[cur], H, W, rth, c->part);
    CNN_TRY(launched(c, "k_in_bwd_w"));
    CNN_TRY(reduce(nth, 9 * 2 * CH + CH, c->G + OFF_W0));
    std::vector<double> lp((size_t)c->lblocks);
    CNN_HIP(c, hipMemcpyAsync(lp.data(), c->lpart, sizeof(double) * lp.size(),
                              hipMemcpyDeviceToHost, c->stream));
    CNN_HIP(c, hipStreamSynchronize(c->stream));
    double s = 0.0;
    for (double v : lp) s += v;
    if (loss) *loss = s / (double)npx;
    return S2D_OK;
}

} // namespace

extern "C" {

const char *s2d_cnn_last_error(const s2d_cnn *c) { return c ? c->err.c_str() : g_cnn_err.c_str(); }

long long s2d_cnn_param_count(void) { return NPARAM; }

int s2d_cnn_destroy(s2d_cnn *c)
{
    if (!c) return S2D_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    void *ps[] = {c->P,     c->G,    c->M,     c->V,     c->wf,     c->wt,    c->dz[0],
                  c->dz[1], c->y,    c->dy,    c->part,  c->lpart,  c->ds[0], c->ds[1],
                  c->ds[2], c->bt[0], c->bt[1], c->bt[2], c->d_idx};
    for (void *p : ps) (void)hipFree(p);
    for (int l = 0; l <= NMID; ++l) {
        (void)hipFree(c->z[l]);
        (void)hipFree(c->a[l]);
    }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return S2D_OK;
}

int s2d_cnn_create(int H, int W, int batch_max, s2d_cnn **out)
{
    if (!out) return cfail(nullptr, S2D_ERR_INVALID_ARG, "null output pointer");
    *out = nullptr;
    if (H < 8 || W < WSEG || W % WSEG || H % 8 || batch_max < 1)
        return cfail(nullptr, S2D_ERR_UNSUPPORTED,
                     "grid %dx%d: W a multiple of %d, H a multiple of 8, batch >= 1", H, W, WSEG);
    if ((long long)batch_max * H * W >= (1LL << 31)) // pixel indices are 32-bit
        return cfail(nullptr, S2D_ERR_UNSUPPORTED, "batch x grid of 2^31 pixels or more");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return cfail(nullptr, S2D_ERR_NO_DEVICE, "no GPU visible");
    s2d_cnn *c = new s2d_cnn();
    c->H = H;
    c->W = W;
    c->Bmax = batch_max;
    auto bail = [&](const char *what) {
        const std::string m = std::string("allocation failed: ") + what;
        s2d_cnn_destroy(c);
        g_cnn_err = m;
        return S2D_ERR_ALLOC;
    };
    if (hipGetDevice(&c->device) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return bail("stream");
    const size_t npx = (size_t)batch_max * H * W;
    const size_t act = npx * CH * sizeof(u16);
    c->lblocks = 1024;
    // per-block partial sums of the weight-gradient kernels, for any batch
    size_t pf = 0;
    for (int b = 1; b <= batch_max; ++b)
        pf = std::max({pf, (size_t)dw_blocks(b, H, W, 512) * (W64 + CH),
                       (size_t)thin_blocks(b, H, W, CNN_THIN_WIDTH) * (9 * 2 * CH + CH)});
    c->part_floats = pf;
    if (hipMalloc(&c->P, NPARAM * 4) || hipMalloc(&c->G, NPARAM * 4) ||
        hipMalloc(&c->M, NPARAM * 4) || hipMalloc(&c->V, NPARAM * 4) ||
        hipMalloc(&c->wf, (size_t)NMID * W64 * 2) || hipMalloc(&c->wt, (size_t)NMID * W64 * 2) ||
        hipMalloc(&c->dz[0], act) || hipMalloc(&c->dz[1], act) || hipMalloc(&c->y, npx * 4) ||
        hipMalloc(&c->dy, npx * 4) || hipMalloc(&c->part, c->part_floats * 4) ||
        hipMalloc(&c->lpart, c->lblocks * sizeof(double)))
        return bail("buffers");
    for (int l = 0; l <= NMID; ++l)
        if (hipMalloc(&c->z[l], act) || hipMalloc(&c->a[l], act)) return bail("activations");
    if (hipMemsetAsync(c->P, 0, NPARAM * 4, c->stream) ||
        hipMemsetAsync(c->M, 0, NPARAM * 4, c->stream) ||
        hipMemsetAsync(c->V, 0, NPARAM * 4, c->stream) ||
        hipStreamSynchronize(c->stream))
        return bail("initialisation");
    *out = c;
    return S2D_OK;
}

int s2d_cnn_set_params(s2d_cnn *c, const float *params)
{
    if (!c || !params) return cfail(c, S2D_ERR_INVALID_ARG, "null argument");
    CNN_HIP(c, hipSetDevice(c->device));
    CNN_HIP(c, hipMemcpyAsync(c->P, params, NPARAM * 4, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_refresh, dim3(grid_for((long long)NMID * W64, 256)), dim3(256), 0,
                       c->stream, c->P, c->wf, c->wt);
    CNN_TRY(launched(c, "k_refresh"));
    CNN_HIP(c, hipMemsetAsync(c->M, 0, NPARAM * 4, c->stream));
    CNN_HIP(c, hipMemsetAsync(c->V, 0, NPARAM * 4, c->stream));
    c->step = 0;
    CNN_HIP(c, hipStreamSynchronize(c->stream));
    return S2D_OK;
}

int s2d_cnn_get_params(s2d_cnn *c, float *params)
{
    if (!c || !params) return cfail(c, S2D_ERR_INVALID_ARG, "null argument");
    CNN_HIP(c, hipSetDevice(c->device));
    CNN_HIP(c, hipMemcpyAsync(params, c->P, NPARAM * 4, hipMemcpyDeviceToHost, c->stream));
    CNN_HIP(c, hipStreamSynchronize(c->stream));
    return S2D_OK;
}

int s2d_cnn_set_scales(s2d_cnn *c, float in0, float in1, float out)
{
    if (!c) return cfail(nullptr, S2D_ERR_INVALID_ARG, "null network");
    if (!(out != 0.f)) return cfail(c, S2D_ERR_INVALID_ARG, "output scale must be nonzero");
    c->in_scale[0] = in0;
    c->in_scale[1] = in1;
    c->out_scale = out;
    return S2D_OK;
}

int s2d_cnn_grid(const s2d_cnn *c, int *H, int *W)
{
    if (!c || !H || !W) return cfail(nullptr, S2D_ERR_INVALID_ARG, "null argument");
    *H = c->H;
    *W = c->W;
    return S2D_OK;
}

int s2d_cnn_set_launch_width(s2d_cnn *c, int blocks)
{
    if (!c) return cfail(nullptr, S2D_ERR_INVALID_ARG, "null network");
    if (blocks < 0 || blocks > 512)
        return cfail(c, S2D_ERR_INVALID_ARG, "launch width %d outside [0, 512]", blocks);
    c->width = blocks;
    return S2D_OK;
}

int s2d_cnn_set_clip(s2d_cnn *c, int clip)
{
    if (!c) return cfail(nullptr, S2D_ERR_INVALID_ARG, "null network");
    c->clip = clip ? 1 : 0;
    return S2D_OK;
}

int s2d_cnn_forward_on(s2d_cnn *c, void *stream, int B, const float *w, const float *psi,
                       long long sN, int ld, float *out, long long oN, int old)
{
    CNN_TRY(check_batch(c, B));
    if (!w || !psi || !out) return cfail(c, S2D_ERR_INVALID_ARG, "null device pointer");
    CNN_HIP(c, hipSetDevice(c->device));
    hipStream_t own = c->stream;
    c->stream = static_cast<hipStream_t>(stream);
    int r = forward(c, B, w, psi, sN, ld, false);
    if (r == S2D_OK) {
        hipLaunchKernelGGL(k_write_out, dim3(c->W / 128, B * c->H),
                           dim3(128), 0, c->stream, c->y, out, oN, old, c->out_scale, B, c->H,
                           c->W, w, sN, ld, c->clip);
        r = launched(c, "k_write_out");
    }
    c->stream = own;
    return r;
}

int s2d_cnn_forward(s2d_cnn *c, int B, const float *w, const float *psi, long long sN, int ld,
                    float *out, long long oN, int old)
{
    CNN_TRY(check_batch(c, B));
    if (!w || !psi || !out) return cfail(c, S2D_ERR_INVALID_ARG, "null device pointer");
    CNN_HIP(c, hipSetDevice(c->device));
    CNN_TRY(forward(c, B, w, psi, sN, ld, false));
    hipLaunchKernelGGL(k_write_out, dim3(c->W / 128, B * c->H), dim3(128), 0,
                       c->stream, c->y, out, oN, old, c->out_scale, B, c->H, c->W, w, sN, ld,
                       c->clip);
    CNN_TRY(launched(c, "k_write_out"));
    CNN_HIP(c, hipStreamSynchronize(c->stream));
    return S2D_OK;
}

int s2d_cnn_gradient(s2d_cnn *c, int B, const float *w, const float *psi, long long sN, int ld,
                     const float *target, long long tN, int tld, double *loss, float *grad)
{
    CNN_TRY(check_batch(c, B));
    if (!w || !psi || !target) return cfail(c, S2D_ERR_INVALID_ARG, "null device pointer");
    CNN_HIP(c, hipSetDevice(c->device));
    CNN_TRY(forward(c, B, w, psi, sN, ld, true));
    CNN_TRY(backward(c, B, w, psi, sN, ld, target, tN, tld, loss));
    if (grad) {
        CNN_HIP(c, hipMemcpyAsync(grad, c->G, NPARAM * 4, hipMemcpyDeviceToHost, c->stream));
        CNN_HIP(c, hipStreamSynchronize(c->stream));
    }
    return S2D_OK;
}

int s2d_cnn_set_dataset(s2d_cnn *c, long long n, const float *w, const float *psi,
                        const float *tau)
{
    if (!c || !w || !psi || !tau || n < 1)
        return cfail(c, S2D_ERR_INVALID_ARG, "null argument or empty dataset");
    CNN_HIP(c, hipSetDevice(c->device));
    const size_t plane = (size_t)c->H * c->W;
    for (int k = 0; k < 3; ++k) {
        (void)hipFree(c->ds[k]);
        c->ds[k] = nullptr;
        if (hipMalloc(&c->ds[k], sizeof(float) * plane * n) != hipSuccess)
            return cfail(c, S2D_ERR_ALLOC, "dataset of %lld samples does not fit", n);
        if (!c->bt[k] && hipMalloc(&c->bt[k], sizeof(float) * plane * c->Bmax) != hipSuccess)
            return cfail(c, S2D_ERR_ALLOC, "batch buffers");
    }
    if (!c->d_idx && hipMalloc(&c->d_idx, sizeof(int) * c->Bmax) != hipSuccess)
        return cfail(c, S2D_ERR_ALLOC, "batch index buffer");
    const float *src[3] = {w, psi, tau};
    for (int k = 0; k < 3; ++k)
        CNN_HIP(c, hipMemcpyAsync(c->ds[k], src[k], sizeof(float) * plane * n,
                                  hipMemcpyHostToDevice, c->stream));
    CNN_HIP(c, hipStreamSynchronize(c->stream));
    c->ds_n = n;
    return S2D_OK;
}

int s2d_cnn_train_batch(s2d_cnn *c, int B, const int *idx, float lr, double *loss)
{
    CNN_TRY(check_batch(c, B));
    if (!idx) return cfail(c, S2D_ERR_INVALID_ARG, "null index array");
    if (c->ds_n < 1) return cfail(c, S2D_ERR_INVALID_ARG, "no dataset (s2d_cnn_set_dataset)");
    for (int b = 0; b < B; ++b)
        if (idx[b] < 0 || idx[b] >= c->ds_n)
            return cfail(c, S2D_ERR_INVALID_ARG, "sample index %d outside [0, %lld)", idx[b], c->ds_n);
    CNN_HIP(c, hipSetDevice(c->device));
    CNN_HIP(c, hipMemcpyAsync(c->d_idx, idx, sizeof(int) * B, hipMemcpyHostToDevice, c->stream));
    const long long plane = (long long)c->H * c->W;
    for (int k = 0; k < 3; ++k) {
        hipLaunchKernelGGL(k_gather, dim3(grid_for(B * plane, 256)), dim3(256), 0, c->stream,
                           c->ds[k], c->d_idx, c->bt[k], B, plane);
        CNN_TRY(launched(c, "k_gather"));
    }
    return s2d_cnn_train_step(c, B, c->bt[0], c->bt[1], plane, c->W, c->bt[2], plane, c->W, lr,
                              loss);
}

int s2d_cnn_predict(s2d_cnn *c, long long i, float *host_tau)
{
    if (!c || !host_tau) return cfail(c, S2D_ERR_INVALID_ARG, "null argument");
    if (i < 0 || i >= c->ds_n) return cfail(c, S2D_ERR_INVALID_ARG, "sample %lld outside the dataset", i);
    const long long plane = (long long)c->H * c->W;
    CNN_HIP(c, hipSetDevice(c->device));
    // the scaled prediction into the dy scratch plane
    CNN_TRY(s2d_cnn_forward(c, 1, c->ds[0] + i * plane, c->ds[1] + i * plane, plane, c->W,
                            c->dy, plane, c->W));
    CNN_HIP(c, hipMemcpyAsync(host_tau, c->dy, sizeof(float) * plane, hipMemcpyDeviceToHost,
                              c->stream));
    CNN_HIP(c, hipStreamSynchronize(c->stream));
    return S2D_OK;
}

int s2d_cnn_train_step(s2d_cnn *c, int B, const float *w, const float *psi, long long sN, int ld,
                       const float *target, long long tN, int tld, float lr, double *loss)
{
    CNN_TRY(s2d_cnn_gradient(c, B, w, psi, sN, ld, target, tN, tld, loss, nullptr));
    c->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)c->step), c2 = 1.f - std::pow(b2, (float)c->step);
    hipLaunchKernelGGL(k_adam, dim3(grid_for(NPARAM, 256)), dim3(256), 0, c->stream, c->P, c->G,
                       c->M, c->V, (long long)NPARAM, lr, b1, b2, 1e-8f, c1, c2);
    CNN_TRY(launched(c, "k_adam"));
    hipLaunchKernelGGL(k_refresh, dim3(grid_for((long long)NMID * W64, 256)), dim3(256), 0,
                       c->stream, c->P, c->wf, c->wt);
    CNN_TRY(launched(c, "k_refresh"));
    CNN_HIP(c, hipStreamSynchronize(c->stream));
    return S2D_OK;
}

} // extern "C"
