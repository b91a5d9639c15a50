// cnn.hip -- the learned subgrid closure: a periodic 3x3 convolutional
// network (inputs omega, psi; output tau) on BF16 activations and weights
// with FP32 accumulation, its backward pass and an Adam step, behind the
// s2d_cnn_* entry points of include/s2d.h. No reference counterpart.
//
// Network: layer 0 conv 2 -> 64, layers 1..4 conv 64 -> 64, layer 5
// conv 64 -> 1; GELU (exact, erf) after layers 0..4. Activations are NHWC
// BF16 (64 channels = 128 bytes per pixel). Layers 1..4 (97 % of the
// arithmetic) run on the matrix cores (v_mfma_f32_16x16x16bf16_1k) as
// implicit GEMMs; the thin first and last layers run on the vector ALUs.
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
typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

constexpr int CH = 64;       // channels of the hidden layers
constexpr int NMID = 4;      // 64 -> 64 layers
constexpr int TPX = 128;     // pixels per weight-gradient tile (one row segment)
constexpr int LP = CH + 8;   // LDS pitch (BF16) of a staged pixel
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

__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float z)
{
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) +
           z * expf(-0.5f * z * z) * 0.39894228040143268f;
}

__device__ inline int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

#define S2D_GRID_STRIDE(i, n) \
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (n); \
         i += (long long)gridDim.x * blockDim.x)

// ------------------------------------------------------------ layer 0
// Input planes: channel c of sample n at xc + n*sN, row pitch ld (floats).
// Grid (W / 32, B * H): a block is 32 pixels of one row x 8 groups of 8
// output channels.
__global__ __launch_bounds__(256) void k_conv_in(const float *__restrict__ x0,
                                                 const float *__restrict__ x1, long long sN,
                                                 int ld, float s0, float s1,
                                                 const float *__restrict__ P, u16 *__restrict__ z,
                                                 u16 *__restrict__ a, int B, int H, int W)
{
    constexpr int NW = 9 * 2 * CH + CH, NWI = (NW + 255) / 256;
    __shared__ float w[NW];
    const int row = blockIdx.y, n = row / H, yy = row - n * H; // uniform
    const int xx = blockIdx.x * 32 + (threadIdx.x >> 3), c8 = 8 * (threadIdx.x & 7);
    // every load first (weights and the 18 inputs), then the arithmetic
    float wv[NWI], i0[9], i1[9];
#pragma unroll
    for (int i = 0; i < NWI; ++i) {
        const int j = threadIdx.x + 256 * i;
        if (j < NW) wv[i] = j < 9 * 2 * CH ? P[OFF_W0 + j] : P[OFF_B0 + j - 9 * 2 * CH];
    }
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        const int y2 = wrap(yy + t / 3 - 1, H), x2 = wrap(xx + t % 3 - 1, W);
        const long long o = n * sN + (long long)y2 * ld + x2;
        i0[t] = x0[o];
        i1[t] = x1[o];
    }
#pragma unroll
    for (int i = 0; i < NWI; ++i) {
        const int j = threadIdx.x + 256 * i;
        if (j < NW) w[j] = j < 9 * 2 * CH ? bfr(wv[i]) : wv[i];
    }
    __syncthreads();
    float acc[8];
    for (int k = 0; k < 8; ++k) acc[k] = w[9 * 2 * CH + c8 + k];
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        const float a0 = bfr(s0 * i0[t]), a1 = bfr(s1 * i1[t]);
        for (int k = 0; k < 8; ++k)
            acc[k] += a0 * w[(t * 2 + 0) * CH + c8 + k] + a1 * w[(t * 2 + 1) * CH + c8 + k];
    }
    u16 zo[8], ao[8];
    for (int k = 0; k < 8; ++k) {
        zo[k] = f2bf(acc[k]);
        ao[k] = f2bf(gelu(acc[k]));
    }
    const size_t p = ((size_t)row * W + xx) * CH + c8;
    *reinterpret_cast<uint4 *>(z + p) = *reinterpret_cast<const uint4 *>(zo);
    *reinterpret_cast<uint4 *>(a + p) = *reinterpret_cast<const uint4 *>(ao);
}

// ------------------------------------------------- layers 1..4 (MFMA)
// One block: a segment of CT = 64 MT pixels of one row x 64 output
// channels; wave w computes pixels 16 MT w .. (MT 16-pixel M tiles) x 4 N
// tiles of 16 channels. K = 9 taps x 64 input channels, in chunks of 16. The
// input rows y-1..y+1 (with a one-pixel periodic halo) are staged in LDS.
// Every thread issues its share of all nine taps' weights at the start (18
// 16-byte loads in flight with the input staging: one memory latency per
// block), then copies tap by tap into a double-buffered LDS tile.
//   mode 0: out = acc + bias -> zo (BF16), gelu(acc + bias) -> ao (BF16)
//   mode 1: out = acc * gelu'(zp) -> zo (BF16)   (the backward data pass:
//           in = dZ_l, wt = the flipped transpose of W_l, zp = z_{l-1})
// wt: BF16 [9][64 co][64 ci].
// f(std::integral_constant<int, i>) for i = I..N-1, unrolled at compile time.
template <int N, int I = 0, typename F>
__device__ __forceinline__ void for_each_c(F &&f)
{
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        for_each_c<N, I + 1>(f);
    }
}

template <int MT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_conv64(const u16 *__restrict__ in,
                                                const u16 *__restrict__ wt,
                                                const float *__restrict__ bias,
                                                const u16 *__restrict__ zp, u16 *__restrict__ zo,
                                                u16 *__restrict__ ao, int H, int W, int mode)
{
    constexpr int CT = 64 * MT; // pixels of the tile
    __shared__ __align__(16) u16 tin[3 * (CT + 2) * LP];
    __shared__ __align__(16) u16 tw[2][CH * LP];
    const int segs = W / CT;
    int b = blockIdx.x;
    const int seg = b % segs;
    b /= segs;
    const int y = b % H, n = b / H;
    const int x0 = seg * CT;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    uint4 wr[9][2]; // this thread's 16-byte chunks of every tap: q = tid + 256 i
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int q = tid + 256 * i, co = q >> 3, chunk = q & 7;
            wr[t][i] = *reinterpret_cast<const uint4 *>(wt + ((size_t)t * CH + co) * CH + 8 * chunk);
        }
    // the input rows: all loads of the thread in flight together, then the
    // LDS stores
    constexpr int NST = 3 * (CT + 2) * 8, NSI = (NST + 255) / 256;
    uint4 st[NSI];
    for_each_c<NSI>([&](auto iI) {
        constexpr int i = decltype(iI)::value;
        const int q = tid + 256 * i;
        st[i] = uint4{0, 0, 0, 0};
        if (q < NST) {
            const int chunk = q & 7, pr = q >> 3;
            const int r = pr / (CT + 2), px = pr % (CT + 2);
            const int yy = wrap(y + r - 1, H), xx = wrap(x0 + px - 1, W);
            st[i] = *reinterpret_cast<const uint4 *>(in + (((size_t)n * H + yy) * W + xx) * CH +
                                                     8 * chunk);
        }
    });
    for_each_c<NSI>([&](auto iI) {
        constexpr int i = decltype(iI)::value;
        const int q = tid + 256 * i;
        if (q < NST) *reinterpret_cast<uint4 *>(tin + (q >> 3) * LP + 8 * (q & 7)) = st[i];
    });
    // thread tid holds chunks q = tid and tid + 256: co = q >> 3 (and co + 32)
    const int wco = tid >> 3, wch = tid & 7;
    auto store_w = [&](uint4 lo, uint4 hi, int buf) {
        *reinterpret_cast<uint4 *>(tw[buf] + wco * LP + 8 * wch) = lo;
        *reinterpret_cast<uint4 *>(tw[buf] + (wco + 32) * LP + 8 * wch) = hi;
    };
    store_w(wr[0][0], wr[0][1], 0);
    __syncthreads();
    float4_t acc[MT][4];
    for (int m = 0; m < MT; ++m)
        for (int j = 0; j < 4; ++j) acc[m][j] = float4_t{0.f, 0.f, 0.f, 0.f};
    const int row16 = lane & 15, kq = 4 * (lane >> 4);
    // taps at compile time, so that wr stays in registers
    for_each_c<9>([&](auto tI) {
        constexpr int t = decltype(tI)::value;
        if constexpr (t + 1 < 9) store_w(wr[t + 1][0], wr[t + 1][1], (t + 1) & 1);
        constexpr int dy = t / 3, dx = t % 3;
        const u16 *wb = tw[t & 1];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int kk = 16 * c + kq;
            short4_t av[MT], bv[4];
#pragma unroll
            for (int m = 0; m < MT; ++m) {
                const int px = 16 * MT * wv + 16 * m + row16 + dx;
                av[m] = *reinterpret_cast<const short4_t *>(tin + (dy * (CT + 2) + px) * LP + kk);
            }
#pragma unroll
            for (int j = 0; j < 4; ++j)
                bv[j] = *reinterpret_cast<const short4_t *>(wb + (16 * j + row16) * LP + kk);
#pragma unroll
            for (int m = 0; m < MT; ++m)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[m][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[m], bv[j], acc[m][j],
                                                                          0, 0, 0);
        }
        __syncthreads();
    });
    // results through LDS (the input tile is free): [px][co] FP32, pitch 68
    constexpr int SOP = CH + 4;
    float *so = reinterpret_cast<float *>(tin);
    for (int m = 0; m < MT; ++m)
        for (int j = 0; j < 4; ++j)
            for (int e = 0; e < 4; ++e) {
                const int px = 16 * MT * wv + 16 * m + kq + e, co = 16 * j + row16;
                so[px * SOP + co] = acc[m][j][e] + (bias ? bias[co] : 0.f);
            }
    __syncthreads();
    for (int q = tid; q < CT * 8; q += 256) {
        const int px = q >> 3, c0 = 8 * (q & 7);
        const size_t gi = (((size_t)n * H + y) * W + x0 + px) * CH + c0;
        u16 o1[8], o2[8];
        if (mode == 0) {
            for (int k = 0; k < 8; ++k) {
                const float v = so[px * SOP + c0 + k];
                o1[k] = f2bf(v);
                o2[k] = f2bf(gelu(v));
            }
            *reinterpret_cast<uint4 *>(zo + gi) = *reinterpret_cast<const uint4 *>(o1);
            *reinterpret_cast<uint4 *>(ao + gi) = *reinterpret_cast<const uint4 *>(o2);
        } else {
            const uint4 zz = *reinterpret_cast<const uint4 *>(zp + gi);
            const u16 *zh = reinterpret_cast<const u16 *>(&zz);
            for (int k = 0; k < 8; ++k) o1[k] = f2bf(so[px * SOP + c0 + k] * gelu_grad(bf2f(zh[k])));
            *reinterpret_cast<uint4 *>(zo + gi) = *reinterpret_cast<const uint4 *>(o1);
        }
    }
}

// Weight gradient of a 64 -> 64 layer on the matrix cores:
// dW[t][co][ci] = sum_p dZ[p][co] a[p + off_t][ci], db[co] = sum_p dZ[p][co].
// One block: ROWS rows of one sample, segment by segment; wave w owns the
// output rows co in [16w, 16w + 16) for all 9 taps and 4 ci tiles
// (144 accumulator registers). K = pixels: both operands are staged
// transposed ([channel][pixel]); the tap's x shift makes the input operand's
// four pixels start at any 2-byte offset, so it is read element by element.
// Writes per-block partial sums (reduced by k_reduce).
constexpr int DW_ROWS = 8;
constexpr int TP = TPX + 8; // pitch of a transposed pixel row

__device__ inline short4_t ld4_any(const u16 *p)
{
    short4_t r;
    r[0] = (short)p[0];
    r[1] = (short)p[1];
    r[2] = (short)p[2];
    r[3] = (short)p[3];
    return r;
}

__global__ __launch_bounds__(256) void k_dw64(const u16 *__restrict__ dz,
                                              const u16 *__restrict__ a, int H, int W,
                                              float *__restrict__ part)
{
    __shared__ __align__(16) u16 zt[CH * TP];          // dZ [co][px]
    __shared__ __align__(16) u16 at[CH * 3 * TP];      // a [ci][dy][px'] (px' = px + 1)
    __shared__ float dbs[256 * 8];
    const int rb = H / DW_ROWS;
    const int n = blockIdx.x / rb, y0 = (blockIdx.x % rb) * DW_ROWS;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int row16 = lane & 15, kq = 4 * (lane >> 4);
    float4_t acc[9][4];
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[t][j] = float4_t{0.f, 0.f, 0.f, 0.f};
    float dsum[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int y = y0; y < y0 + DW_ROWS; ++y)
        for (int x0 = 0; x0 < W; x0 += TPX) {
            __syncthreads();
            for (int q = tid; q < TPX * 8; q += 256) { // dZ, transposed
                const int px = q >> 3, c0 = 8 * (q & 7);
                const uint4 v = *reinterpret_cast<const uint4 *>(
                    dz + (((size_t)n * H + y) * W + x0 + px) * CH + c0);
                const u16 *h = reinterpret_cast<const u16 *>(&v);
                for (int k = 0; k < 8; ++k) {
                    zt[(c0 + k) * TP + px] = h[k];
                    dsum[k] += bf2f(h[k]); // q & 7 == tid & 7: fixed channels per thread
                }
            }
            for (int q = tid; q < 3 * (TPX + 2) * 8; q += 256) { // a, transposed
                const int c0 = 8 * (q & 7), pr = q >> 3;
                const int r = pr / (TPX + 2), px = pr % (TPX + 2);
                const int yy = wrap(y + r - 1, H), xx = wrap(x0 + px - 1, W);
                const uint4 v = *reinterpret_cast<const uint4 *>(
                    a + (((size_t)n * H + yy) * W + xx) * CH + c0);
                const u16 *h = reinterpret_cast<const u16 *>(&v);
                for (int k = 0; k < 8; ++k) at[((c0 + k) * 3 + r) * TP + px] = h[k];
            }
            __syncthreads();
#pragma unroll
            for (int t = 0; t < 9; ++t) {
                const int dy = t / 3, dx = t % 3;
                for (int ks = 0; ks < TPX / 16; ++ks) {
                    const int p0 = 16 * ks + kq;
                    const short4_t av =
                        *reinterpret_cast<const short4_t *>(zt + (16 * wv + row16) * TP + p0);
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const u16 *src = at + ((16 * j + row16) * 3 + dy) * TP + p0 + dx;
                        const short4_t bv = dx == 0 ? *reinterpret_cast<const short4_t *>(src)
                                                    : ld4_any(src);
                        acc[t][j] =
                            __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[t][j], 0, 0, 0);
                    }
                }
            }
        }
    // partial: D[co][ci] per tap; lane holds rows co = 16w + kq + e, col ci = 16j + row16
    float *pb = part + (size_t)blockIdx.x * (W64 + CH);
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j)
            for (int e = 0; e < 4; ++e) {
                const int co = 16 * wv + kq + e, ci = 16 * j + row16;
                pb[((size_t)t * CH + co) * CH + ci] = acc[t][j][e];
            }
    for (int k = 0; k < 8; ++k) dbs[tid * 8 + k] = dsum[k];
    __syncthreads();
    if (tid < CH) { // channel tid: threads with (thread & 7) == tid / 8, element tid % 8
        float s = 0.f;
        for (int th = tid / 8; th < 256; th += 8) s += dbs[th * 8 + tid % 8];
        pb[W64 + tid] = s;
    }
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
// gelu'(z4[q][ci]) (BF16).
__global__ __launch_bounds__(256) void k_out_bwd_data(const float *__restrict__ dy,
                                                      const float *__restrict__ P,
                                                      const u16 *__restrict__ z4,
                                                      u16 *__restrict__ dz, int B, int H, int W)
{
    __shared__ float w[9 * CH];
    for (int i = threadIdx.x; i < 9 * CH; i += blockDim.x) w[i] = bfr(P[OFF_W5 + i]);
    __syncthreads();
    const long long npx = (long long)B * H * W;
    S2D_GRID_STRIDE(q, npx)
    {
        const int xx = (int)(q % W), yy = (int)((q / W) % H), n = (int)(q / ((long long)W * H));
        float d[9];
        for (int t = 0; t < 9; ++t) {
            const int y2 = wrap(yy - (t / 3 - 1), H), x2 = wrap(xx - (t % 3 - 1), W);
            d[t] = dy[((size_t)n * H + y2) * W + x2];
        }
        for (int c8 = 0; c8 < CH; c8 += 8) {
            const uint4 zz = *reinterpret_cast<const uint4 *>(z4 + q * CH + c8);
            const u16 *zh = reinterpret_cast<const u16 *>(&zz);
            u16 o[8];
            for (int k = 0; k < 8; ++k) {
                float s = 0.f;
                for (int t = 0; t < 9; ++t) s += d[t] * w[t * CH + c8 + k];
                o[k] = f2bf(s * gelu_grad(bf2f(zh[k])));
            }
            *reinterpret_cast<uint4 *>(dz + q * CH + c8) = *reinterpret_cast<const uint4 *>(o);
        }
    }
}

// Weight gradient of layer 5: one block per (sample, row band); thread j in
// [0, 576) owns (t, ci) = (j / 64, j % 64); thread 576 owns db5. Partial
// sums per block (reduced by k_reduce).
constexpr int OUT_ROWS = 8;
__global__ __launch_bounds__(640) void k_out_bwd_w(const float *__restrict__ dy,
                                                   const u16 *__restrict__ a, int H, int W,
                                                   float *__restrict__ part)
{
    const int rb = H / OUT_ROWS;
    const int n = blockIdx.x / rb, y0 = (blockIdx.x % rb) * OUT_ROWS;
    const int j = threadIdx.x;
    if (j > 9 * CH) return;
    const int t = j / CH, ci = j % CH;
    float s = 0.f;
    for (int yy = y0; yy < y0 + OUT_ROWS; ++yy)
        for (int xx = 0; xx < W; ++xx) {
            const float d = dy[((size_t)n * H + yy) * W + xx];
            if (j == 9 * CH) {
                s += d;
            } else {
                const int y2 = wrap(yy + t / 3 - 1, H), x2 = wrap(xx + t % 3 - 1, W);
                s += d * bf2f(a[(((size_t)n * H + y2) * W + x2) * CH + ci]);
            }
        }
    part[(size_t)blockIdx.x * (9 * CH + 1) + j] = s;
}

// Weight gradient of layer 0: thread (g, co) of 192 owns the six (t, ci)
// pairs 6g .. 6g + 5 (pair = 2 t + ci) and, for g == 0, db0[co].
__global__ __launch_bounds__(192) void k_in_bwd_w(const float *__restrict__ x0,
                                                  const float *__restrict__ x1, long long sN,
                                                  int ld, float s0, float s1,
                                                  const u16 *__restrict__ dz, int H, int W,
                                                  float *__restrict__ part)
{
    const int rb = H / OUT_ROWS;
    const int n = blockIdx.x / rb, y0 = (blockIdx.x % rb) * OUT_ROWS;
    const int co = threadIdx.x % CH, g = threadIdx.x / CH;
    float s[6] = {0, 0, 0, 0, 0, 0}, sb = 0.f;
    for (int yy = y0; yy < y0 + OUT_ROWS; ++yy)
        for (int xx = 0; xx < W; ++xx) {
            const float d = bf2f(dz[(((size_t)n * H + yy) * W + xx) * CH + co]);
            sb += d;
            for (int k = 0; k < 6; ++k) {
                const int pr = 6 * g + k, t = pr / 2, ci = pr % 2;
                const int y2 = wrap(yy + t / 3 - 1, H), x2 = wrap(xx + t % 3 - 1, W);
                const float v = (ci == 0 ? x0 : x1)[n * sN + (long long)y2 * ld + x2];
                s[k] += bfr((ci == 0 ? s0 : s1) * v) * d;
            }
        }
    float *pb = part + (size_t)blockIdx.x * (9 * 2 * CH + CH);
    for (int k = 0; k < 6; ++k) {
        const int pr = 6 * g + k, t = pr / 2, ci = pr % 2;
        pb[(t * 2 + ci) * CH + co] = s[k];
    }
    if (g == 0) pb[9 * 2 * CH + co] = sb;
}

// out[i] = sum_b part[b * len + i] (FP64, in block order) for i < len.
__global__ void k_reduce(const float *__restrict__ part, int nb, int len, float *__restrict__ out)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < len; i += gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int b = 0; b < nb; ++b) s += (double)part[(size_t)b * len + i];
        out[i] = (float)s;
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

unsigned grid_for(long long n, int blk)
{
    long long g = (n + blk - 1) / blk;
    if (g > 65535LL * 16) g = 65535LL * 16;
    return (unsigned)(g < 1 ? 1 : g);
}

// A 64 -> 64 layer in 128-pixel row segments (W is a multiple of 128): at
// 256^2 that is 512 blocks, two per CU, all resident at once.
void conv64(s2d_cnn *c, int B, const u16 *in, const u16 *wt, const float *bias, const u16 *zp,
            u16 *zo, u16 *ao, int mode)
{
    const int H = c->H, W = c->W;
    hipLaunchKernelGGL(k_conv64<2>, dim3(B * H * (W / 128)), dim3(256), 0, c->stream, in, wt,
                       bias, zp, zo, ao, H, W, mode);
}

// Forward of B samples into c->y (raw network output).
int forward(s2d_cnn *c, int B, const float *x0, const float *x1, long long sN, int ld)
{
    const int H = c->H, W = c->W;
    const long long npx = (long long)B * H * W;
    hipLaunchKernelGGL(k_conv_in, dim3(W / 32, B * H), dim3(256), 0, c->stream, x0, x1, sN,
                       ld, c->in_scale[0], c->in_scale[1], c->P, c->z[0], c->a[0], B, H, W);
    CNN_TRY(launched(c, "k_conv_in"));
    for (int l = 1; l <= NMID; ++l) {
        conv64(c, B, c->a[l - 1], c->wf + (size_t)(l - 1) * W64, c->P + off_bl(l),
               (const u16 *)nullptr, c->z[l], c->a[l], 0);
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
    const int nrb = B * (H / OUT_ROWS);
    // layer 5
    hipLaunchKernelGGL(k_out_bwd_w, dim3(nrb), dim3(640), 0, c->stream, c->dy, c->a[NMID], H, W,
                       c->part);
    CNN_TRY(launched(c, "k_out_bwd_w"));
    hipLaunchKernelGGL(k_reduce, dim3(grid_for(9 * CH + 1, 256)), dim3(256), 0, c->stream,
                       c->part, nrb, 9 * CH + 1, c->G + OFF_W5);
    CNN_TRY(launched(c, "k_reduce"));
    hipLaunchKernelGGL(k_out_bwd_data, dim3(grid_for(npx, 256)), dim3(256), 0, c->stream, c->dy,
                       c->P, c->z[NMID], c->dz[0], B, H, W);
    CNN_TRY(launched(c, "k_out_bwd_data"));
    int cur = 0;
    const int ndw = B * (H / DW_ROWS);
    for (int l = NMID; l >= 1; --l) {
        hipLaunchKernelGGL(k_dw64, dim3(ndw), dim3(256), 0, c->stream, c->dz[cur], c->a[l - 1],
                           H, W, c->part);
        CNN_TRY(launched(c, "k_dw64"));
        hipLaunchKernelGGL(k_reduce, dim3(grid_for(W64 + CH, 256)), dim3(256), 0, c->stream,
                           c->part, ndw, W64 + CH, c->G + off_wl(l));
        CNN_TRY(launched(c, "k_reduce"));
        conv64(c, B, c->dz[cur], c->wt + (size_t)(l - 1) * W64, (const float *)nullptr,
               c->z[l - 1], c->dz[1 - cur], (u16 *)nullptr, 1);
        CNN_TRY(launched(c, "k_conv64 (backward data)"));
        cur = 1 - cur;
    }
    hipLaunchKernelGGL(k_in_bwd_w, dim3(nrb), dim3(192), 0, c->stream, x0, x1, sN, ld,
                       c->in_scale[0], c->in_scale[1], c->dz[cur], H, W, c->part);
    CNN_TRY(launched(c, "k_in_bwd_w"));
    hipLaunchKernelGGL(k_reduce, dim3(grid_for(9 * 2 * CH + CH, 256)), dim3(256), 0, c->stream,
                       c->part, nrb, 9 * 2 * CH + CH, c->G + OFF_W0);
    CNN_TRY(launched(c, "k_reduce"));
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
    if (H < 8 || W < TPX || W % TPX || H % 8 || batch_max < 1)
        return cfail(nullptr, S2D_ERR_UNSUPPORTED,
                     "grid %dx%d: W a multiple of %d, H a multiple of 8, batch >= 1", H, W, TPX);
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
    const size_t nrb = (size_t)batch_max * (H / 8);
    c->part_floats = nrb * (W64 + CH);
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
    int r = forward(c, B, w, psi, sN, ld);
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
    CNN_TRY(forward(c, B, w, psi, sN, ld));
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
    CNN_TRY(forward(c, B, w, psi, sN, ld));
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
