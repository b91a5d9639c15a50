// kernels_fused.hpp -- the fused, device-resident TVD-RK3 stage pipeline.
//
// State: the vorticity on the kept modes only, a compact (K, KP) array
// W[kx*KP + c] with kx = 0..N/2 (K = N/2+1) and compact row c = ky mod
// (N+1) for ky = -N/2..N/2 (modes outside the disk |k| <= N/2 are held at
// zero). psi = omega/|k|^2 is never stored: it is formed where needed.
//
// One RK stage = three kernels over the 3/2-padded grid (Np = 3N/2):
//  A  (per kx column)  omega -> u_hat = i ky psi, v_hat = -i kx psi,
//     w_hat; zero-pad to Np in FFTW ky order; inverse FFT along y; write the
//     mixed (kx, y) fields G[f][y*KG + kx] (row-major, so that pass B reads
//     whole rows).
//  B  (per pair of rows) two rows' real transforms share one complex
//     transform: z3 = w_a + i w_b from the G_w spectra (Hermitian
//     extension), then for each row z1 = u + i v, the products z1*w =
//     (uw + i vw) in registers, one forward FFT, and the uw/vw spectra are
//     separated from P_k and conj P_{Np-k}; writes H[f][kx*Np + y] for
//     kx <= N/2. Also the max of |u|+|v| for the CFL step.
//  C  (per kx column)  forward FFT along y of H_uw, H_vw, then on the kept
//     modes: /(Np*Np) (TransformR2C), truncation, i kx, i ky, dW =
//     -N - nu |k|^2 omega_stage, and the Shu-Osher stage update
//     (src/TimeIntegration.F90:156-202), writing the next compact omega.
// In exact arithmetic this is src/Spectral.F90:396-482 composed with
// src/TimeIntegration.F90:130-210; the reference's FFT(IFFT(omega)) round
// trip of omega inside ComputeRHS is the identity on a Hermitian, truncated
// field and is not repeated here.
#pragma once

#include "fft_block.hpp"
#include "kernels_ref.hpp"
#include "s2d_internal.hpp"

namespace s2d {

struct FusedDims {
    int N;   // resolved grid
    int Np;  // padded grid (rows and columns)
    int K;   // kept kx columns, N/2 + 1
    int KP;  // compact column pitch
    int cI;  // reference complex row length (Np/2 + 1)
    int KG;  // row pitch of the G fields (K rounded up to 16)
};

template <typename T>
__device__ __forceinline__ T absv(T x) { return x < (T)0 ? -x : x; }

// compact row of FFTW ky index j (0..Np-1), or -1 when |ky| > N/2
__device__ __forceinline__ int compact_row(int j, const FusedDims &d)
{
    const int h = d.N / 2;
    if (j <= h) return j;
    if (j >= d.Np - h) return j - d.Np + d.N + 1;
    return -1;
}

// ------------------------------------------------------------- pass A
template <typename T, int E>
__global__ void k_passA(const cpx<T> *__restrict__ Wsrc, cpx<T> *__restrict__ Gu,
                        cpx<T> *__restrict__ Gv, cpx<T> *__restrict__ Gw, FusedDims d,
                        FftPlan p, const cpx<T> *__restrict__ tw, DevScalars *sc,
                        int reset_max)
{
#pragma clang fp contract(off)
    extern __shared__ __align__(16) unsigned char smem[];
    cpx<T> *lds = reinterpret_cast<cpx<T> *>(smem);
    const int t = threadIdx.x % p.T, f = threadIdx.x / p.T;
    const int kx = blockIdx.x;
    if (reset_max && kx == 0 && threadIdx.x == 0) {
        sc->vmax_bits = 0ull;
        sc->t_start = sc->time;
        sc->bad = 0u;
    }
    const cpx<T> *src = Wsrc + (size_t)kx * d.KP;
    const T qx = (T)kx;
    cpx<T> v[E];
#pragma unroll
    for (int s = 0; s < E; ++s) {
        const int j = t + p.T * s;
        const int c = compact_row(j, d);
        cpx<T> w = (c >= 0) ? src[c] : mk<T>((T)0, (T)0);
        const int kyi = (j <= d.Np / 2) ? j : j - d.Np;
        const T qy = (T)kyi;
        if (f == 2) {
            v[s] = w;
        } else {
            cpx<T> ps;
            if (kx == 0 && kyi == 0) {
                ps = mk<T>((T)0, (T)0);
            } else {
                const T dd = (kx == 0) ? qy * qy : (kyi == 0) ? qx * qx : qx * qx + qy * qy;
                ps = mk<T>(w.re / dd, w.im / dd);
            }
            if (f == 0) {
                v[s] = mk<T>(-(qy * ps.im), qy * ps.re); // II*ky*psi
            } else {
                cpx<T> r = mk<T>(-(qx * ps.im), qx * ps.re); // II*kx*psi
                v[s] = mk<T>((T)-1 * r.re, (T)-1 * r.im);
            }
        }
    }
    block_fft<T, E, 1>(v, lds + (size_t)f * d.Np, p, t, true, tw);
    cpx<T> *dst = (f == 0 ? Gu : f == 1 ? Gv : Gw) + kx;
#pragma unroll
    for (int s = 0; s < E; ++s) dst[(size_t)(t + p.T * s) * d.KG] = v[s];
}

// ------------------------------------------------------------- pass B

// Hermitian-extended spectrum of rows (a, b) packed as A + i B at x
// frequency k of a length-Np complex transform: for k > Np/2 the stored
// conjugate at Np - k is used; only Re of k = 0 is read (c2r convention).
template <typename T>
__device__ __forceinline__ cpx<T> packed_spectrum(const cpx<T> *__restrict__ GA,
                                                  const cpx<T> *__restrict__ GB, int k,
                                                  int ya, int yb, const FusedDims &d)
{
    const int h = d.N / 2;
    cpx<T> A, B;
    if (k <= h) {
        A = GA[(size_t)ya * d.KG + k];
        B = GB[(size_t)yb * d.KG + k];
        if (k == 0) {
            A.im = (T)0;
            B.im = (T)0;
        }
    } else if (k >= d.Np - h) {
        const int m = d.Np - k;
        A = GA[(size_t)ya * d.KG + m];
        B = GB[(size_t)yb * d.KG + m];
        A.im = -A.im;
        B.im = -B.im;
    } else {
        return mk<T>((T)0, (T)0);
    }
    return mk<T>(A.re - B.im, A.im + B.re); // A + i B
}

template <typename T, int E>
__global__ __launch_bounds__(512) void k_passB(const cpx<T> *__restrict__ Gu, const cpx<T> *__restrict__ Gv,
                        const cpx<T> *__restrict__ Gw, cpx<T> *__restrict__ Huw,
                        cpx<T> *__restrict__ Hvw, FusedDims d, FftPlan p,
                        const cpx<T> *__restrict__ tw, DevScalars *sc, int B, int forward,
                        int track_max)
{
#pragma clang fp contract(off)
    extern __shared__ __align__(16) unsigned char smem[];
    const int b = threadIdx.x % B, t = threadIdx.x / B;
    cpx<T> *lds = reinterpret_cast<cpx<T> *>(smem) + (size_t)b * d.Np;
    const int y0 = blockIdx.x * 2 * B;
    const int ya = y0 + b, yb = y0 + b + B;
    const bool active = yb < d.Np;
    const int h = d.N / 2;

    cpx<T> z3[E];
    if (active) {
#pragma unroll
        for (int s = 0; s < E; ++s) z3[s] = packed_spectrum(Gw, Gw, t + p.T * s, ya, yb, d);
    }
    block_fft<T, E, 1>(z3, lds, p, t, active, tw); // w_a + i w_b
    __syncthreads();

    T vmax = (T)0;
    for (int r = 0; r < 2; ++r) {
        const int y = r == 0 ? ya : yb;
        cpx<T> z[E];
        if (active) {
#pragma unroll
            for (int s = 0; s < E; ++s) z[s] = packed_spectrum(Gu, Gv, t + p.T * s, y, y, d);
        }
        block_fft<T, E, 1>(z, lds, p, t, active, tw); // u + i v
        if (active) {
#pragma unroll
            for (int s = 0; s < E; ++s) {
                const T ww = r == 0 ? z3[s].re : z3[s].im;
                const T m = absv(z[s].re) + absv(z[s].im);
                vmax = (m > vmax || m != m) ? m : vmax; // NaN sticks
                z[s] = mk<T>(z[s].re * ww, z[s].im * ww); // uw + i vw
            }
        }
        if (forward) {
            __syncthreads();
            block_fft<T, E, -1>(z, lds, p, t, active, tw);
            if (active) {
#pragma unroll
                for (int s = 0; s < E; ++s) lds[t + p.T * s] = z[s];
            }
            __syncthreads();
            if (active) {
#pragma unroll
                for (int s = 0; s < E; ++s) {
                    const int k = t + p.T * s;
                    if (k <= h) {
                        cpx<T> P = lds[k];
                        cpx<T> Pc = lds[(d.Np - k) % d.Np];
                        Pc.im = -Pc.im;
                        cpx<T> su = add(P, Pc), df = sub(P, Pc);
                        Huw[(size_t)k * d.Np + y] = mk<T>((T)0.5 * su.re, (T)0.5 * su.im);
                        Hvw[(size_t)k * d.Np + y] = mk<T>((T)0.5 * df.im, (T)-0.5 * df.re);
                    }
                }
            }
        }
        __syncthreads();
    }

    // max of |u|+|v| over the block -> device max (positive doubles order
    // like their bit patterns; NaN bits sort above +Inf). Only the first
    // stage of a step reports it: every stage's dt derives from that value.
    if (!track_max) return;
    double dm = (double)vmax;
    if (dm != dm) dm = __longlong_as_double(0x7ff8000000000000ll);
    for (int off = 32; off > 0; off >>= 1) {
        double o = __shfl_xor(dm, off);
        dm = (o > dm || o != o) ? o : dm;
    }
    if ((threadIdx.x & 63) == 0)
        atomicMax(&sc->vmax_bits, (unsigned long long)__double_as_longlong(dm));
}

// ------------------------------------------------------------- pass C

struct StageArgs {
    int stage;      // 1, 2, 3
    double nu;
    double dt;      // fixed dt
    double cfl;     // > 0: dt from the CFL max of this step
    double dx;      // 2 pi / N
    double norm;    // Np * Np
    double inv_norm; // 1 / norm
};

// The step's dt: the CFL dt of this step's velocity maximum (or the fixed
// dt), clamped so that the step ends at t_end when one is set (t_start is
// the time at the step's start, fixed for its three stages); 0 once the run
// has reached t_end (the stage update then keeps the state as it is).
// Written with selects, not branches on the loaded scalars, so that the
// loads need not complete before the kernel's transforms (measured, FP64
// pass C at 1536^2: 13.5 us; with a branch on t_end 14.6 us).
__device__ __forceinline__ double stage_dt(const StageArgs &a, const DevScalars *sc,
                                           unsigned long long vmax_bits)
{
    const double m = __longlong_as_double((long long)vmax_bits);
    const double dcfl = a.cfl * a.dx / m;
    const double d = (a.cfl > 0.0 && m > 0.0 && m < 1.0e300) ? dcfl : a.dt;
    const double te = sc->t_end, rem = te - sc->t_start;
    const double dc = rem > 1.0e-12 * te ? (d < rem ? d : rem) : 0.0;
    return te > 0.0 ? dc : d;
}
__device__ __forceinline__ double stage_dt(const StageArgs &a, const DevScalars *sc)
{
    return stage_dt(a, sc, sc->vmax_bits);
}

template <typename T, int E>
__global__ void k_passC(const cpx<T> *__restrict__ Huw, const cpx<T> *__restrict__ Hvw,
                        const cpx<T> *Wn, const cpx<T> *Wst, cpx<T> *Wout, FusedDims d, FftPlan p,
                        const cpx<T> *__restrict__ tw, const int *__restrict__ trunc_i1,
                        DevScalars *sc, StageArgs a)
{
#pragma clang fp contract(off)
    extern __shared__ __align__(16) unsigned char smem[];
    cpx<T> *lds = reinterpret_cast<cpx<T> *>(smem);
    const int t = threadIdx.x % p.T, f = threadIdx.x / p.T;
    const int kx = blockIdx.x;
    const cpx<T> *src = (f == 0 ? Huw : Hvw) + (size_t)kx * d.Np;
    cpx<T> v[E];
#pragma unroll
    for (int s = 0; s < E; ++s) v[s] = src[t + p.T * s];
    block_fft<T, E, -1>(v, lds + (size_t)f * d.Np, p, t, true, tw);
    __syncthreads();
#pragma unroll
    for (int s = 0; s < E; ++s) lds[(size_t)f * d.Np + t + p.T * s] = v[s];
    __syncthreads();

    const double dtd = stage_dt(a, sc);
    const T dt = (T)dtd;
    const T qx = (T)kx;
    const size_t col = (size_t)kx * d.KP;
    for (int s = f; s < E; s += 2) {
        const int j = t + p.T * s;
        const int c = compact_row(j, d);
        if (c < 0) continue;
        const int kyi = (j <= d.Np / 2) ? j : j - d.Np;
        const T qy = (T)kyi;
        cpx<T> out = mk<T>((T)0, (T)0);
        if (kx < trunc_i1[j]) {
            cpx<T> P = div_real(lds[j], a.norm);
            cpx<T> Q = div_real(lds[d.Np + j], a.norm);
            cpx<T> du = mk<T>(-(qx * P.im), qx * P.re); // DuDx of (uw)^
            cpx<T> dv = mk<T>(-(qy * Q.im), qy * Q.re); // DuDy of (vw)^
            const T sn = (T)a.nu * (qx * qx + qy * qy);
            cpx<T> ws = Wst[col + c];
            cpx<T> dW = mk<T>((-du.re - dv.re) - sn * ws.re, (-du.im - dv.im) - sn * ws.im);
            cpx<T> w0 = Wn[col + c];
            if (dtd == 0.0) {
                out = w0; // a run at its end time keeps its state
            } else if (a.stage == 1) {
                out = mk<T>(w0.re + dt * dW.re, w0.im + dt * dW.im);
            } else if (a.stage == 2) {
                const T cc = (T)0.25 * dt;
                out = mk<T>(((T)0.75 * w0.re + (T)0.25 * ws.re) + cc * dW.re,
                            ((T)0.75 * w0.im + (T)0.25 * ws.im) + cc * dW.im);
            } else {
                const T cc = dt * (T)2;
                out = mk<T>((w0.re / (T)3 + ((T)2 * ws.re) / (T)3) + (cc * dW.re) / (T)3,
                            (w0.im / (T)3 + ((T)2 * ws.im) / (T)3) + (cc * dW.im) / (T)3);
            }
        }
        Wout[col + c] = out;
    }
    if (a.stage == 3 && kx == 0 && threadIdx.x == 0) {
        sc->dt = dtd;
        sc->time = sc->time + dtd;
        if (!(dtd == 0.0 && sc->t_end > 0.0)) sc->nadv = sc->nadv + 1;
    }
}

// Flags a non-finite value anywhere in a complex array (sc->bad).
template <typename T>
__global__ void k_check_finite(const cpx<T> *__restrict__ a, size_t n, long long per_run,
                               DevScalars *sc)
{
    S2D_GRID_LOOP(k, n)
    {
        const cpx<T> x = a[k];
        if (!((x.re - x.re == (T)0) && (x.im - x.im == (T)0)))
            atomicOr(&sc[(long long)k / per_run].bad, 1u);
    }
}

// ---------------------------------------------------- pack / unpack

// Reference-layout W slot -> compact state (kept modes; the kx = 0 column
// is projected on its Hermitian part, which is what the reference's first
// c2r/r2c round trip of omega does).
template <typename T>
__global__ void k_pack(const cpx<T> *__restrict__ Wref, cpx<T> *__restrict__ Wc, FusedDims d,
                       const int *__restrict__ trunc_i1, long long sref, long long sc)
{
#pragma clang fp contract(off)
    Wref += (size_t)blockIdx.y * sref; // blockIdx.y: one of a batch of runs
    Wc += (size_t)blockIdx.y * sc;
    S2D_GRID_LOOP(idx, (size_t)d.K * d.KP)
    {
        const int kx = (int)(idx / d.KP), c = (int)(idx % d.KP);
        cpx<T> out = mk<T>((T)0, (T)0);
        if (c <= d.N) {
            const int kyi = (c <= d.N / 2) ? c : c - (d.N + 1);
            const int j = kyi >= 0 ? kyi : kyi + d.Np;
            if (kx < trunc_i1[j]) {
                out = Wref[kx + (size_t)j * d.cI];
                if (kx == 0) {
                    const int jm = kyi > 0 ? d.Np - kyi : -kyi;
                    cpx<T> m = Wref[(size_t)jm * d.cI];
                    out = mk<T>((T)0.5 * (out.re + m.re), (T)0.5 * (out.im - m.im));
                }
            }
        }
        Wc[idx] = out;
    }
}

// compact state -> reference-layout W and PSI slots (ComputePsi + Truncate).
template <typename T>
__global__ void k_unpack(const cpx<T> *__restrict__ Wc, cpx<T> *__restrict__ Wref,
                         cpx<T> *__restrict__ Pref, FusedDims d, long long sc, long long sref)
{
#pragma clang fp contract(off)
    Wc += (size_t)blockIdx.y * sc; // blockIdx.y: one of a batch of runs
    Wref += (size_t)blockIdx.y * sref;
    Pref += (size_t)blockIdx.y * sref;
    S2D_GRID_LOOP(idx, (size_t)d.cI * d.Np)
    {
        const int i = (int)(idx % d.cI), j = (int)(idx / d.cI);
        const int kyi = (j <= d.Np / 2) ? j : j - d.Np;
        cpx<T> w = mk<T>((T)0, (T)0), ps = mk<T>((T)0, (T)0);
        if (i < d.K && kyi <= d.N / 2 && kyi >= -d.N / 2) {
            const int c = kyi >= 0 ? kyi : kyi + d.N + 1;
            w = Wc[(size_t)i * d.KP + c];
            if (!(i == 0 && kyi == 0)) {
                const T a = (T)i, b = (T)kyi;
                const T dd = (i == 0) ? b * b : (kyi == 0) ? a * a : a * a + b * b;
                ps = mk<T>(w.re / dd, w.im / dd);
            }
        }
        Wref[idx] = w;
        Pref[idx] = ps;
    }
}

// -------------------------------------------------------- diagnostics

// Per-block FP64 partial sums over the reference-layout vorticity:
// [tke, enstrophy, palinstrophy, max|w|, nonfinite count].
template <typename T>
__global__ void k_diag_partial(const cpx<T> *__restrict__ W, int cI, int ny,
                               const int *__restrict__ kx, const int *__restrict__ ky,
                               int kxMG, double *__restrict__ part)
{
    __shared__ double sh[5][256];
    double a0 = 0, a1 = 0, a2 = 0, a3 = 0, a4 = 0;
    S2D_GRID_LOOP(k, (size_t)cI * ny)
    {
        const int i = (int)(k % cI), j = (int)(k / cI);
        const cpx<T> w = W[k];
        const double re = (double)w.re, im = (double)w.im;
        if (!(re == re && im == im) || fabs(re) > 1.0e300 || fabs(im) > 1.0e300) a4 += 1.0;
        const double q = (double)kx[i] * kx[i] + (double)ky[j] * ky[j];
        const double wt = (kx[i] == 0 || kx[i] == kxMG) ? 0.5 : 1.0; // weight/2 folded
        const double e = re * re + im * im;
        if (q > 0.0) a0 += wt * e / q;
        a1 += wt * e;
        a2 += wt * e * q;
        const double m = sqrt(e);
        a3 = m > a3 ? m : a3;
    }
    sh[0][threadIdx.x] = a0;
    sh[1][threadIdx.x] = a1;
    sh[2][threadIdx.x] = a2;
    sh[3][threadIdx.x] = a3;
    sh[4][threadIdx.x] = a4;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            for (int q = 0; q < 5; ++q) {
                if (q == 3)
                    sh[q][threadIdx.x] = fmax(sh[q][threadIdx.x], sh[q][threadIdx.x + o]);
                else
                    sh[q][threadIdx.x] += sh[q][threadIdx.x + o];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 5; ++q) part[(size_t)blockIdx.x * 5 + q] = sh[q][0];
}

} // namespace s2d
