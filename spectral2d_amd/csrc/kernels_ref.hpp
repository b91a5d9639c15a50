// kernels_ref.hpp -- kernels on the reference layout (one field =
// cISize x nyG complex, x fastest; real view rISize x nyG), used by the
// per-routine entry points (TransformR2C/C2R, DuDx, DuDy, Truncate,
// ComputePsi, ComputeVelocity, ComputeVorticity, ComputeRHS and the literal
// IntegrateOneStep composition).
//
// Pointwise kernels keep the reference's expression order with FP
// contraction off, so they agree bitwise with an FP64 CPU evaluation of
// src/Spectral.F90 / src/TimeIntegration.F90.
#pragma once

#include "fft_block.hpp"

namespace s2d {

template <typename T>
__device__ __forceinline__ cpx<T> div_real(cpx<T> a, double d)
{
    if constexpr (sizeof(T) == 8) {
        return mk<T>(a.re / d, a.im / d);
    } else {
        const T r = (T)(1.0 / d);
        return mk<T>(a.re * r, a.im * r);
    }
}

// ---------------------------------------------------------------- FFT

// Column c2c (along y) of the reference layout: B columns per workgroup,
// lane = b + B*t so a load of element s reads B adjacent columns of one row.
// normalize != 0 divides by norm afterwards (TransformR2C, :761-762).
// blockIdx.y selects a plane ps elements further on (a batch of fields).
template <typename T, int E, int D>
__global__ void k_col_fft(cpx<T> *__restrict__ data, int cI, FftPlan p,
                          const cpx<T> *__restrict__ tw, int B, int normalize,
                          double norm, long long ps)
{
    extern __shared__ __align__(16) unsigned char smem[];
    cpx<T> *lds = reinterpret_cast<cpx<T> *>(smem);
    data += (size_t)blockIdx.y * ps;
    const int b = threadIdx.x % B, t = threadIdx.x / B;
    const int i = blockIdx.x * B + b;
    const bool active = i < cI;
    cpx<T> v[E];
    if (active) {
#pragma unroll
        for (int s = 0; s < E; ++s) v[s] = data[i + (size_t)(t + p.T * s) * cI];
    }
    block_fft<T, E, D>(v, lds + (size_t)b * p.n, p, t, active, tw);
    if (active) {
#pragma unroll
        for (int s = 0; s < E; ++s) {
            cpx<T> x = v[s];
            if (normalize) x = div_real(x, norm);
            data[i + (size_t)(t + p.T * s) * cI] = x;
        }
    }
}

// Row r2c of length n = 2h through a length-h complex transform, in place
// on each row (the real row and its complex row share memory). B rows per
// workgroup, lane = t + T*b. twx[k] = exp(-2 pi i k / n). blockIdx.y
// selects a plane ps elements further on.
template <typename T, int E>
__global__ void k_row_r2c(cpx<T> *__restrict__ data, int cI, int nrows, FftPlan p,
                          const cpx<T> *__restrict__ tw, const cpx<T> *__restrict__ twx,
                          int B, long long ps)
{
    extern __shared__ __align__(16) unsigned char smem[];
    data += (size_t)blockIdx.y * ps;
    const int h = p.n;
    cpx<T> *lds = reinterpret_cast<cpx<T> *>(smem) + (size_t)(threadIdx.x / p.T) * h;
    const int t = threadIdx.x % p.T, b = threadIdx.x / p.T;
    const int j = blockIdx.x * B + b;
    const bool active = j < nrows;
    cpx<T> *row = data + (size_t)j * cI;
    cpx<T> v[E];
    if (active) {
#pragma unroll
        for (int s = 0; s < E; ++s) v[s] = row[t + p.T * s];
    }
    block_fft<T, E, -1>(v, lds, p, t, active, tw);
    if (active) {
#pragma unroll
        for (int s = 0; s < E; ++s) lds[t + p.T * s] = v[s];
    }
    __syncthreads();
    if (active) {
        const int kmax = (t == 0) ? E + 1 : E;
        for (int s = 0; s < kmax; ++s) {
            const int k = (s < E) ? t + p.T * s : h;
            cpx<T> zk = lds[k % h];
            cpx<T> zc = lds[(h - k) % h];
            zc.im = -zc.im;
            cpx<T> Ev = mk<T>((T)0.5 * (zk.re + zc.re), (T)0.5 * (zk.im + zc.im));
            cpx<T> Od = mk<T>((T)0.5 * (zk.im - zc.im), (T)-0.5 * (zk.re - zc.re));
            cpx<T> w = twx[k];
            cpx<T> tt = mul(w, Od);
            row[k] = mk<T>(Ev.re + tt.re, Ev.im + tt.im);
        }
    }
}

// Row c2r of length n = 2h (unnormalized), in place. Only Re of X_0 and
// X_h is read, as for a c2r transform that assumes a real result.
template <typename T, int E>
__global__ void k_row_c2r(cpx<T> *__restrict__ data, int cI, int nrows, FftPlan p,
                          const cpx<T> *__restrict__ tw, const cpx<T> *__restrict__ twx,
                          int B, long long ps)
{
    extern __shared__ __align__(16) unsigned char smem[];
    data += (size_t)blockIdx.y * ps;
    const int h = p.n;
    cpx<T> *lds = reinterpret_cast<cpx<T> *>(smem) + (size_t)(threadIdx.x / p.T) * h;
    const int t = threadIdx.x % p.T, b = threadIdx.x / p.T;
    const int j = blockIdx.x * B + b;
    const bool active = j < nrows;
    cpx<T> *row = data + (size_t)j * cI;
    cpx<T> v[E];
    if (active) {
#pragma unroll
        for (int s = 0; s < E; ++s) {
            const int k = t + p.T * s;
            cpx<T> a = row[k];
            cpx<T> bb = row[h - k];
            if (k == 0) {
                a.im = (T)0;
                bb.im = (T)0;
            }
            bb.im = -bb.im;
            cpx<T> sm = add(a, bb);
            cpx<T> df = sub(a, bb);
            cpx<T> tt = mulc(df, twx[k]); // df * conj(W_n^k)
            v[s] = mk<T>(sm.re - tt.im, sm.im + tt.re);
        }
    }
    __syncthreads(); // every read of the row precedes every write
    block_fft<T, E, 1>(v, lds, p, t, active, tw);
    if (active) {
#pragma unroll
        for (int s = 0; s < E; ++s) row[t + p.T * s] = v[s];
        if (t == 0) row[h] = mk<T>((T)0, (T)0); // the two padding reals
    }
}

// ---------------------------------------------------------- pointwise

#define S2D_GRID_LOOP(idx, total) \
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < (total); \
         idx += (size_t)gridDim.x * blockDim.x)

template <typename T>
__global__ void k_fill(cpx<T> *__restrict__ d, size_t n)
{
    S2D_GRID_LOOP(k, n) d[k] = mk<T>((T)0, (T)0);
}

template <typename T>
__global__ void k_copy(const cpx<T> *__restrict__ s, cpx<T> *__restrict__ d, size_t n)
{
    S2D_GRID_LOOP(k, n) d[k] = s[k];
}

// DuDx, src/Spectral.F90:678-692: out = II*kx*in, then the kx = kxMG column
// is zeroed. (0,kx)*(a,b) = (-(kx*b), kx*a).
template <typename T>
__global__ void k_dudx(const cpx<T> *in, cpx<T> *out, int cI, int ny, const int *__restrict__ kx,
                       int xMaskInd)
{
#pragma clang fp contract(off)
    S2D_GRID_LOOP(k, (size_t)cI * ny)
    {
        const int i = (int)(k % cI);
        const T q = (T)kx[i];
        cpx<T> a = in[k];
        cpx<T> r = mk<T>(-(q * a.im), q * a.re);
        if (i == xMaskInd) r = mk<T>((T)0, (T)0);
        out[k] = r;
    }
}

// DuDy, src/Spectral.F90:724-737.
template <typename T>
__global__ void k_dudy(const cpx<T> *in, cpx<T> *out, int cI, int ny, const int *__restrict__ ky,
                       int yMaskInd)
{
#pragma clang fp contract(off)
    S2D_GRID_LOOP(k, (size_t)cI * ny)
    {
        const int j = (int)(k / cI);
        const T q = (T)ky[j];
        cpx<T> a = in[k];
        cpx<T> r = mk<T>(-(q * a.im), q * a.re);
        if (j == yMaskInd) r = mk<T>((T)0, (T)0);
        out[k] = r;
    }
}

// Truncate, src/Spectral.F90:799-801.
template <typename T>
__global__ void k_truncate(cpx<T> *d, int cI, int ny, const int *__restrict__ i1)
{
    S2D_GRID_LOOP(k, (size_t)cI * ny)
    {
        const int i = (int)(k % cI), j = (int)(k / cI);
        if (i >= i1[j]) d[k] = mk<T>((T)0, (T)0);
    }
}

// ComputePsi, src/Spectral.F90:528-559: Qp = Qw/(kx^2+ky^2), Qp(0,0) = 0.
// Row 0 is ky = 0 (one process), so the four chunks reduce to this.
template <typename T>
__global__ void k_psi(const cpx<T> *w, cpx<T> *psi, int cI, int ny, const int *__restrict__ kx,
                      const int *__restrict__ ky)
{
#pragma clang fp contract(off)
    S2D_GRID_LOOP(k, (size_t)cI * ny)
    {
        const int i = (int)(k % cI), j = (int)(k / cI);
        if (i == 0 && j == 0) {
            psi[k] = mk<T>((T)0, (T)0);
        } else {
            const T a = (T)kx[i], b = (T)ky[j];
            const T d = (i == 0) ? b * b : (j == 0) ? a * a : a * a + b * b;
            cpx<T> q = w[k];
            psi[k] = mk<T>(q.re / d, q.im / d);
        }
    }
}

// ComputeVelocity (spectral part), src/Spectral.F90:591-603:
// u = II*ky*psi (masked row), v = -(II*kx*psi) (masked column).
template <typename T>
__global__ void k_velocity(const cpx<T> *psi, cpx<T> *u, cpx<T> *v, int cI, int ny,
                           const int *__restrict__ kx, const int *__restrict__ ky,
                           int xMaskInd, int yMaskInd)
{
#pragma clang fp contract(off)
    S2D_GRID_LOOP(k, (size_t)cI * ny)
    {
        const int i = (int)(k % cI), j = (int)(k / cI);
        cpx<T> a = psi[k];
        const T qy = (T)ky[j], qx = (T)kx[i];
        cpx<T> ru = mk<T>(-(qy * a.im), qy * a.re);
        if (j == yMaskInd) ru = mk<T>((T)0, (T)0);
        cpx<T> rv = mk<T>(-(qx * a.im), qx * a.re);
        if (i == xMaskInd) rv = mk<T>((T)0, (T)0);
        u[k] = ru;
        v[k] = mk<T>((T)-1 * rv.re, (T)-1 * rv.im);
    }
}

// ComputeVorticity, src/Spectral.F90:636-646 (v <- DuDx v, u <- DuDy u,
// w = v - u).
template <typename T>
__global__ void k_vorticity(cpx<T> *u, cpx<T> *v, cpx<T> *w, int cI, int ny,
                            const int *__restrict__ kx, const int *__restrict__ ky,
                            int xMaskInd, int yMaskInd)
{
#pragma clang fp contract(off)
    S2D_GRID_LOOP(k, (size_t)cI * ny)
    {
        const int i = (int)(k % cI), j = (int)(k / cI);
        const T qy = (T)ky[j], qx = (T)kx[i];
        cpx<T> a = v[k], b = u[k];
        cpx<T> dv = mk<T>(-(qx * a.im), qx * a.re);
        if (i == xMaskInd) dv = mk<T>((T)0, (T)0);
        cpx<T> du = mk<T>(-(qy * b.im), qy * b.re);
        if (j == yMaskInd) du = mk<T>((T)0, (T)0);
        v[k] = dv;
        u[k] = du;
        w[k] = mk<T>(dv.re - du.re, dv.im - du.im);
    }
}

// Physical-space products, src/Spectral.F90:442-447 (i < nxG of each row).
template <typename T>
__global__ void k_product(T *u, T *v, const T *w, int nxG, int rI, int ny)
{
#pragma clang fp contract(off)
    S2D_GRID_LOOP(k, (size_t)nxG * ny)
    {
        const size_t i = k % nxG, j = k / nxG;
        const size_t o = i + j * rI;
        const T ww = w[o];
        u[o] = u[o] * ww;
        v[o] = v[o] * ww;
    }
}

// RHS assembly, src/Spectral.F90:471-481: dW = -u - v - nu*(kx^2+ky^2)*w.
template <typename T>
__global__ void k_rhs(const cpx<T> *u, const cpx<T> *v, const cpx<T> *w, cpx<T> *dW, int cI,
                      int ny, const int *__restrict__ kx, const int *__restrict__ ky, double nu)
{
#pragma clang fp contract(off)
    S2D_GRID_LOOP(k, (size_t)cI * ny)
    {
        const int i = (int)(k % cI), j = (int)(k / cI);
        const T a = (T)kx[i], b = (T)ky[j];
        const T s = (T)nu * (a * a + b * b);
        cpx<T> x = u[k], y = v[k], z = w[k];
        dW[k] = mk<T>((-x.re - y.re) - s * z.re, (-x.im - y.im) - s * z.im);
    }
}

// Stage combinations, src/TimeIntegration.F90:156-160, 175-181, 196-202.
template <typename T>
__global__ void k_rk3_stage(int stage, cpx<T> *w, cpx<T> *wS, const cpx<T> *dW, size_t n,
                            double dt_)
{
#pragma clang fp contract(off)
    const T dt = (T)dt_;
    S2D_GRID_LOOP(k, n)
    {
        cpx<T> a = w[k], b = wS[k], d = dW[k];
        if (stage == 1) {
            wS[k] = mk<T>(a.re + dt * d.re, a.im + dt * d.im);
        } else if (stage == 2) {
            const T c = (T)0.25 * dt;
            wS[k] = mk<T>(((T)0.75 * a.re + (T)0.25 * b.re) + c * d.re,
                          ((T)0.75 * a.im + (T)0.25 * b.im) + c * d.im);
        } else {
            const T c = dt * (T)2;
            w[k] = mk<T>((a.re / (T)3 + ((T)2 * b.re) / (T)3) + (c * d.re) / (T)3,
                         (a.im / (T)3 + ((T)2 * b.im) / (T)3) + (c * d.im) / (T)3);
        }
    }
}

} // namespace s2d
