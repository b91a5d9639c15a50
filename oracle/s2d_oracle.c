/*
 * s2d_oracle.c -- CPU FP64 oracle for the Spectral2D TVD-RK3 time step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product). See
 * s2d_oracle.h for the layout contract and the reference files restated.
 *
 * Arithmetic follows the reference's expression forms so that pointwise
 * results can be compared bitwise: gfortran evaluates mixed real*complex and
 * complex/real component-wise, left to right, without FMA contraction
 * (build with -ffp-contract=off). All FFT arithmetic is this file's own
 * recursive decimation-in-time mixed-radix (4, 2, 3) transform; the
 * reference used FFTW 3.3.3, which is not available.
 */
#define _GNU_SOURCE /* sincos */
#include "s2d_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double re, im; } cplx;

/* ------------------------------------------------------------------ FFT */

typedef struct fft_plan {
    int n;
    cplx *tw; /* tw[k] = exp(-2 pi i k / n), k = 0..n-1 */
} fft_plan;

static int is_2a3b(int n)
{
    if (n < 1) return 0;
    while (n % 2 == 0) n /= 2;
    while (n % 3 == 0) n /= 3;
    return n == 1;
}

static void plan_init(fft_plan *p, int n)
{
    p->n = n;
    p->tw = (cplx *)malloc(sizeof(cplx) * (size_t)(n > 0 ? n : 1));
    /* Library sin/cos in long double, then rounded: ~0.5 ulp twiddles. */
    const long double two_pi = 6.283185307179586476925286766559005768L;
    for (int k = 0; k < n; ++k) {
        long double a = two_pi * (long double)k / (long double)n;
        p->tw[k].re = (double)cosl(a);
        p->tw[k].im = (double)(-sinl(a));
    }
}

static void plan_free(fft_plan *p)
{
    free(p->tw);
    p->tw = NULL;
    p->n = 0;
}

static inline cplx cmul(cplx a, cplx b)
{
    cplx r;
    r.re = a.re * b.re - a.im * b.im;
    r.im = a.re * b.im + a.im * b.re;
    return r;
}

/*
 * Recursive DIT: X[k + s m] = sum_q W_n^{q (k + s m)} Y_q[k], n = r m, with
 * Y_q = DFT_m(x[q + r t]). Sub-transforms land in out[q m .. q m + m - 1];
 * the combine step is in place over out. tw is the plan table of length N;
 * W_n^e = tw[e * (N / n)]. sign = -1 forward, +1 inverse (conjugate).
 */
static void fft_rec(const cplx *in, int istride, cplx *out, int n,
                    const fft_plan *p, int sign)
{
    if (n == 1) {
        out[0] = in[0];
        return;
    }
    int r = (n % 4 == 0) ? 4 : (n % 2 == 0) ? 2 : 3;
    int m = n / r;
    for (int q = 0; q < r; ++q)
        fft_rec(in + (size_t)q * istride, istride * r, out + (size_t)q * m,
                m, p, sign);
    int ts = p->n / n;
    for (int k = 0; k < m; ++k) {
        cplx a[4];
        for (int q = 0; q < r; ++q) {
            cplx w = p->tw[(size_t)q * k * ts];
            if (sign > 0) w.im = -w.im;
            a[q] = (q == 0) ? out[k] : cmul(out[(size_t)q * m + k], w);
        }
        if (r == 2) {
            out[k].re = a[0].re + a[1].re;
            out[k].im = a[0].im + a[1].im;
            out[k + m].re = a[0].re - a[1].re;
            out[k + m].im = a[0].im - a[1].im;
        } else if (r == 4) {
            /* forward: X1 = a0 - i a1 - a2 + i a3; inverse conjugates i */
            double s = (sign < 0) ? 1.0 : -1.0;
            cplx t0 = {a[0].re + a[2].re, a[0].im + a[2].im};
            cplx t1 = {a[0].re - a[2].re, a[0].im - a[2].im};
            cplx t2 = {a[1].re + a[3].re, a[1].im + a[3].im};
            cplx t3 = {a[1].re - a[3].re, a[1].im - a[3].im};
            /* -i*s*t3 = (s t3.im, -s t3.re) */
            out[k].re = t0.re + t2.re;
            out[k].im = t0.im + t2.im;
            out[k + m].re = t1.re + s * t3.im;
            out[k + m].im = t1.im - s * t3.re;
            out[k + 2 * m].re = t0.re - t2.re;
            out[k + 2 * m].im = t0.im - t2.im;
            out[k + 3 * m].re = t1.re - s * t3.im;
            out[k + 3 * m].im = t1.im + s * t3.re;
        } else {
            /* radix 3: X1,2 = a0 - (a1+a2)/2 -/+ i*sign'*(sqrt3/2)(a1-a2) */
            const double h = 0.86602540378443864676372317075294;
            double s = (sign < 0) ? 1.0 : -1.0;
            cplx t1 = {a[1].re + a[2].re, a[1].im + a[2].im};
            cplx t2 = {a[0].re - 0.5 * t1.re, a[0].im - 0.5 * t1.im};
            cplx t3 = {s * h * (a[1].re - a[2].re), s * h * (a[1].im - a[2].im)};
            out[k].re = a[0].re + t1.re;
            out[k].im = a[0].im + t1.im;
            /* -i*t3 = (t3.im, -t3.re) */
            out[k + m].re = t2.re + t3.im;
            out[k + m].im = t2.im - t3.re;
            out[k + 2 * m].re = t2.re - t3.im;
            out[k + 2 * m].im = t2.im + t3.re;
        }
    }
}

int or_fft_c2c(int n, int sign, const double *in, double *out)
{
    if (!is_2a3b(n)) return -1;
    fft_plan p;
    plan_init(&p, n);
    fft_rec((const cplx *)in, 1, (cplx *)out, n, &p, sign);
    plan_free(&p);
    return 0;
}

/* Real-to-complex of one row of even length n through a length-n/2 complex
 * FFT: z_j = x_{2j} + i x_{2j+1}, X_k = E_k + W_n^k O_k with
 * E_k = (Z_k + conj Z_{n/2-k})/2, O_k = (Z_k - conj Z_{n/2-k})/(2i). */
static void row_r2c(const fft_plan *ph, const cplx *wn, int n,
                    const double *x, cplx *X, cplx *zbuf, cplx *Zbuf)
{
    int h = n / 2;
    for (int j = 0; j < h; ++j) {
        zbuf[j].re = x[2 * j];
        zbuf[j].im = x[2 * j + 1];
    }
    fft_rec(zbuf, 1, Zbuf, h, ph, -1);
    for (int k = 0; k <= h; ++k) {
        cplx zk = Zbuf[k % h];
        cplx zc = Zbuf[(h - k) % h];
        zc.im = -zc.im;
        cplx E = {0.5 * (zk.re + zc.re), 0.5 * (zk.im + zc.im)};
        /* (zk - zc)/(2i) = (d.im/2, -d.re/2) */
        cplx O = {0.5 * (zk.im - zc.im), -0.5 * (zk.re - zc.re)};
        cplx t = cmul(wn[k], O);
        X[k].re = E.re + t.re;
        X[k].im = E.im + t.im;
    }
}

/* Complex-to-real of one row (unnormalized). Only the real parts of X_0 and
 * X_{n/2} are read, as a c2r transform that assumes a real output does.
 * Z_k = (X_k + conj X_{h-k}) + i (X_k - conj X_{h-k}) conj(W_n^k),
 * z = IDFT_h(Z), x_{2j} = Re z_j, x_{2j+1} = Im z_j. */
static void row_c2r(const fft_plan *ph, const cplx *wn, int n, const cplx *Xin,
                    double *x, cplx *Zbuf, cplx *zbuf)
{
    int h = n / 2;
    for (int k = 0; k < h; ++k) {
        cplx a = Xin[k];
        cplx b = Xin[h - k];
        if (k == 0) { a.im = 0.0; b.im = 0.0; }
        b.im = -b.im;
        cplx s = {a.re + b.re, a.im + b.im};
        cplx d = {a.re - b.re, a.im - b.im};
        cplx wc = wn[k];
        wc.im = -wc.im;
        cplx t = cmul(d, wc);
        /* s + i t */
        Zbuf[k].re = s.re - t.im;
        Zbuf[k].im = s.im + t.re;
    }
    fft_rec(Zbuf, 1, zbuf, h, ph, +1);
    for (int j = 0; j < h; ++j) {
        x[2 * j] = zbuf[j].re;
        x[2 * j + 1] = zbuf[j].im;
    }
}

/* ------------------------------------------------------------ context */

struct or_ctx {
    /* SetupMask, src/Spectral.F90:67-73 */
    int xMaskBool, yMaskBool, xMaskInd, yMaskInd;
    /* SetupTruncation, src/Spectral.F90:62 */
    int nTrunc;
    int *trunc_i1;
    /* SetupPlans, src/Spectral.F90:78-80 */
    int nxG, nyG;
    fft_plan px; /* length nxG/2 complex, rows */
    fft_plan py; /* length nyG complex, columns */
    cplx *wn;    /* exp(-2 pi i k / nxG), k = 0..nxG/2 */
    /* TimeIntegration_m scratch, src/TimeIntegration.F90:47-64 */
    int scheme;
    size_t nC;
    double *dW, *wS, *psiS;
};

static int g_threads = 1;

int or_set_threads(int n)
{
#ifdef _OPENMP
    if (n < 1) n = 1;
    g_threads = n;
#else
    (void)n;
    g_threads = 1;
#endif
    return g_threads;
}

or_ctx_t *or_create(void)
{
    or_ctx_t *c = (or_ctx_t *)calloc(1, sizeof(or_ctx_t));
    c->xMaskInd = c->yMaskInd = -1;
    return c;
}

void or_destroy(or_ctx_t *c)
{
    if (!c) return;
    free(c->trunc_i1);
    plan_free(&c->px);
    plan_free(&c->py);
    free(c->wn);
    free(c->dW);
    free(c->wS);
    free(c->psiS);
    free(c);
}

/* ------------------------------------------------------- grid / setup */

int or_get_grid_size(int nx, int ny, int dealias, or_grid_t *g)
{
    memset(g, 0, sizeof(*g));
    g->nx = nx;
    g->ny = ny;
    g->dealias = dealias;
    switch (dealias) {
    case OR_NO_DEALIASING: /* Spectral.F90:141-159 */
        g->nxG = nx;
        g->nyG = ny;
        g->kxLG = 0;
        g->kyLG = -g->nyG / 2 + 1;
        g->kxMG = g->nxG / 2;
        g->kyMG = g->nyG / 2;
        g->kxLU = 0;
        g->kyLU = -ny / 2 + 1;
        g->kxMU = nx / 2;
        g->kyMU = ny / 2;
        g->kMax = (double)g->kxMU;
        break;
    case OR_DEALIAS_3_2_RULE: /* Spectral.F90:160-179 */
        g->nxG = 3 * nx / 2;
        g->nyG = 3 * ny / 2;
        g->kxLG = 0;
        g->kyLG = -g->nyG / 2 + 1;
        g->kxMG = g->nxG / 2;
        g->kyMG = g->nyG / 2;
        g->kxLU = 0;
        g->kyLU = -ny / 2 + 1;
        g->kxMU = nx / 2;
        g->kyMU = ny / 2;
        g->kMax = (double)g->kxMU;
        break;
    default: /* Spectral.F90:180-185: 2/3 rule and grid shift are empty */
        return -1;
    }
    g->rISize = 2 * (g->nxG / 2 + 1); /* Spectral.F90:189-190 */
    g->cISize = g->nxG / 2 + 1;
    return 0;
}

void or_wavenumbers(const or_grid_t *g, int *kxG, int *kyG)
{
    for (int i = 0; i < g->cISize; ++i) kxG[i] = i; /* :204-207 */
    for (int i = 0; i < g->nyG; ++i)                /* :210-218 */
        kyG[i] = (i > g->nyG / 2) ? i - g->nyG : i;
}

void or_setup_mask(or_ctx_t *c, int cISize, int nyP, const int *kxP,
                   const int *kyP, int kxMG, int kyMG)
{
    c->xMaskBool = 0;
    c->xMaskInd = -1;
    for (int i = 0; i < cISize; ++i)
        if (kxP[i] == kxMG) { c->xMaskBool = 1; c->xMaskInd = i; break; }
    c->yMaskBool = 0;
    c->yMaskInd = -1;
    for (int j = 0; j < nyP; ++j)
        if (kyP[j] == kyMG) { c->yMaskBool = 1; c->yMaskInd = j; break; }
}

void or_get_mask(const or_ctx_t *c, int *xb, int *xi, int *yb, int *yi)
{
    *xb = c->xMaskBool;
    *xi = c->xMaskInd;
    *yb = c->yMaskBool;
    *yi = c->yMaskInd;
}

void or_setup_truncation(or_ctx_t *c, int cISize, int nyP, double kMax,
                         const int *kxP, const int *kyP)
{
    const double eps = 1.0e-10; /* Spectral.F90:285 */
    free(c->trunc_i1);
    c->nTrunc = nyP;
    c->trunc_i1 = (int *)malloc(sizeof(int) * (size_t)(nyP > 0 ? nyP : 1));
    for (int j = 0; j < nyP; ++j) {
        double ky = (double)kyP[j];
        /* A row with no |k| beyond kMax is left unset by the reference
         * (undefined); here it truncates nothing. */
        c->trunc_i1[j] = cISize;
        for (int i = 0; i < cISize; ++i) {
            double kx = (double)kxP[i];
            double kMag = sqrt(kx * kx + ky * ky);
            if (kMag > kMax + eps) { c->trunc_i1[j] = i; break; }
        }
    }
}

void or_get_truncation(const or_ctx_t *c, int *i1)
{
    for (int j = 0; j < c->nTrunc; ++j) i1[j] = c->trunc_i1[j];
}

void or_setup_plans(or_ctx_t *c, int nxG, int nyG, int nyP, int rISize,
                    int cISize, double *data)
{
    (void)rISize;
    plan_free(&c->px);
    plan_free(&c->py);
    free(c->wn);
    c->nxG = nxG;
    c->nyG = nyG;
    plan_init(&c->px, nxG / 2);
    plan_init(&c->py, nyG);
    c->wn = (cplx *)malloc(sizeof(cplx) * (size_t)(nxG / 2 + 1));
    const long double two_pi = 6.283185307179586476925286766559005768L;
    for (int k = 0; k <= nxG / 2; ++k) {
        long double a = two_pi * (long double)k / (long double)nxG;
        c->wn[k].re = (double)cosl(a);
        c->wn[k].im = (double)(-sinl(a));
    }
    if (data) /* Spectral.F90:352 */
        memset(data, 0, sizeof(double) * 2 * (size_t)cISize * (size_t)nyP);
}

/* --------------------------------------------------------- transforms */

static void columns_c2c(or_ctx_t *c, int cISize, int nyP, cplx *d, int sign)
{
    int n = nyP;
#pragma omp parallel num_threads(g_threads)
    {
        cplx *a = (cplx *)malloc(sizeof(cplx) * (size_t)n);
        cplx *b = (cplx *)malloc(sizeof(cplx) * (size_t)n);
#pragma omp for schedule(static)
        for (int i = 0; i < cISize; ++i) {
            for (int j = 0; j < n; ++j) a[j] = d[i + (size_t)j * cISize];
            fft_rec(a, 1, b, n, &c->py, sign);
            for (int j = 0; j < n; ++j) d[i + (size_t)j * cISize] = b[j];
        }
        free(a);
        free(b);
    }
}

void or_transform_r2c(or_ctx_t *c, int nxG, int nyG, int nyP, int rISize,
                      int cISize, double *data)
{
    cplx *d = (cplx *)data;
    int h = nxG / 2;
#pragma omp parallel num_threads(g_threads)
    {
        double *x = (double *)malloc(sizeof(double) * (size_t)nxG);
        cplx *z = (cplx *)malloc(sizeof(cplx) * (size_t)h);
        cplx *Z = (cplx *)malloc(sizeof(cplx) * (size_t)h);
        cplx *X = (cplx *)malloc(sizeof(cplx) * (size_t)(h + 1));
#pragma omp for schedule(static)
        for (int j = 0; j < nyP; ++j) {
            memcpy(x, data + (size_t)j * rISize, sizeof(double) * (size_t)nxG);
            row_r2c(&c->px, c->wn, nxG, x, X, z, Z);
            memcpy(d + (size_t)j * cISize, X, sizeof(cplx) * (size_t)cISize);
        }
        free(x);
        free(z);
        free(Z);
        free(X);
    }
    columns_c2c(c, cISize, nyP, d, -1);
    /* Spectral.F90:761-762: cData = cData/norm, component-wise. */
    double norm = (double)(nxG * nyG);
    size_t nc = (size_t)cISize * nyP;
    for (size_t k = 0; k < nc; ++k) {
        d[k].re = d[k].re / norm;
        d[k].im = d[k].im / norm;
    }
}

void or_transform_c2r(or_ctx_t *c, int rISize, int cISize, int nyP,
                      double *data)
{
    cplx *d = (cplx *)data;
    int nxG = c->nxG;
    int h = nxG / 2;
    columns_c2c(c, cISize, nyP, d, +1);
#pragma omp parallel num_threads(g_threads)
    {
        cplx *X = (cplx *)malloc(sizeof(cplx) * (size_t)(h + 1));
        cplx *Z = (cplx *)malloc(sizeof(cplx) * (size_t)h);
        cplx *z = (cplx *)malloc(sizeof(cplx) * (size_t)h);
#pragma omp for schedule(static)
        for (int j = 0; j < nyP; ++j) {
            memcpy(X, d + (size_t)j * cISize, sizeof(cplx) * (size_t)(h + 1));
            double *x = data + (size_t)j * rISize;
            row_c2r(&c->px, c->wn, nxG, X, x, Z, z);
            for (int i = nxG; i < rISize; ++i) x[i] = 0.0; /* padding */
        }
        free(X);
        free(Z);
        free(z);
    }
}

/* --------------------------------------------------- pointwise kernels */

void or_dudx(const or_ctx_t *c, int cISize, int nyP, const int *kxP,
             const int *kyP, const double *in, double *out)
{
    (void)kyP;
    for (int j = 0; j < nyP; ++j)
        for (int i = 0; i < cISize; ++i) {
            size_t k = 2 * ((size_t)i + (size_t)j * cISize);
            double kx = (double)kxP[i];
            /* II*kx*sigIn: (0,kx)*(a,b) = (-(kx*b), kx*a) */
            double a = in[k], b = in[k + 1];
            out[k] = -(kx * b);
            out[k + 1] = kx * a;
        }
    if (c->xMaskBool) /* Spectral.F90:690-692 */
        for (int j = 0; j < nyP; ++j) {
            size_t k = 2 * ((size_t)c->xMaskInd + (size_t)j * cISize);
            out[k] = 0.0;
            out[k + 1] = 0.0;
        }
}

void or_dudy(const or_ctx_t *c, int cISize, int nyP, const int *kxP,
             const int *kyP, const double *in, double *out)
{
    (void)kxP;
    for (int j = 0; j < nyP; ++j) {
        double ky = (double)kyP[j];
        for (int i = 0; i < cISize; ++i) {
            size_t k = 2 * ((size_t)i + (size_t)j * cISize);
            double a = in[k], b = in[k + 1];
            out[k] = -(ky * b);
            out[k + 1] = ky * a;
        }
    }
    if (c->yMaskBool) /* Spectral.F90:735-737 */
        for (int i = 0; i < cISize; ++i) {
            size_t k = 2 * ((size_t)i + (size_t)c->yMaskInd * cISize);
            out[k] = 0.0;
            out[k + 1] = 0.0;
        }
}

void or_truncate(const or_ctx_t *c, int cISize, int nyP, double *data)
{
    for (int j = 0; j < nyP; ++j)
        for (int i = c->trunc_i1[j]; i < cISize; ++i) {
            size_t k = 2 * ((size_t)i + (size_t)j * cISize);
            data[k] = 0.0;
            data[k + 1] = 0.0;
        }
}

void or_compute_psi(int cISize, int nyP, const int *kxP, const int *kyP,
                    const double *Qw, double *Qp)
{
    /* Spectral.F90:528-559, four chunks; row 0 is ky = 0 on one process. */
    for (int j = 1; j < nyP; ++j) {
        double ky = (double)kyP[j];
        for (int i = 1; i < cISize; ++i) {
            double kx = (double)kxP[i];
            double d = kx * kx + ky * ky;
            size_t k = 2 * ((size_t)i + (size_t)j * cISize);
            Qp[k] = Qw[k] / d;
            Qp[k + 1] = Qw[k + 1] / d;
        }
    }
    for (int i = 1; i < cISize; ++i) {
        double kx = (double)kxP[i];
        double d = kx * kx;
        size_t k = 2 * (size_t)i;
        Qp[k] = Qw[k] / d;
        Qp[k + 1] = Qw[k + 1] / d;
    }
    for (int j = 1; j < nyP; ++j) {
        double ky = (double)kyP[j];
        double d = ky * ky;
        size_t k = 2 * ((size_t)j * cISize);
        Qp[k] = Qw[k] / d;
        Qp[k + 1] = Qw[k + 1] / d;
    }
    Qp[0] = 0.0;
    Qp[1] = 0.0;
}

void or_compute_velocity(or_ctx_t *c, int nxP, int nyP, int rISize,
                         int cISize, const int *kxP, const int *kyP,
                         double *u, double *v, const double *psi, int trans)
{
    (void)nxP;
    size_t nd = 2 * (size_t)cISize * nyP;
    memset(u, 0, sizeof(double) * nd); /* Spectral.F90:591 */
    or_dudy(c, cISize, nyP, kxP, kyP, psi, u);
    memset(v, 0, sizeof(double) * nd); /* :599 */
    or_dudx(c, cISize, nyP, kxP, kyP, psi, v);
    for (size_t k = 0; k < nd; ++k) v[k] = -1.0 * v[k]; /* :603 */
    if (trans) {
        or_transform_c2r(c, rISize, cISize, nyP, u);
        or_transform_c2r(c, rISize, cISize, nyP, v);
    }
}

void or_compute_vorticity(const or_ctx_t *c, int nxP, int nyP, int cISize,
                          const int *kxP, const int *kyP, double *u,
                          double *v, double *w)
{
    (void)nxP;
    or_dudx(c, cISize, nyP, kxP, kyP, v, v); /* :636 */
    or_dudy(c, cISize, nyP, kxP, kyP, u, u); /* :639 */
    size_t nd = 2 * (size_t)cISize * nyP;
    for (size_t k = 0; k < nd; ++k) w[k] = v[k] - u[k]; /* :642-646 */
}

void or_compute_rhs(or_ctx_t *c, int nxG, int nyG, int nxP, int nyP,
                    int rISize, int cISize, const int *kxP, const int *kyP,
                    double nu, double *u, double *v, double *w,
                    const double *psi, double *dW)
{
    or_compute_velocity(c, nxP, nyP, rISize, cISize, kxP, kyP, u, v, psi, 0);
    or_transform_c2r(c, rISize, cISize, nyP, u); /* :433-435 */
    or_transform_c2r(c, rISize, cISize, nyP, v);
    or_transform_c2r(c, rISize, cISize, nyP, w);
    for (int j = 0; j < nyP; ++j) /* :442-447 */
        for (int i = 0; i < nxP; ++i) {
            size_t k = (size_t)i + (size_t)j * rISize;
            u[k] = u[k] * w[k];
            v[k] = v[k] * w[k];
        }
    or_transform_r2c(c, nxG, nyG, nyP, rISize, cISize, u); /* :451-453 */
    or_transform_r2c(c, nxG, nyG, nyP, rISize, cISize, v);
    or_transform_r2c(c, nxG, nyG, nyP, rISize, cISize, w);
    or_truncate(c, cISize, nyP, u); /* :457-458 */
    or_truncate(c, cISize, nyP, v);
    or_dudx(c, cISize, nyP, kxP, kyP, u, u); /* :464-465 */
    or_dudy(c, cISize, nyP, kxP, kyP, v, v);
    for (int j = 0; j < nyP; ++j) { /* :471-481 */
        double ky = (double)kyP[j];
        for (int i = 0; i < cISize; ++i) {
            double kx = (double)kxP[i];
            double s = nu * (kx * kx + ky * ky);
            size_t k = 2 * ((size_t)i + (size_t)j * cISize);
            dW[k] = (-u[k] - v[k]) - s * w[k];
            dW[k + 1] = (-u[k + 1] - v[k + 1]) - s * w[k + 1];
        }
    }
}

/* ------------------------------------------------------ time stepping */

int or_time_integration_setup(or_ctx_t *c, int scheme, int rISize,
                              int cISize, int nyP, int allocLocal)
{
    (void)rISize;
    if (scheme != OR_RK3_TVD_SHU) return -1;
    c->scheme = scheme;
    free(c->dW);
    free(c->wS);
    free(c->psiS);
    size_t nAlloc = (size_t)(allocLocal > cISize * nyP ? allocLocal
                                                        : cISize * nyP);
    c->nC = (size_t)cISize * nyP;
    c->dW = (double *)calloc(2 * c->nC, sizeof(double));
    c->wS = (double *)calloc(2 * nAlloc, sizeof(double));
    c->psiS = (double *)calloc(2 * nAlloc, sizeof(double));
    return 0;
}

void or_integrate_one_step(or_ctx_t *c, int nxG, int nyG, int nxP, int nyP,
                           int rISize, int cISize, const int *kxP,
                           const int *kyP, double nu, double dt, double *u,
                           double *v, double *w, double *psi)
{
    size_t nd = 2 * (size_t)cISize * nyP;
    double *wS = c->wS, *psiS = c->psiS, *dW = c->dW;
    memset(wS, 0, sizeof(double) * nd); /* TimeIntegration.F90:147-149 */
    memset(psiS, 0, sizeof(double) * nd);
    memset(dW, 0, sizeof(double) * nd);
    /* stage 1, :152-167 */
    or_compute_rhs(c, nxG, nyG, nxP, nyP, rISize, cISize, kxP, kyP, nu, u, v,
                   w, psi, dW);
    for (size_t k = 0; k < nd; ++k) wS[k] = w[k] + dt * dW[k];
    or_compute_psi(cISize, nyP, kxP, kyP, wS, psiS);
    or_truncate(c, cISize, nyP, wS);
    or_truncate(c, cISize, nyP, psiS);
    /* stage 2, :170-188 */
    memset(dW, 0, sizeof(double) * nd);
    or_compute_rhs(c, nxG, nyG, nxP, nyP, rISize, cISize, kxP, kyP, nu, u, v,
                   wS, psiS, dW);
    for (size_t k = 0; k < nd; ++k)
        wS[k] = (0.75 * w[k] + 0.25 * wS[k]) + (0.25 * dt) * dW[k];
    or_compute_psi(cISize, nyP, kxP, kyP, wS, psiS);
    or_truncate(c, cISize, nyP, wS);
    or_truncate(c, cISize, nyP, psiS);
    /* stage 3, :191-209 */
    memset(dW, 0, sizeof(double) * nd);
    or_compute_rhs(c, nxG, nyG, nxP, nyP, rISize, cISize, kxP, kyP, nu, u, v,
                   wS, psiS, dW);
    for (size_t k = 0; k < nd; ++k)
        w[k] = (w[k] / 3.0 + (2.0 * wS[k]) / 3.0) + ((dt * 2.0) * dW[k]) / 3.0;
    or_compute_psi(cISize, nyP, kxP, kyP, w, psi);
    or_truncate(c, cISize, nyP, w);
    or_truncate(c, cISize, nyP, psi);
}

/* --------------------------------------------------- initial condition */

void or_set_taylor_green(or_ctx_t *c, int nxG, int nyG, int nxP, int nyP,
                         int j1, int j2, int rISize, int cISize, double *w,
                         double *psi)
{
    const double PI = acos(-1.0); /* Parameters.F90:44 */
    for (int j = j1; j <= j2; ++j) { /* SetIC.F90:88-96, 1-based i, j */
        int jI = j - j1;
        for (int i = 1; i <= nxP; ++i) {
            double sx = sin(2.0 * PI * (double)i / (double)nxG);
            double sy = sin(2.0 * PI * (double)j / (double)nyG);
            psi[(size_t)(i - 1) + (size_t)jI * rISize] = sx * sy;
            w[(size_t)(i - 1) + (size_t)jI * rISize] = 2.0 * sx * sy;
        }
    }
    or_transform_r2c(c, nxG, nyG, nyP, rISize, cISize, psi);
    or_transform_r2c(c, nxG, nyG, nyP, rISize, cISize, w);
}

static uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

double or_uniform(uint64_t seed, int kx, int ky)
{
    uint64_t key = ((uint64_t)(uint32_t)kx << 32) | (uint64_t)(uint32_t)ky;
    uint64_t h = splitmix64(splitmix64(seed) ^ key);
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

void or_set_turbulence(const or_grid_t *g, const int *kxP, const int *kyP,
                       uint64_t seed, double k0, double ke, double *w,
                       double *psi)
{
    const double PI = acos(-1.0);
    int cI = g->cISize, ny = g->nyG;
    double kMax = g->kMax;
    size_t nd = 2 * (size_t)cI * ny;
    memset(w, 0, sizeof(double) * nd);
    double tke = 0.0;
    for (int j = 0; j < ny; ++j) {
        int ky = kyP[j];
        for (int i = 0; i < cI; ++i) {
            int kx = kxP[i];
            if (kx == 0 && ky == 0) continue;
            double k2 = (double)kx * kx + (double)ky * ky;
            double km = sqrt(k2);
            if (km > kMax + 1.0e-10) continue;
            int sx = kx, sy = ky;
            int conj = 0;
            if (kx == 0 && ky < 0) { sy = -ky; conj = 1; }
            double q = km / k0;
            double E = k2 * k2 * exp(-q * q);
            double au = sqrt(E / (PI * km)); /* |u_k| */
            double th = 2.0 * PI * or_uniform(seed, sx, sy);
            double amp = au * km;            /* |w_k| = |k| |u_k| */
            size_t k = 2 * ((size_t)i + (size_t)j * cI);
            /* glibc sincos, called explicitly so that any compiler's build
             * of this recipe gives the same bits */
            double sn, cs;
            sincos(th, &sn, &cs);
            w[k] = amp * cs;
            w[k + 1] = conj ? -(amp * sn) : amp * sn;
            double wt = (kx == 0 || kx == g->kxMG) ? 1.0 : 2.0;
            tke += wt * 0.5 * au * au;
        }
    }
    double s = (tke > 0.0) ? sqrt(ke / tke) : 0.0;
    for (size_t k = 0; k < nd; ++k) w[k] = w[k] * s;
    or_compute_psi(cI, ny, kxP, kyP, w, psi);
}

double or_cfl_dt(or_ctx_t *c, const or_grid_t *g, const int *kxP,
                 const int *kyP, double cfl, const double *psi, double *u,
                 double *v)
{
    const double PI = acos(-1.0);
    or_compute_velocity(c, g->nxG, g->nyG, g->rISize, g->cISize, kxP, kyP, u,
                        v, psi, 1);
    double m = 0.0;
    for (int j = 0; j < g->nyG; ++j)
        for (int i = 0; i < g->nxG; ++i) {
            size_t k = (size_t)i + (size_t)j * g->rISize;
            double s = fabs(u[k]) + fabs(v[k]);
            if (s > m) m = s;
        }
    double dx = 2.0 * PI / (double)g->nx;
    return (m > 0.0) ? cfl * dx / m : 0.0;
}

/* ------------------------------------------------------------ spectrum */

int or_spectrum_bins(int kxMG, int kyMG)
{
    double kMax = sqrt((double)(kxMG * kxMG + kyMG * kyMG)); /* :326 */
    int kIntMax = (int)(kMax + 1.0);                       /* :329 */
    return kIntMax + 1;                                    /* :335 */
}

/* BinarySearch, src/Analysis.F90:475-513; returns the 1-based bin. */
static int binary_search(int n, const double *fVec, double f)
{
    int iL = 0, iU = n;
    while (iU - iL > 1) {
        int iM = (iL + iU) / 2;
        if (f >= fVec[iM - 1]) iL = iM;
        else iU = iM;
    }
    return iL;
}

int or_compute_spectrum(double nu, int cISize, int nyP, const int *kxP,
                        const int *kyP, int kxMG, int kyMG, const double *u,
                        const double *v, double *ek, int *cnt, double *tke,
                        double *eps, double *enstDiss)
{
    int numBin = or_spectrum_bins(kxMG, kyMG);
    int kIntMax = numBin - 1;
    int numEdges = numBin + 1;
    double *edges = (double *)malloc(sizeof(double) * (size_t)numEdges);
    edges[0] = 0.0; /* :358-362 */
    for (int i = 1; i <= kIntMax; ++i) edges[i] = 0.5 + (double)(i - 1) * 1.0;
    edges[numEdges - 1] = (double)kIntMax + 0.1;
    for (int b = 0; b < numBin; ++b) { ek[b] = 0.0; cnt[b] = 0; }
    for (int j = 0; j < nyP; ++j) {
        double ky = (double)kyP[j];
        for (int i = 0; i < cISize; ++i) {
            double kx = (double)kxP[i];
            double kMag = sqrt(kx * kx + ky * ky);
            size_t k = 2 * ((size_t)i + (size_t)j * cISize);
            double Ehat = (u[k] * u[k] + u[k + 1] * u[k + 1]) +
                          (v[k] * v[k] + v[k + 1] * v[k + 1]);
            int b = binary_search(numEdges, edges, kMag) - 1;
            if (kxP[i] == 0 || kxP[i] == kxMG) { /* :389-395 */
                ek[b] = ek[b] + 0.5 * Ehat;
                cnt[b] += 1;
            } else {
                ek[b] = ek[b] + Ehat;
                cnt[b] += 2;
            }
        }
    }
    double t = 0.0, e = 0.0, z = 0.0; /* :439-448, dk = 1 */
    for (int b = 0; b < numBin; ++b) {
        double kb = (double)b;
        t = t + ek[b];
        e = e + kb * kb * ek[b];
        z = z + kb * kb * kb * kb * ek[b];
    }
    *tke = t;
    *eps = 2.0 * nu * e;
    *enstDiss = 2.0 * nu * z;
    free(edges);
    return numBin;
}
