/*
 * s2d_oracle.h -- CPU FP64 oracle for the Spectral2D TVD-RK3 time step.
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker for the HIP library, not a
 * product path: only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it.
 *
 * It restates, in plain C with its own radix-2/3/4 FFT, the algorithm of
 * mpclay/Spectral2D (Fortran 2003 + FFTW-MPI):
 *   src/Spectral.F90        grid, masks, truncation, transforms, psi-solve,
 *                           derivatives, velocity, vorticity, RHS
 *   src/TimeIntegration.F90 Shu-Osher TVD-RK3 (IntegrateOneStep)
 *   src/SetIC.F90           Taylor-Green sampling (SetTaylorGreen)
 *   src/Analysis.F90        binned energy spectrum (ComputeSpectrum)
 * One process (P = 1), so the local slab is the whole grid: nyP = nyG,
 * j1 = 1, j2 = nyG, kyP = kyG (src/Alloc.F90:62-122).
 *
 * Layout is the reference's: one buffer per field holding allocLocal =
 * cISize*nyP complex words, viewed either as complex (cISize, nyP) or as
 * real (rISize, nyP) with rISize = 2*cISize, Fortran column-major, x fastest
 * (src/Spectral2D.F90:293-304). In C a field is a double* of 2*cISize*nyP
 * doubles; complex element (i, j) (0-based) is at [2*(i + j*cISize)], real
 * element (i, j) at [i + j*rISize].
 *
 * Indices in this API are 0-based; wavenumber arrays hold the same integer
 * values as the reference's kxP/kyP.
 */
#ifndef S2D_ORACLE_H
#define S2D_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Dealiasing methods, src/Spectral.F90:41-47. */
#define OR_NO_DEALIASING 0
#define OR_DEALIAS_3_2_RULE 1
#define OR_DEALIAS_2_3_RULE 2
#define OR_DEALIAS_GRID_SHIFT 3
/* Time integrator, src/TimeIntegration.F90:42. */
#define OR_RK3_TVD_SHU 1

typedef struct or_grid {
    int nx, ny, dealias;
    int nxG, nyG;
    int kxLG, kyLG, kxMG, kyMG;
    int kxLU, kyLU, kxMU, kyMU;
    double kMax;
    int rISize, cISize;
} or_grid_t;

typedef struct or_ctx or_ctx_t;

/* GetGridSize, src/Spectral.F90:117-219. Returns 0, or -1 for a dealiasing
 * method the reference leaves undefined (2/3 rule, grid shift). */
int or_get_grid_size(int nx, int ny, int dealias, or_grid_t *g);
/* kxG (cISize) and kyG (nyG) in FFTW order, src/Spectral.F90:203-218. */
void or_wavenumbers(const or_grid_t *g, int *kxG, int *kyG);

/* Module state of Spectral_m and TimeIntegration_m (plans, masks,
 * truncation table, RK scratch) lives in one context. */
or_ctx_t *or_create(void);
void or_destroy(or_ctx_t *c);
/* Threads used by the 2D transforms (OpenMP builds); returns the count. */
int or_set_threads(int n);

/* SetupMask, src/Spectral.F90:229-260. */
void or_setup_mask(or_ctx_t *c, int cISize, int nyP, const int *kxP,
                   const int *kyP, int kxMG, int kyMG);
/* Query the mask: x/y flags and 0-based indices (-1 when absent). */
void or_get_mask(const or_ctx_t *c, int *xMaskBool, int *xMaskInd,
                 int *yMaskBool, int *yMaskInd);
/* SetupTruncation, src/Spectral.F90:272-313. */
void or_setup_truncation(or_ctx_t *c, int cISize, int nyP, double kMax,
                         const int *kxP, const int *kyP);
/* 0-based first truncated i per row (cISize when the reference leaves the
 * row's entry unset: nothing is truncated there). */
void or_get_truncation(const or_ctx_t *c, int *i1);
/* SetupPlans, src/Spectral.F90:330-353 (twiddle tables instead of FFTW
 * plans; zeroes dataC afterwards like the reference). */
void or_setup_plans(or_ctx_t *c, int nxG, int nyG, int nyP, int rISize,
                    int cISize, double *data);
/* TransformR2C / TransformC2R, src/Spectral.F90:749-781. In place. */
void or_transform_r2c(or_ctx_t *c, int nxG, int nyG, int nyP, int rISize,
                      int cISize, double *data);
void or_transform_c2r(or_ctx_t *c, int rISize, int cISize, int nyP,
                      double *data);
/* DuDx / DuDy, src/Spectral.F90:662-738. in and out may alias. */
void or_dudx(const or_ctx_t *c, int cISize, int nyP, const int *kxP,
             const int *kyP, const double *in, double *out);
void or_dudy(const or_ctx_t *c, int cISize, int nyP, const int *kxP,
             const int *kyP, const double *in, double *out);
/* Truncate, src/Spectral.F90:788-802. */
void or_truncate(const or_ctx_t *c, int cISize, int nyP, double *data);
/* ComputePsi, src/Spectral.F90:508-560. */
void or_compute_psi(int cISize, int nyP, const int *kxP, const int *kyP,
                    const double *Qw, double *Qp);
/* ComputeVelocity, src/Spectral.F90:576-610. */
void or_compute_velocity(or_ctx_t *c, int nxP, int nyP, int rISize,
                         int cISize, const int *kxP, const int *kyP,
                         double *u, double *v, const double *psi, int trans);
/* ComputeVorticity, src/Spectral.F90:624-647 (destroys u and v). */
void or_compute_vorticity(const or_ctx_t *c, int nxP, int nyP, int cISize,
                          const int *kxP, const int *kyP, double *u,
                          double *v, double *w);
/* ComputeRHS, src/Spectral.F90:396-482. Overwrites u, v; round-trips w. */
void or_compute_rhs(or_ctx_t *c, int nxG, int nyG, int nxP, int nyP,
                    int rISize, int cISize, const int *kxP, const int *kyP,
                    double nu, double *u, double *v, double *w,
                    const double *psi, double *dW);
/* TimeIntegrationSetup, src/TimeIntegration.F90:79-108. Returns 0, or -1
 * for an unknown scheme. */
int or_time_integration_setup(or_ctx_t *c, int scheme, int rISize,
                              int cISize, int nyP, int allocLocal);
/* IntegrateOneStep, src/TimeIntegration.F90:130-210. */
void or_integrate_one_step(or_ctx_t *c, int nxG, int nyG, int nxP, int nyP,
                           int rISize, int cISize, const int *kxP,
                           const int *kyP, double nu, double dt, double *u,
                           double *v, double *w, double *psi);
/* SetTaylorGreen, src/SetIC.F90:70-101 (x_i = i*dx sampling, then r2c). */
void or_set_taylor_green(or_ctx_t *c, int nxG, int nyG, int nxP, int nyP,
                         int j1, int j2, int rISize, int cISize, double *w,
                         double *psi);

/* ---- Additions without a reference body (the reference's IC seed is
 * /dev/urandom and ComputeTimeStep is empty). ---- */

/* Counter-based uniform in [0,1) keyed on (seed, kx, ky). */
double or_uniform(uint64_t seed, int kx, int ky);
/* Seeded decaying-turbulence IC on the padded reference layout: kept modes
 * 0 < |k| <= kMax, E(k) ~ k^4 exp(-(k/k0)^2), |u_k|^2 = E/(pi k), phases from
 * or_uniform, Hermitian on kx = 0, scaled so the ComputeSpectrum TKE equals
 * ke. Writes w (vorticity) and psi = ComputePsi(w). */
void or_set_turbulence(const or_grid_t *g, const int *kxP, const int *kyP,
                       uint64_t seed, double k0, double ke, double *w,
                       double *psi);
/* CFL time step: cfl * (2 pi / nx) / max over the padded grid of
 * |u| + |v|, with u, v from psi (ComputeVelocity + c2r). u, v are scratch. */
double or_cfl_dt(or_ctx_t *c, const or_grid_t *g, const int *kxP,
                 const int *kyP, double cfl, const double *psi, double *u,
                 double *v);
/* ComputeSpectrum, src/Analysis.F90:267-460, on spectral u, v. ek and cnt
 * hold numBin = int(sqrt(kxMG^2+kyMG^2)+1)+1 bins; returns numBin and the
 * scalars tke, eps, enstDiss. */
int or_spectrum_bins(int kxMG, int kyMG);
int or_compute_spectrum(double nu, int cISize, int nyP, const int *kxP,
                        const int *kyP, int kxMG, int kyMG, const double *u,
                        const double *v, double *ek, int *cnt, double *tke,
                        double *eps, double *enstDiss);

/* Plain 1D complex DFT through the oracle's own FFT (sign -1 forward,
 * +1 inverse, unnormalized); n must be 2^a 3^b. Exposed for pinning. */
int or_fft_c2c(int n, int sign, const double *in, double *out);

#ifdef __cplusplus
}
#endif
#endif
