/*
 * s2d.h -- C ABI of libs2d, the MI355X (gfx950) library behind the
 * Spectral_m / TimeIntegration_m module interfaces of mpclay/Spectral2D.
 *
 * Every function returns an int status (S2D_OK == 0). On failure the
 * context's message is available from s2d_last_error(ctx) (or
 * s2d_last_error(NULL) for errors raised before a context exists).
 *
 * Data contract (src/Alloc.F90:62-122, src/Spectral2D.F90:287-304, P = 1):
 *   one field = allocLocal = cISize*nyP complex words, viewed as complex
 *   (cISize, nyP) or real (rISize, nyP), Fortran column-major, x fastest.
 *   Host buffers passed to s2d_set_field_host / s2d_get_field_host are
 *   2*cISize*nyP doubles in exactly that layout (C_DOUBLE / C_DOUBLE_COMPLEX,
 *   src/Parameters.F90:31-35). The library keeps device copies; it retains
 *   no host pointer after a call returns.
 *
 * Fields are device slots named by S2D_FIELD_*. The per-routine entry
 * points act on slots with the reference's semantics (in-place transforms,
 * normalization placement, masks, truncation). s2d_rk3_step is the fused
 * device-resident time step: it keeps the vorticity in a compact
 * kept-modes-only layout between calls and refreshes the W and PSI slots
 * only when a later call reads them.
 *
 * Calls on one context are synchronous with respect to the host at the
 * ABI (results are complete when the call returns), not re-entrant, and
 * use one HIP stream per context. Different contexts may be driven from
 * different host threads.
 */
#ifndef S2D_H
#define S2D_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define S2D_OK 0
#define S2D_ERR_INVALID_ARG 1   /* bad size, field id, null pointer */
#define S2D_ERR_UNSUPPORTED 2   /* dealiasing method or grid size not supported */
#define S2D_ERR_HIP 3           /* HIP runtime error */
#define S2D_ERR_ALLOC 4         /* device allocation failed */
#define S2D_ERR_NONFINITE 5     /* NaN/Inf in the state, or CFL blow-up */
#define S2D_ERR_NO_DEVICE 6     /* no GPU visible */

/* ---- constants mirrored from the reference ---- */
#define S2D_NO_DEALIASING 0     /* src/Spectral.F90:41 */
#define S2D_DEALIAS_3_2_RULE 1  /* src/Spectral.F90:43 */
#define S2D_DEALIAS_2_3_RULE 2  /* src/Spectral.F90:45 (empty in the reference: rejected) */
#define S2D_DEALIAS_GRID_SHIFT 3 /* src/Spectral.F90:47 (empty in the reference: rejected) */
#define S2D_RK3_TVD_SHU 1       /* src/TimeIntegration.F90:42 */

/* ---- arithmetic of a context ---- */
#define S2D_FP64 0              /* the reference's C_DOUBLE path */
#define S2D_FP32 1              /* FP32 transforms/products/updates, FP64 twiddles and reductions */

/* ---- device field slots ---- */
#define S2D_FIELD_W 0           /* vorticity (wC / wR) */
#define S2D_FIELD_PSI 1         /* streamfunction (psiC / psiR) */
#define S2D_FIELD_U 2           /* u velocity / nonlinear scratch (uC / uR) */
#define S2D_FIELD_V 3           /* v velocity / nonlinear scratch (vC / vR) */
#define S2D_FIELD_WS 4          /* RK scratch wS (src/TimeIntegration.F90:54-58) */
#define S2D_FIELD_PSIS 5        /* RK scratch psiS (src/TimeIntegration.F90:60-64) */
#define S2D_FIELD_DW 6          /* RK increment dW (src/TimeIntegration.F90:52) */
#define S2D_FIELD_TMP 7         /* free scratch for callers */
#define S2D_NFIELDS 8

typedef struct s2d_ctx s2d_ctx;

/* Output of GetGridSize + GetAllocSize for one process. */
typedef struct s2d_grid {
    int nx, ny, dealias;
    int nxG, nyG;
    int kxLG, kyLG, kxMG, kyMG;
    int kxLU, kyLU, kxMU, kyMU;
    double kMax;
    int rISize, cISize;
    int nxP, nyP, j1, j2, localJ, allocLocal;
} s2d_grid;

/* FP64 diagnostics of the current vorticity (all sums over the full
 * Hermitian spectrum, src/Analysis.F90:380-395 weights). */
typedef struct s2d_diag {
    double tke;         /* 1/2 <u^2 + v^2> */
    double enstrophy;   /* 1/2 <omega^2> */
    double palinstrophy;/* 1/2 <|grad omega|^2> */
    double max_abs_w;   /* max |w_hat| over stored modes (the driver's print, src/Spectral2D.F90:358) */
    double time;        /* simulation time accumulated by the step calls */
    long long nadv;     /* steps taken by the step calls */
} s2d_diag;

/* Message of the last failure on ctx (or of a failed s2d_create when
 * ctx is NULL). Never NULL. */
const char *s2d_last_error(const s2d_ctx *ctx);

/* GetGridSize, src/Spectral.F90:117-219, plus GetAllocSize with one
 * process, src/Alloc.F90:62-122. Host only (no GPU needed). */
int s2d_get_grid_size(int nx, int ny, int dealias, s2d_grid *g);
/* kxG (cISize ints) and kyG (nyG ints) in FFTW order,
 * src/Spectral.F90:203-218. Host only. */
int s2d_wavenumbers(const s2d_grid *g, int *kxG, int *kyG);

/* Replaces the module setup sequence GetGridSize -> SetupMask
 * (src/Spectral.F90:229) -> SetupTruncation (:272) -> SetupPlans (:330)
 * -> TimeIntegrationSetup (src/TimeIntegration.F90:79): allocates the
 * device slots (zeroed), twiddle tables and transform plans.
 * precision is S2D_FP64 or S2D_FP32. */
int s2d_create(int nx, int ny, int dealias, int precision, s2d_ctx **ctx);
/* SpectralFinalize (src/Spectral.F90:805) + TimeIntegrationFinalize
 * (src/TimeIntegration.F90:218). Frees everything. NULL is a no-op. */
int s2d_destroy(s2d_ctx *ctx);

int s2d_get_grid(const s2d_ctx *ctx, s2d_grid *g);
/* SetupMask result, 0-based indices (-1 when absent). */
int s2d_get_mask(const s2d_ctx *ctx, int *xMaskBool, int *xMaskInd,
                 int *yMaskBool, int *yMaskInd);
/* SetupTruncation result: 0-based first truncated i of each row (nyP
 * ints); cISize where the reference leaves the row unset. */
int s2d_get_truncation(const s2d_ctx *ctx, int *i1);

/* Host <-> device copies of one field in the reference layout
 * (2*cISize*nyP doubles). */
int s2d_set_field_host(s2d_ctx *ctx, int field, const double *buf);
int s2d_get_field_host(s2d_ctx *ctx, int field, double *buf);
/* Device pointer of a slot (T = double for S2D_FP64, float for S2D_FP32;
 * cISize*nyP complex T). Reading W or PSI through it after s2d_rk3_step
 * requires s2d_sync_fields first. */
int s2d_field_device_ptr(s2d_ctx *ctx, int field, void **ptr);
/* Refresh the W and PSI slots from the fused step's compact state. */
int s2d_sync_fields(s2d_ctx *ctx);
int s2d_zero_field(s2d_ctx *ctx, int field);
int s2d_copy_field(s2d_ctx *ctx, int src, int dst);

/* TransformR2C, src/Spectral.F90:749-763: forward 2D r2c in place,
 * then divide by nxG*nyG. */
int s2d_transform_r2c(s2d_ctx *ctx, int field);
/* TransformC2R, src/Spectral.F90:772-781: unnormalized 2D c2r in place
 * (destroys the spectral input; Im of the kx = 0 and kx = nxG/2 columns is
 * ignored). */
int s2d_transform_c2r(s2d_ctx *ctx, int field);
/* DuDx / DuDy, src/Spectral.F90:662-693 / 708-738 (in and out may be the
 * same slot). */
int s2d_dudx(s2d_ctx *ctx, int in, int out);
int s2d_dudy(s2d_ctx *ctx, int in, int out);
/* Truncate, src/Spectral.F90:788-802. */
int s2d_truncate(s2d_ctx *ctx, int field);
/* ComputePsi, src/Spectral.F90:508-560 (one process: row 1 is ky = 0). */
int s2d_compute_psi(s2d_ctx *ctx, int w, int psi);
/* ComputeVelocity, src/Spectral.F90:576-610. */
int s2d_compute_velocity(s2d_ctx *ctx, int u, int v, int psi, int trans);
/* ComputeVorticity, src/Spectral.F90:624-647 (destroys u and v). */
int s2d_compute_vorticity(s2d_ctx *ctx, int u, int v, int w);
/* ComputeRHS, src/Spectral.F90:396-482 (overwrites u, v; round-trips w). */
int s2d_compute_rhs(s2d_ctx *ctx, double nu, int u, int v, int w, int psi,
                    int dW);
/* IntegrateOneStep, src/TimeIntegration.F90:130-210, composed literally
 * from the per-routine kernels on slots W, PSI, U, V with scratch WS,
 * PSIS, DW. */
int s2d_integrate_one_step(s2d_ctx *ctx, double nu, double dt);

/* The fused device-resident TVD-RK3 step (same mathematics as
 * IntegrateOneStep; 3 kernels per stage on a compact kept-modes state).
 * After it, W and PSI hold the new state (refreshed lazily); U and V are
 * undefined, as after the reference step (src/Spectral2D.F90:341). */
int s2d_rk3_step(s2d_ctx *ctx, double nu, double dt);
/* nsteps fused steps. cfl > 0 selects a CFL time step recomputed on the
 * device every step from the state at its start:
 * dt = cfl * (2 pi / nx) / max over the padded grid of (|u| + |v|);
 * cfl <= 0 uses the fixed dt. The state is checked for NaN/Inf at the
 * end (S2D_ERR_NONFINITE). */
int s2d_rk3_steps(s2d_ctx *ctx, int nsteps, double nu, double dt, double cfl);
/* ComputeTimeStep, src/TimeIntegration.F90:213-215 (an empty stub in the
 * reference): the CFL dt of the current state, computed on the device. */
int s2d_cfl_dt(s2d_ctx *ctx, double cfl, double *dt);
/* FP64 device reductions of the current state. */
int s2d_diagnostics(s2d_ctx *ctx, s2d_diag *out);
/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): the reductions of s2d_diagnostics on a caller's field, with
 * no allocation of any kind. w: device pointer to a field of the context's
 * grid and precision in the reference layout (cISize * nyG complex);
 * out: device pointer to 5 doubles, receiving tke, enstrophy, palinstrophy,
 * max |w_hat| and the count of non-finite values. Enqueued on `stream` (a
 * hipStream_t; NULL is the null stream) without waiting. Its partial sums
 * use one fixed device array per GPU (nyG <= 6144, i.e. N <= 4096 under the
 * 3/2 rule; larger grids fail with S2D_ERR_UNSUPPORTED): calls that overlap
 * on one GPU race. */
int s2d_diagnostics_noalloc(s2d_ctx *ctx, void *stream, const void *w, double *out);
/* ComputeSpectrum, src/Analysis.F90:267-460, on the device in FP64: shell
 * energies ek[nbins] of the current state (shell b: floor(|k| + 1/2) = b,
 * weights 1/2 at kx = 0 and kx = Np/2), the padded-spectrum mode counts
 * cnt[nbins] (weights 1 / 2), TKE = sum ek, eps = 2 nu sum b^2 ek and the
 * enstrophy dissipation 2 nu sum b^4 ek. nbins = s2d_spectrum_bins(grid).
 * Any output pointer may be NULL. */
int s2d_spectrum_bins(const s2d_grid *g);
int s2d_spectrum(s2d_ctx *ctx, double nu, double *ek, int *cnt, double *tke, double *eps,
                 double *ediss);
/* ComputeSpectrum(..., uC, vC, tke, eps, enstDiss), src/Analysis.F90:267-460,
 * from the velocity spectra held in slots u and v (Ehat = |u|^2 + |v|^2 of
 * each stored mode, src/Analysis.F90:384), with the outputs of s2d_spectrum. */
int s2d_spectrum_uv(s2d_ctx *ctx, int u, int v, double nu, double *ek, int *cnt, double *tke,
                    double *eps, double *ediss);
/* Simulation time / step count held by the context (advanced by the step
 * calls); s2d_set_time resets them. */
int s2d_set_time(s2d_ctx *ctx, double time, long long nadv);
/* End time of the fused steps (s2d_rk3_steps, s2d_closure_steps): t_end > 0
 * clamps every step's dt (fixed or CFL) so that the run lands on t_end; once
 * there, further steps have dt = 0 and keep the state exactly (time and the
 * step count stay). t_end <= 0 removes the limit. No reference counterpart:
 * the reference's driver clamps its last step on the host
 * (src/Spectral2D.F90 DESIRED_END_TIME), which a caller can still do. */
int s2d_set_end_time(s2d_ctx *ctx, double t_end);
/* time, step count and dt of the last step (any pointer may be NULL) */
int s2d_get_time(s2d_ctx *ctx, double *time, long long *nadv, double *dt);

/* ---- filtered DNS and closure coupling (no reference counterpart) ----
 * s2d_restrict: dst slot = the modes kx^2 + ky^2 <= kcut^2 of the src slot
 * (another context's grid; same device), zero elsewhere. Both are in the
 * reference spectral layout with normalised coefficients, which carry over
 * unchanged; where one grid resolves +k and -k and the other has a single
 * Nyquist sample, the Nyquist sample holds their sum. This is the sharp
 * spectral cutoff filter and the sampling onto a coarser grid. */
int s2d_restrict(s2d_ctx *src, int src_field, s2d_ctx *dst, int dst_field, double kcut);
/* slot y = a * slot x + b * slot y */
int s2d_axpby(s2d_ctx *ctx, double a, int x, double b, int y);
/* RHS forcing of the fused step: from now on every stage computes
 * dW = -N - nu k^2 w - F, with F the kept modes of slot f at the time of
 * the call (frozen until the next s2d_set_forcing). The closure term of a
 * coarse run enters as F = tau_hat. s2d_clear_forcing turns it off.
 * Single-run contexts with the compiled fused kernels only. */
int s2d_set_forcing(s2d_ctx *ctx, int f);
int s2d_clear_forcing(s2d_ctx *ctx);
/* The closure forcing of a coarse run in one call: the state's w and psi
 * sampled on `grid` (a context of the same N without padding, FP32) by the
 * sharp cutoff |k| <= N/2, tau = net(w, psi) there (s2d_cnn_*, declared
 * below), its transform restricted back, then s2d_set_forcing. One stream,
 * one wait at the end. Overwrites grid's W, PSI and TMP slots and ctx's TMP
 * slot. */
struct s2d_cnn;
int s2d_closure_forcing(s2d_ctx *ctx, s2d_ctx *grid, struct s2d_cnn *net);
/* nsteps fused steps of a closure run, the forcing refreshed (as
 * s2d_closure_forcing) before a step once `every` steps have been taken with
 * the current one; the count runs on across calls and is reset by
 * s2d_closure_forcing / s2d_set_forcing, so chunks of CFL steps and single
 * fixed steps keep one refresh period. cfl > 0: CFL steps, each step (with
 * or without the refresh) captured once into a hipGraph and replayed
 * (re-captured when the grid, network, nu or cfl change); cfl <= 0: fixed
 * steps of dt, launched eagerly. Checked for NaN/Inf at the end. */
int s2d_closure_steps(s2d_ctx *ctx, s2d_ctx *grid, struct s2d_cnn *net, int nsteps, double nu,
                      double dt, double cfl, int every);
/* on != 0: the closure forcing is zeroed on every shell floor(|k| + 1/2) =
 * b where -F would raise the shell's enstrophy (sum over the shell of
 * Re(conj(w_hat) F) < 0): the closure keeps its net drain, shell by shell. */
int s2d_set_closure_guard(s2d_ctx *ctx, int on);

/* Checkpoint / restart (the neighbour of WriteRestart / ReadRestart,
 * src/IO.F90:98-143, 306-424, which write HDF5): a little-endian binary
 * with the state in the reference spectral layout (W and PSI, FP64) plus,
 * for fused contexts, the compact kept-modes state, simulation time, step
 * count and last dt. Reading it into a context of the same grid restores
 * the state exactly (a fused context of the same precision continues
 * bitwise). spectral2d_amd/checkpoint.py reads the format. */
int s2d_checkpoint_write(s2d_ctx *ctx, const char *path);
int s2d_checkpoint_read(s2d_ctx *ctx, const char *path);

/* Seeded decaying-turbulence initial condition into W and PSI. Replaces
 * SetTurbulentVelocity (src/SetIC.F90:125-151), whose phases come from
 * /dev/urandom, with a reproducible recipe: kept modes 0 < |k| <= kMax,
 * E(k) ~ k^4 exp(-(k/k0)^2), |u_k|^2 = E/(pi |k|), phase 2 pi U where U is a
 * splitmix64 hash of (seed, kx, ky), Hermitian on kx = 0, scaled so that
 * the ComputeSpectrum TKE (src/Analysis.F90:380-395) equals ke; PSI =
 * ComputePsi(W). */
int s2d_set_turbulence(s2d_ctx *ctx, unsigned long long seed, double k0, double ke);
/* Taylor-Green vortex psi = sin x sin y, w = 2 psi sampled at x_i = i dx
 * (1-based i), then TransformR2C: SetTaylorGreen, src/SetIC.F90:70-101. */
int s2d_set_taylor_green(s2d_ctx *ctx);

/* ---- schedule of the fused step and its autotuner ----
 * No reference counterpart beyond FFTW's plan-time measurement
 * (FFTW_MEASURE in SetupPlans, src/Spectral.F90:340-349). A schedule picks,
 * per pass of the fused step, one of the kernel variants compiled for the
 * context's padded size (index 0 is the default; the single-run sizes
 * Np = 768, 1536, 3072 carry alternatives: the other twiddle batching, the
 * other pass B row pairing and, in FP32, 24 elements per lane for passes A
 * and C), and whether multi-step calls replay one captured step (hipGraph).
 * Every variant computes the same mathematics; results differ at rounding
 * level only. */
typedef struct s2d_schedule {
    int pass_a, pass_b, pass_c; /* variant index of each pass */
    int step_graph;             /* 1: multi-step calls replay a captured step */
    double step_us;             /* s2d_tune: measured microseconds per step */
    double default_us;          /* s2d_tune: the default schedule's, eager */
    int candidates, rejected;   /* s2d_tune: schedules timed / failing the check */
    int from_cache;             /* s2d_tune: 1 when read from the cache */
} s2d_schedule;
#define S2D_TUNE_MEASURE 1      /* measure even when the cache has an entry */
#define S2D_TUNE_CACHE_ONLY 2   /* never measure: the cache's entry or the default */
/* n[3]: the variants compiled for each pass at this context's size. */
int s2d_schedule_variants(const s2d_ctx *ctx, int *n);
int s2d_get_schedule(const s2d_ctx *ctx, s2d_schedule *s);
int s2d_set_schedule(s2d_ctx *ctx, const s2d_schedule *s);
/* Time every schedule on the context's own buffers from its current state
 * (steps with nu, dt, cfl as in s2d_rk3_steps), reject any whose step
 * differs from the default's by more than rounding (relative L2 of the new
 * state 1e-12 in FP64, 1e-5 in FP32), set the fastest and restore the state
 * (time and step count included). cache: a JSON-lines file (NULL or "": none)
 * keyed by device, precision and grid; without S2D_TUNE_MEASURE an entry in
 * it is used as is. out (may be NULL) receives the choice. */
int s2d_tune(s2d_ctx *ctx, double nu, double dt, double cfl, const char *cache, int flags,
             s2d_schedule *out);

/* ---- kernel timing (benchmarks) ----
 * Between s2d_profile_begin and s2d_profile_end every launch of the fused
 * stage kernels is bracketed by HIP events recorded on the context's stream
 * (up to max_launches launches). s2d_profile_end returns per-kernel total
 * milliseconds and launch counts, indexed by S2D_KERNEL_*. No reference
 * counterpart (the reference has no timers). */
#define S2D_KERNEL_PASS_A 0     /* column inverse transform */
#define S2D_KERNEL_PASS_B 1     /* row c2r . product . r2c */
#define S2D_KERNEL_PASS_C 2     /* column forward + RHS + RK3 stage */
#define S2D_NKERNELS 3
int s2d_profile_begin(s2d_ctx *ctx, int max_launches);
/* Within a timing session: bracket only the launches of every every-th step
 * (counted from s2d_profile_begin), so that the event pairs' dispatch gaps
 * stay out of the other steps. */
int s2d_profile_every(s2d_ctx *ctx, int every);
int s2d_profile_end(s2d_ctx *ctx, double *ms, long long *count);

/* ---- ensemble of independent runs ----
 * No reference counterpart: the reference runs one simulation per process
 * group (src/Spectral2D.F90). An ensemble holds nruns decaying-turbulence
 * runs of the 3/2-rule fused step, run r on an n[r] x n[r] grid with its
 * own viscosity, CFL time step, simulation time and step count. Runs of
 * one size are batched; one launch per kernel pass advances every run of
 * every size (9 launches per RK3 step). Each run computes exactly what a
 * context of its size computes alone (s2d_create + s2d_set_turbulence +
 * s2d_rk3_steps with the same arguments). Supported N: 64, 128, 256, 512. */
typedef struct s2d_ensemble s2d_ensemble;
#define S2D_ENS_GRAPH 1         /* capture one step into a hipGraph and replay it */
const char *s2d_ensemble_last_error(const s2d_ensemble *ens);
int s2d_ensemble_create(int nruns, const int *n, int precision, s2d_ensemble **out);
int s2d_ensemble_destroy(s2d_ensemble *ens);
int s2d_ensemble_size(const s2d_ensemble *ens, int *nruns, int *nclasses);
int s2d_ensemble_run_grid(s2d_ensemble *ens, int run, s2d_grid *g);
/* nu[nruns] */
int s2d_ensemble_set_viscosity(s2d_ensemble *ens, const double *nu);
/* s2d_set_end_time of every run: t_end[nruns] (<= 0: none), or NULL to
 * clear them all; each run lands on its own end time and then keeps its
 * state while the others go on. */
int s2d_ensemble_set_end_time(s2d_ensemble *ens, const double *t_end);
/* s2d_set_turbulence of every run: seed[nruns], k0[nruns], ke[nruns]. */
int s2d_ensemble_set_turbulence(s2d_ensemble *ens, const unsigned long long *seed,
                                const double *k0, const double *ke);
/* Vorticity of one run from a reference-layout host buffer (as
 * s2d_set_field_host for S2D_FIELD_W). */
int s2d_ensemble_set_vorticity(s2d_ensemble *ens, int run, const double *w);
/* S2D_FIELD_W or S2D_FIELD_PSI of one run into a reference-layout host
 * buffer (as s2d_get_field_host). */
int s2d_ensemble_get_field(s2d_ensemble *ens, int run, int field, double *buf);
/* nsteps RK3 steps of every run; dt and cfl as in s2d_rk3_steps (the CFL
 * step is per run). flags: 0 or S2D_ENS_GRAPH. Every run is checked for
 * NaN/Inf at the end. */
int s2d_ensemble_steps(s2d_ensemble *ens, int nsteps, double dt, double cfl, int flags);
/* s2d_closure_steps for every run of an ensemble of one size: nsteps
 * graph-replayed steps with the closure forcing of each run, refreshed every
 * `every` steps (the period runs on across calls) by one network call at
 * B = runs (the network's batch size must allow it); guard != 0: the shell
 * guard (s2d_set_closure_guard). Each run's forcing, and so its trajectory,
 * is bitwise the one s2d_closure_steps computes for that run alone. grid:
 * an FP32 S2D_NO_DEALIASING context of the runs' size (its transform plans
 * are used, its fields are left alone); net: a network of that grid. At
 * most 32767 runs. The forcing stays on for later s2d_ensemble_steps until
 * s2d_ensemble_clear_forcing. No reference counterpart. */
int s2d_ensemble_closure_steps(s2d_ensemble *ens, s2d_ctx *grid, s2d_cnn *net, int nsteps,
                               double dt, double cfl, int every, int guard);
int s2d_ensemble_clear_forcing(s2d_ensemble *ens);
/* time[nruns], nadv[nruns] (either may be NULL): each run's simulation time
 * and step count (s2d_get_time of each run, one read-back). */
int s2d_ensemble_get_time(s2d_ensemble *ens, double *time, long long *nadv);
/* s2d_spectrum of one run. */
int s2d_ensemble_spectrum(s2d_ensemble *ens, int run, double nu, double *ek, int *cnt,
                          double *tke, double *eps, double *ediss);
/* out[nruns]: s2d_diagnostics of each run. */
int s2d_ensemble_diagnostics(s2d_ensemble *ens, s2d_diag *out);
/* Launch schedule of a step: S2D_ENS_UNIFIED (default), one launch per
 * kernel pass covering every size class; S2D_ENS_UNIFIED_FUSED_A, the same
 * with pass A's three fields of a column tile in one block; S2D_ENS_PER_CLASS,
 * one launch per pass and size class of the kernels compiled for that size
 * alone; S2D_ENS_STAGE_QUEUE, one launch per RK stage whose blocks take the
 * stage's pass A, B and C items from a queue, a run's pass B items waiting
 * for its pass A items and its pass C items for its pass B items. All compute
 * the same bits. */
#define S2D_ENS_UNIFIED 0
#define S2D_ENS_PER_CLASS 1
#define S2D_ENS_PER_CLASS_TUNED 2 /* s2d_ensemble_tune only: per class, each class's variants */
#define S2D_ENS_UNIFIED_FUSED_A 3 /* unified, pass A with the three fields of a tile per block */
#define S2D_ENS_STAGE_QUEUE 4     /* one kernel per RK stage over a work queue (per-run pass
                                     dependencies instead of kernel boundaries) */
int s2d_ensemble_set_schedule(s2d_ensemble *ens, int schedule);
/* The schedule by measurement (the s2d_tune of an ensemble, FFTW_MEASURE's
 * part): the unified launches (with either pass A), the stage queue
 * kernel, the per-class launches with the default variants, and the
 * per-class launches with each class's variants tuned by s2d_tune's search,
 * each checked against the unified step (bitwise, but to rounding for the
 * last) and timed by graph replay from the
 * current state, which is restored; the fastest is set when it beats the
 * unified launches by more than 1 % (else those). cache and flags as
 * in s2d_tune (the key holds the classes' sizes and run counts). *schedule
 * receives the chosen S2D_ENS_* schedule code, *step_us the measured
 * microseconds per step (either may be NULL). */
int s2d_ensemble_tune(s2d_ensemble *ens, double dt, double cfl, const char *cache, int flags,
                      int *schedule, double *step_us);
/* Kernel timing of the ensemble launches (eager steps only), as
 * s2d_profile_begin / s2d_profile_end. */
int s2d_ensemble_profile_begin(s2d_ensemble *ens, int max_launches);
int s2d_ensemble_profile_end(s2d_ensemble *ens, double *ms, long long *count);

/* ---- learned subgrid closure: a periodic 3x3 CNN (no reference
 * counterpart) ----
 * Six 3x3 convolutions with periodic padding: 2 -> 64 -> 64 -> 64 -> 64 ->
 * 64 -> 1 channels, GELU after all but the last; inputs (omega, psi), output
 * tau, all on an H x W grid (W a multiple of 128, H of 8). BF16 weights and
 * activations, FP32 accumulation: the 64 -> 64 layers forward, backward-data
 * and backward-weight on the matrix cores (MFMA). The parameter vector
 * (s2d_cnn_param_count() FP32 values) is W0 [9][2][64] (tap, ci, co), b0
 * [64]; for l = 1..4 Wl [9][64][64] (tap, co, ci), bl [64]; W5 [9][64], b5;
 * tap t = 3 dy + dx is the offset (dy - 1, dx - 1). The inputs are scaled
 * (in0, in1) before the first layer; the output is out * (network value);
 * the training loss is the mean of (network value - target / out)^2.
 * Data pointers are device pointers to FP32 planes: sample n's plane at
 * ptr + n * stride, row pitch ld floats. Calls are synchronous. */
typedef struct s2d_cnn s2d_cnn;
const char *s2d_cnn_last_error(const s2d_cnn *net);
long long s2d_cnn_param_count(void);
int s2d_cnn_create(int H, int W, int batch_max, s2d_cnn **out);
int s2d_cnn_destroy(s2d_cnn *net);
/* host FP32 parameter vectors; set_params also resets the Adam state */
int s2d_cnn_set_params(s2d_cnn *net, const float *params);
int s2d_cnn_get_params(s2d_cnn *net, float *params);
int s2d_cnn_set_scales(s2d_cnn *net, float in0, float in1, float out);
/* clip != 0: s2d_cnn_forward keeps tau only where omega * tau > 0 (the
 * closure term -tau then only removes enstrophy, pointwise); 0: raw */
int s2d_cnn_set_clip(s2d_cnn *net, int clip);
/* Launch width: blocks >= 1 makes every kernel use as few blocks as it can
 * while keeping at least `blocks` (row-tiled kernels take more rows per
 * block) and grid-stride kernels use at most `blocks`; 0 (the default)
 * restores each kernel's own width (256 or 512 blocks). Results are the
 * same to rounding (the forward bitwise); a width below the GPU's CU count
 * only serves tests of the kernels' multi-row paths at small sizes.
 * blocks <= 512 (the partial sums are sized for the default widths). */
int s2d_cnn_set_launch_width(s2d_cnn *net, int blocks);
/* tau (= out * network value) of B samples */
int s2d_cnn_forward(s2d_cnn *net, int B, const float *w, const float *psi, long long stride,
                    int ld, float *tau, long long tau_stride, int tau_ld);
/* the same on a caller's HIP stream (a hipStream_t), without waiting */
int s2d_cnn_forward_on(s2d_cnn *net, void *stream, int B, const float *w, const float *psi,
                       long long stride, int ld, float *tau, long long tau_stride, int tau_ld);
int s2d_cnn_grid(const s2d_cnn *net, int *H, int *W);
/* loss and its gradient (host FP32 vector, may be NULL) at the current
 * parameters, for B samples against target planes */
int s2d_cnn_gradient(s2d_cnn *net, int B, const float *w, const float *psi, long long stride,
                     int ld, const float *target, long long target_stride, int target_ld,
                     double *loss, float *grad);
/* gradient, then one Adam step (beta 0.9 / 0.999, eps 1e-8) of rate lr */
int s2d_cnn_train_step(s2d_cnn *net, int B, const float *w, const float *psi, long long stride,
                       int ld, const float *target, long long target_stride, int target_ld,
                       float lr, double *loss);
/* A training set held on the device: n samples of dense H x W FP32 planes
 * (host arrays w[n][H][W], psi, tau); s2d_cnn_train_batch gathers samples
 * idx[0..B-1] and takes one s2d_cnn_train_step on them; s2d_cnn_predict
 * writes the network's tau of sample i to a host plane. */
int s2d_cnn_set_dataset(s2d_cnn *net, long long n, const float *w, const float *psi,
                        const float *tau);
int s2d_cnn_train_batch(s2d_cnn *net, int B, const int *idx, float lr, double *loss);
int s2d_cnn_predict(s2d_cnn *net, long long i, float *tau);

#ifdef __cplusplus
}
#endif
#endif
