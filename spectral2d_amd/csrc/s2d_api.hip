// s2d_api.hip -- the C ABI of libs2d (declared in include/s2d.h).
//
// Host side of the library: grid set-up (GetGridSize / GetAllocSize /
// SetupMask / SetupTruncation restated for one process), transform plans
// and FP64-generated twiddle tables, device slots, and the per-routine and
// fused-step entry points. All GPU work of a context goes to its own HIP
// stream; every entry point synchronises that stream before returning.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "kernels_fused.hpp"
#include "kernels_ensemble.hpp"
#include "kernels_fused_v2.hpp"
#include "kernels_ref.hpp"
#include "kernels_spectrum.hpp"
#include "kernels_closure.hpp"
#include "s2d_internal.hpp"

using namespace s2d;

namespace {

thread_local std::string g_err = "no error";

int vfail(s2d_ctx *c, int code, const char *fmt, va_list ap)
{
    char buf[512];
    vsnprintf(buf, sizeof(buf), fmt, ap);
    if (c)
        c->err = buf;
    else
        g_err = buf;
    return code;
}

int fail(s2d_ctx *c, int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    int r = vfail(c, code, fmt, ap);
    va_end(ap);
    return r;
}

#define S2D_HIP(c, call)                                                                   \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail((c), S2D_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));  \
    } while (0)

#define S2D_LAUNCHED(c, what)                                                              \
    do {                                                                                   \
        hipError_t e_ = hipGetLastError();                                                 \
        if (e_ != hipSuccess)                                                              \
            return fail((c), S2D_ERR_HIP, "launch of %s failed: %s", what,                 \
                        hipGetErrorString(e_));                                            \
    } while (0)

#define S2D_TRY(expr)             \
    do {                          \
        int r_ = (expr);          \
        if (r_ != S2D_OK) return r_; \
    } while (0)

constexpr int kRadices[4] = {8, 4, 3, 2};

bool factorable(int n, int E)
{
    int m = n;
    while (m > 1) {
        bool ok = false;
        for (int r : kRadices)
            if (E % r == 0 && m % r == 0) {
                m /= r;
                ok = true;
                break;
            }
        if (!ok) return false;
    }
    return true;
}

int grid_blocks(size_t n, int threads)
{
    size_t b = (n + threads - 1) / threads;
    if (b > 8192) b = 8192;
    if (b < 1) b = 1;
    return (int)b;
}

bool valid_field(int f) { return f >= 0 && f < S2D_NFIELDS; }

// Counter-based uniform in [0,1): splitmix64 of the seed mixed with the
// packed (kx, ky) key, top 53 bits.
uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

double hash_uniform(uint64_t seed, int kx, int ky)
{
    const uint64_t key = ((uint64_t)(uint32_t)kx << 32) | (uint64_t)(uint32_t)ky;
    const uint64_t h = splitmix64(splitmix64(seed) ^ key);
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

} // namespace

namespace s2d {

int choose_elems(int n)
{
    const int cand[6] = {12, 8, 6, 4, 3, 2};
    for (int E : cand)
        if (n % E == 0 && factorable(n, E)) return E;
    return 0;
}

bool make_plan(int n, int E, FftPlan *p)
{
    if (E <= 0 || n % E != 0 || !factorable(n, E)) return false;
    p->n = n;
    p->T = n / E;
    p->npass = 0;
    int m = n;
    while (m > 1) {
        for (int r : kRadices)
            if (E % r == 0 && m % r == 0) {
                if (p->npass >= kMaxPasses) return false;
                p->radix[p->npass++] = r;
                m /= r;
                break;
            }
    }
    if (p->npass == 0) { // n == 1 cannot happen for E >= 2
        return false;
    }
    return true;
}

} // namespace s2d

// ------------------------------------------------------------ helpers

namespace {

template <typename T>
std::vector<cpx<T>> host_twiddles(int n)
{
    std::vector<cpx<T>> tw((size_t)n);
    const long double two_pi = 6.283185307179586476925286766559005768L;
    for (int e = 0; e < n; ++e) {
        long double a = two_pi * (long double)e / (long double)n;
        tw[(size_t)e].re = (T)(double)cosl(a);
        tw[(size_t)e].im = (T)(double)(-sinl(a));
    }
    return tw;
}

int make_devfft(s2d_ctx *c, int n, DevFft *f)
{
    f->E = choose_elems(n);
    if (f->E == 0 || !make_plan(n, f->E, &f->plan))
        return fail(c, S2D_ERR_UNSUPPORTED, "transform length %d is not 2^a 3^b (>= 2)", n);
    if (c->precision == S2D_FP64) {
        auto tw = host_twiddles<double>(n);
        S2D_HIP(c, hipMalloc(&f->tw, tw.size() * sizeof(tw[0])));
        S2D_HIP(c, hipMemcpy(f->tw, tw.data(), tw.size() * sizeof(tw[0]), hipMemcpyHostToDevice));
    } else {
        auto tw = host_twiddles<float>(n);
        S2D_HIP(c, hipMalloc(&f->tw, tw.size() * sizeof(tw[0])));
        S2D_HIP(c, hipMemcpy(f->tw, tw.data(), tw.size() * sizeof(tw[0]), hipMemcpyHostToDevice));
    }
    return S2D_OK;
}

size_t nc_of(const s2d_ctx *c) { return (size_t)c->g.cISize * (size_t)c->g.nyG; }

template <typename T>
cpx<T> *F(s2d_ctx *c, int f) { return reinterpret_cast<cpx<T> *>(c->field[f]); }

// Launch helper: dynamic-LDS kernels may exceed the 64 KiB default.
template <typename K>
void allow_lds(K kern, size_t bytes)
{
    if (bytes > 65536)
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

// While profiling is on, a ProfScope reserves an event pair for the next
// fused launch of its context; launch() hands the pair to
// hipExtLaunchKernelGGL, which records them at the kernel's own start and
// end (no dispatch gaps inside the interval).
struct ProfScope {
    s2d_ctx *c;
    int slot = -1;
    ProfScope(s2d_ctx *c_, int kind) : c(c_)
    {
        if (c->prof_on && c->prof_step_on && c->prof_used < (int)c->prof_kind.size()) {
            slot = c->prof_used++;
            c->prof_kind[slot] = kind;
            c->prof_slot = slot;
        }
    }
    ~ProfScope() { c->prof_slot = -1; }
};

// Marks the start of a fused step: only every prof_every-th step of a
// timing session records events (the event pairs add a few microseconds of
// dispatch gap per launch, which sampling keeps out of the other steps).
inline void prof_next_step(s2d_ctx *c)
{
    if (!c->prof_on) return;
    c->prof_step_on = (c->prof_step % c->prof_every) == 0;
    ++c->prof_step;
}

template <typename K, typename... Args>
void launch(s2d_ctx *c, K kern, dim3 grid, dim3 block, size_t shmem, Args... args)
{
    if (c->prof_slot >= 0) {
        const int u = c->prof_slot;
        c->prof_slot = -1;
        hipExtLaunchKernelGGL(kern, grid, block, (uint32_t)shmem, c->stream, c->prof_ev[2 * u],
                              c->prof_ev[2 * u + 1], 0, args...);
    } else {
        hipLaunchKernelGGL(kern, grid, block, shmem, c->stream, args...);
    }
}

// ----------------------------------------------- reference-layout FFTs

// np planes ps elements apart (one by default)
template <typename T, int E, int D>
int launch_col_E(s2d_ctx *c, cpx<T> *data, int normalize, int np, long long ps)
{
    const FftPlan &p = c->col.plan;
    int B = 1024 / p.T;
    int Bl = (int)(c->max_lds / ((size_t)p.n * sizeof(cpx<T>)));
    if (Bl < B) B = Bl;
    if (B > 16) B = 16;
    if (B < 1) return fail(c, S2D_ERR_UNSUPPORTED, "column transform of %d does not fit in LDS", p.n);
    const size_t lds = (size_t)B * p.n * sizeof(cpx<T>);
    auto kern = k_col_fft<T, E, D>;
    allow_lds(kern, lds);
    const int blocks = (c->g.cISize + B - 1) / B;
    const double norm = (double)c->g.nxG * (double)c->g.nyG;
    launch(c, kern, dim3(blocks, np), dim3(B * p.T), lds, data, c->g.cISize, p,
           (const cpx<T> *)c->col.tw, B, normalize, norm, ps);
    S2D_LAUNCHED(c, "k_col_fft");
    return S2D_OK;
}

template <typename T, int D>
int launch_col(s2d_ctx *c, cpx<T> *data, int normalize, int np = 1, long long ps = 0)
{
    switch (c->col.E) {
    case 12: return launch_col_E<T, 12, D>(c, data, normalize, np, ps);
    case 8: return launch_col_E<T, 8, D>(c, data, normalize, np, ps);
    case 6: return launch_col_E<T, 6, D>(c, data, normalize, np, ps);
    case 4: return launch_col_E<T, 4, D>(c, data, normalize, np, ps);
    case 3: return launch_col_E<T, 3, D>(c, data, normalize, np, ps);
    case 2: return launch_col_E<T, 2, D>(c, data, normalize, np, ps);
    }
    return fail(c, S2D_ERR_UNSUPPORTED, "no column kernel for E=%d", c->col.E);
}

template <typename T, int E, bool R2C>
int launch_row_E(s2d_ctx *c, cpx<T> *data, int np, long long ps)
{
    const FftPlan &p = c->rowh.plan;
    int B = 256 / p.T;
    if (B < 1) B = 1;
    int Bl = (int)(c->max_lds / ((size_t)p.n * sizeof(cpx<T>)));
    if (Bl < B) B = Bl;
    if (B < 1) return fail(c, S2D_ERR_UNSUPPORTED, "row transform of %d does not fit in LDS", p.n);
    const size_t lds = (size_t)B * p.n * sizeof(cpx<T>);
    const int blocks = (c->g.nyG + B - 1) / B;
    if constexpr (R2C) {
        auto kern = k_row_r2c<T, E>;
        allow_lds(kern, lds);
        launch(c, kern, dim3(blocks, np), dim3(B * p.T), lds, data, c->g.cISize, c->g.nyG, p,
               (const cpx<T> *)c->rowh.tw, (const cpx<T> *)c->tw_x, B, ps);
    } else {
        auto kern = k_row_c2r<T, E>;
        allow_lds(kern, lds);
        launch(c, kern, dim3(blocks, np), dim3(B * p.T), lds, data, c->g.cISize, c->g.nyG, p,
               (const cpx<T> *)c->rowh.tw, (const cpx<T> *)c->tw_x, B, ps);
    }
    S2D_LAUNCHED(c, R2C ? "k_row_r2c" : "k_row_c2r");
    return S2D_OK;
}

template <typename T, bool R2C>
int launch_row(s2d_ctx *c, cpx<T> *data, int np = 1, long long ps = 0)
{
    switch (c->rowh.E) {
    case 12: return launch_row_E<T, 12, R2C>(c, data, np, ps);
    case 8: return launch_row_E<T, 8, R2C>(c, data, np, ps);
    case 6: return launch_row_E<T, 6, R2C>(c, data, np, ps);
    case 4: return launch_row_E<T, 4, R2C>(c, data, np, ps);
    case 3: return launch_row_E<T, 3, R2C>(c, data, np, ps);
    case 2: return launch_row_E<T, 2, R2C>(c, data, np, ps);
    }
    return fail(c, S2D_ERR_UNSUPPORTED, "no row kernel for E=%d", c->rowh.E);
}

// ------------------------------------------------- per-routine ops (T)

template <typename T>
int op_fill(s2d_ctx *c, int f)
{
    size_t n = nc_of(c);
    hipLaunchKernelGGL(k_fill<T>, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream, F<T>(c, f), n);
    S2D_LAUNCHED(c, "k_fill");
    return S2D_OK;
}

template <typename T>
int op_r2c(s2d_ctx *c, int f)
{
    S2D_TRY((launch_row<T, true>(c, F<T>(c, f))));
    return launch_col<T, -1>(c, F<T>(c, f), 1);
}

template <typename T>
int op_c2r(s2d_ctx *c, int f)
{
    S2D_TRY((launch_col<T, 1>(c, F<T>(c, f), 0)));
    return launch_row<T, false>(c, F<T>(c, f));
}

template <typename T>
int op_dudx(s2d_ctx *c, int in, int out)
{
    size_t n = nc_of(c);
    hipLaunchKernelGGL(k_dudx<T>, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream,
                       F<T>(c, in), F<T>(c, out), c->g.cISize, c->g.nyG, c->d_kx,
                       c->xMaskBool ? c->xMaskInd : -1);
    S2D_LAUNCHED(c, "k_dudx");
    return S2D_OK;
}

template <typename T>
int op_dudy(s2d_ctx *c, int in, int out)
{
    size_t n = nc_of(c);
    hipLaunchKernelGGL(k_dudy<T>, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream,
                       F<T>(c, in), F<T>(c, out), c->g.cISize, c->g.nyG, c->d_ky,
                       c->yMaskBool ? c->yMaskInd : -1);
    S2D_LAUNCHED(c, "k_dudy");
    return S2D_OK;
}

template <typename T>
int op_truncate(s2d_ctx *c, int f)
{
    size_t n = nc_of(c);
    hipLaunchKernelGGL(k_truncate<T>, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream,
                       F<T>(c, f), c->g.cISize, c->g.nyG, c->d_trunc);
    S2D_LAUNCHED(c, "k_truncate");
    return S2D_OK;
}

template <typename T>
int op_psi(s2d_ctx *c, int w, int psi)
{
    size_t n = nc_of(c);
    hipLaunchKernelGGL(k_psi<T>, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream, F<T>(c, w),
                       F<T>(c, psi), c->g.cISize, c->g.nyG, c->d_kx, c->d_ky);
    S2D_LAUNCHED(c, "k_psi");
    return S2D_OK;
}

template <typename T>
int op_velocity(s2d_ctx *c, int u, int v, int psi, int trans)
{
    size_t n = nc_of(c);
    hipLaunchKernelGGL(k_velocity<T>, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream,
                       F<T>(c, psi), F<T>(c, u), F<T>(c, v), c->g.cISize, c->g.nyG, c->d_kx,
                       c->d_ky, c->xMaskBool ? c->xMaskInd : -1,
                       c->yMaskBool ? c->yMaskInd : -1);
    S2D_LAUNCHED(c, "k_velocity");
    if (trans) {
        S2D_TRY(op_c2r<T>(c, u));
        S2D_TRY(op_c2r<T>(c, v));
    }
    return S2D_OK;
}

template <typename T>
int op_vorticity(s2d_ctx *c, int u, int v, int w)
{
    size_t n = nc_of(c);
    hipLaunchKernelGGL(k_vorticity<T>, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream,
                       F<T>(c, u), F<T>(c, v), F<T>(c, w), c->g.cISize, c->g.nyG, c->d_kx,
                       c->d_ky, c->xMaskBool ? c->xMaskInd : -1,
                       c->yMaskBool ? c->yMaskInd : -1);
    S2D_LAUNCHED(c, "k_vorticity");
    return S2D_OK;
}

template <typename T>
int op_rhs(s2d_ctx *c, double nu, int u, int v, int w, int psi, int dW)
{
    S2D_TRY(op_velocity<T>(c, u, v, psi, 0)); // Spectral.F90:424-425
    S2D_TRY(op_c2r<T>(c, u));                 // :433-435
    S2D_TRY(op_c2r<T>(c, v));
    S2D_TRY(op_c2r<T>(c, w));
    {
        size_t n = (size_t)c->g.nxG * c->g.nyG; // :442-447
        hipLaunchKernelGGL(k_product<T>, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream,
                           reinterpret_cast<T *>(c->field[u]), reinterpret_cast<T *>(c->field[v]),
                           reinterpret_cast<const T *>(c->field[w]), c->g.nxG, c->g.rISize,
                           c->g.nyG);
        S2D_LAUNCHED(c, "k_product");
    }
    S2D_TRY(op_r2c<T>(c, u)); // :451-453
    S2D_TRY(op_r2c<T>(c, v));
    S2D_TRY(op_r2c<T>(c, w));
    S2D_TRY(op_truncate<T>(c, u)); // :457-458
    S2D_TRY(op_truncate<T>(c, v));
    S2D_TRY(op_dudx<T>(c, u, u)); // :464-465
    S2D_TRY(op_dudy<T>(c, v, v));
    size_t n = nc_of(c); // :471-481
    hipLaunchKernelGGL(k_rhs<T>, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream, F<T>(c, u),
                       F<T>(c, v), F<T>(c, w), F<T>(c, dW), c->g.cISize, c->g.nyG, c->d_kx,
                       c->d_ky, nu);
    S2D_LAUNCHED(c, "k_rhs");
    return S2D_OK;
}

template <typename T>
int op_stage(s2d_ctx *c, int stage, double dt)
{
    size_t n = nc_of(c);
    hipLaunchKernelGGL(k_rk3_stage<T>, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream, stage,
                       F<T>(c, S2D_FIELD_W), F<T>(c, S2D_FIELD_WS), F<T>(c, S2D_FIELD_DW), n, dt);
    S2D_LAUNCHED(c, "k_rk3_stage");
    return S2D_OK;
}

__global__ void k_advance_time(DevScalars *sc, double dt)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        sc->dt = dt;
        sc->time = sc->time + dt;
        sc->nadv = sc->nadv + 1;
    }
}

// IntegrateOneStep, src/TimeIntegration.F90:130-210, literally.
template <typename T>
int op_integrate(s2d_ctx *c, double nu, double dt)
{
    const int W = S2D_FIELD_W, PSI = S2D_FIELD_PSI, U = S2D_FIELD_U, V = S2D_FIELD_V;
    const int WS = S2D_FIELD_WS, PSIS = S2D_FIELD_PSIS, DW = S2D_FIELD_DW;
    S2D_TRY(op_fill<T>(c, WS)); // :147-149
    S2D_TRY(op_fill<T>(c, PSIS));
    S2D_TRY(op_fill<T>(c, DW));
    S2D_TRY(op_rhs<T>(c, nu, U, V, W, PSI, DW)); // :152-153
    S2D_TRY(op_stage<T>(c, 1, dt));              // :156-160
    S2D_TRY(op_psi<T>(c, WS, PSIS));             // :163
    S2D_TRY(op_truncate<T>(c, WS));              // :166-167
    S2D_TRY(op_truncate<T>(c, PSIS));
    S2D_TRY(op_fill<T>(c, DW)); // :170-172
    S2D_TRY(op_rhs<T>(c, nu, U, V, WS, PSIS, DW));
    S2D_TRY(op_stage<T>(c, 2, dt)); // :175-181
    S2D_TRY(op_psi<T>(c, WS, PSIS));
    S2D_TRY(op_truncate<T>(c, WS));
    S2D_TRY(op_truncate<T>(c, PSIS));
    S2D_TRY(op_fill<T>(c, DW)); // :191-193
    S2D_TRY(op_rhs<T>(c, nu, U, V, WS, PSIS, DW));
    S2D_TRY(op_stage<T>(c, 3, dt)); // :196-202
    S2D_TRY(op_psi<T>(c, W, PSI));  // :205
    S2D_TRY(op_truncate<T>(c, W));  // :208-209
    S2D_TRY(op_truncate<T>(c, PSI));
    hipLaunchKernelGGL(k_advance_time, dim3(1), dim3(64), 0, c->stream, c->d_sc, dt);
    S2D_LAUNCHED(c, "k_advance_time");
    return S2D_OK;
}

// ----------------------------------------------------------- fused ops

FusedDims fdims(const s2d_ctx *c)
{
    FusedDims d;
    d.N = c->N;
    d.Np = c->g.nxG;
    d.K = c->K;
    d.KP = c->KP;
    d.cI = c->g.cISize;
    d.KG = c->KG;
    return d;
}

template <typename T>
int op_pack(s2d_ctx *c, int run = 0)
{
    FusedDims d = fdims(c);
    size_t n = (size_t)d.K * d.KP;
    hipLaunchKernelGGL(k_pack<T>, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream,
                       F<T>(c, S2D_FIELD_W), (cpx<T> *)c->Wn + run * c->sW, d, c->d_trunc, 0LL,
                       0LL);
    S2D_LAUNCHED(c, "k_pack");
    return S2D_OK;
}

template <typename T>
int op_unpack(s2d_ctx *c, int run = 0)
{
    FusedDims d = fdims(c);
    size_t n = nc_of(c);
    hipLaunchKernelGGL(k_unpack<T>, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream,
                       (const cpx<T> *)c->Wn + run * c->sW, F<T>(c, S2D_FIELD_W),
                       F<T>(c, S2D_FIELD_PSI), d, 0LL, 0LL);
    S2D_LAUNCHED(c, "k_unpack");
    return S2D_OK;
}

// rows per pass-B workgroup come in pairs: B pairs = 2B rows; the kernel is
// built for at most 512 threads
int passB_pairs(const s2d_ctx *c)
{
    const int Np = c->g.nxG;
    const int T = c->rowf.plan.T;
    int B = (c->precision == S2D_FP64) ? 4 : 8;
    while (B > 1 && ((size_t)B * T > 512 || (size_t)B * Np * c->celt > (size_t)c->max_lds ||
                     (Np / 2) % B != 0))
        B /= 2;
    return B;
}


template <typename T, int E>
int launch_passA(s2d_ctx *c, const cpx<T> *src, int reset)
{
    ProfScope ps(c, S2D_KERNEL_PASS_A);
    FusedDims d = fdims(c);
    const FftPlan &p = c->col.plan;
    const size_t lds = (size_t)3 * d.Np * sizeof(cpx<T>);
    auto kern = k_passA<T, E>;
    allow_lds(kern, lds);
    launch(c, kern, dim3(d.K), dim3(3 * p.T), lds, src, (cpx<T> *)c->G[0],
           (cpx<T> *)c->G[1], (cpx<T> *)c->G[2], d, p, (const cpx<T> *)c->col.tw,
           c->d_sc, reset);
    S2D_LAUNCHED(c, "k_passA");
    return S2D_OK;
}

template <typename T, int E>
int launch_passB(s2d_ctx *c, int forward, int track_max)
{
    ProfScope ps(c, S2D_KERNEL_PASS_B);
    FusedDims d = fdims(c);
    const FftPlan &p = c->rowf.plan;
    const int B = passB_pairs(c);
    const size_t lds = (size_t)B * d.Np * sizeof(cpx<T>);
    auto kern = k_passB<T, E>;
    allow_lds(kern, lds);
    const int blocks = (d.Np + 2 * B - 1) / (2 * B);
    launch(c, kern, dim3(blocks), dim3(B * p.T), lds, (const cpx<T> *)c->G[0],
           (const cpx<T> *)c->G[1], (const cpx<T> *)c->G[2], (cpx<T> *)c->H[0],
           (cpx<T> *)c->H[1], d, p, (const cpx<T> *)c->rowf.tw, c->d_sc, B, forward,
           track_max);
    S2D_LAUNCHED(c, "k_passB");
    return S2D_OK;
}

template <typename T, int E>
int launch_passC(s2d_ctx *c, const cpx<T> *Wst, cpx<T> *Wout, const StageArgs &a)
{
    ProfScope ps(c, S2D_KERNEL_PASS_C);
    FusedDims d = fdims(c);
    const FftPlan &p = c->col.plan;
    const size_t lds = (size_t)2 * d.Np * sizeof(cpx<T>);
    auto kern = k_passC<T, E>;
    allow_lds(kern, lds);
    launch(c, kern, dim3(d.K), dim3(2 * p.T), lds, (const cpx<T> *)c->H[0],
           (const cpx<T> *)c->H[1], (const cpx<T> *)c->Wn, Wst, Wout, d, p,
           (const cpx<T> *)c->col.tw, c->d_trunc, c->d_sc, a);
    S2D_LAUNCHED(c, "k_passC");
    return S2D_OK;
}

template <typename T, int EC, int ER>
int fused_steps_E(s2d_ctx *c, int nsteps, double nu, double dt, double cfl)
{
    cpx<T> *Wn = (cpx<T> *)c->Wn, *Ws = (cpx<T> *)c->Ws;
    StageArgs a;
    a.nu = nu;
    a.dt = dt;
    a.cfl = cfl;
    a.dx = 2.0 * M_PI / (double)c->N;
    a.norm = (double)c->g.nxG * (double)c->g.nyG;
    a.inv_norm = 1.0 / a.norm;
    for (int st = 0; st < nsteps; ++st) {
        prof_next_step(c);
        for (int stage = 1; stage <= 3; ++stage) {
            const cpx<T> *src = stage == 1 ? Wn : Ws;
            cpx<T> *dst = stage == 3 ? Wn : Ws;
            a.stage = stage;
            S2D_TRY((launch_passA<T, EC>(c, src, stage == 1)));
            S2D_TRY((launch_passB<T, ER>(c, 1, stage == 1)));
            S2D_TRY((launch_passC<T, EC>(c, src, dst, a)));
        }
    }
    return S2D_OK;
}

// ---- lane-group FFT kernels for the compiled padded sizes

// The context's runs as a kernel batch (per-run viscosity for an ensemble
// size class; a plain context takes nu from the call).
Batch batch_of(const s2d_ctx *c)
{
    Batch b;
    b.nruns = c->nruns;
    b.sW = c->sW;
    b.sG = c->sG;
    b.sH = c->sH;
    b.nu = c->member ? c->d_nu : nullptr;
    b.forcing = c->forcing_on ? c->forcing : nullptr;
    b.vslot = c->member ? nullptr : c->vslot;
    return b;
}

template <typename T, int NP, int V = 0>
int launch2_passA(s2d_ctx *c, const cpx<T> *src, int reset)
{
    constexpr int LPB = WaveCfg<T, NP, V>::LPB_A;
    FusedDims d = fdims(c);
    int blocks = passA_blocks<T, NP, LPB, V>(d.K, c->nruns);
    blocks = (blocks + 7) / 8 * 8; // a multiple of 8 for the XCD-major order
    launch(c, k2_passA<T, NP, V>, dim3(blocks), dim3(TileCfg<T, NP, LPB, V>::BLK), 0, src, (cpx<T> *)c->G[0], (cpx<T> *)c->G[1], (cpx<T> *)c->G[2],
                       d, (const cpx<T> *)c->col.tw, c->d_sc, reset, batch_of(c));
    S2D_LAUNCHED(c, "k2_passA");
    return S2D_OK;
}

template <typename T, int NP, int V = 0>
int launch2_passB(s2d_ctx *c, int forward, int track_max)
{
    constexpr int LPB = WaveCfg<T, NP, V>::LPB_B;
    FusedDims d = fdims(c);
    int blocks = passB_blocks<T, NP, LPB>(c->nruns);
    blocks = (blocks + 7) / 8 * 8;
    launch(c, k2_passB<T, NP, V>, dim3(blocks), dim3(WaveCfg<T, NP, V>::THR_B), 0,
           (const cpx<T> *)c->G[0], (const cpx<T> *)c->G[1],
                       (const cpx<T> *)c->G[2], (cpx<T> *)c->H[0], (cpx<T> *)c->H[1], d,
                       (const cpx<T> *)c->rowf.tw, c->d_sc, forward, track_max, batch_of(c));
    S2D_LAUNCHED(c, "k2_passB");
    return S2D_OK;
}

template <typename T, int NP, int V = 0>
int launch2_passC(s2d_ctx *c, const cpx<T> *Wst, cpx<T> *Wout, const StageArgs &a)
{
    using Cf = WaveCfg<T, NP, V>;
    FusedDims d = fdims(c);
    // the transformed columns (with passC_tail, each run's column K - 1 is
    // written by the block holding its column 0)
    const int cols = passC_tail_on<NP>() ? d.K - 1 : d.K;
    const int blocks = (cols * c->nruns + Cf::GPB - 1) / Cf::GPB;
    launch(c, k2_passC<T, NP, V>, dim3(blocks), dim3(2 * Cf::BLK), 0, (const cpx<T> *)c->H[0], (const cpx<T> *)c->H[1], (const cpx<T> *)c->Wn,
                       Wst, Wout, d, (const cpx<T> *)c->col.tw, c->d_trunc, c->d_sc, a,
                       batch_of(c));
    S2D_LAUNCHED(c, "k2_passC");
    return S2D_OK;
}

// The compiled schedule variants (VAR_* bit masks, kernels_fused_v2.hpp) of
// pass P (0 = A, 1 = B, 2 = C) at padded size NP; index 0 is the default.
// The single-run sizes 768, 1536 and 3072 carry alternatives: the other
// twiddle mode for every pass, the other pass B row pairing, (FP32) 24
// elements per lane for passes A and C, and pass A with the three fields of
// a tile in one block. (24 per lane spills FP64 passes and every pass B.)
// The ensemble's smaller sizes have the default only.
template <typename T, int NP>
constexpr bool tune_size()
{
    return NP == 768 || NP == 1536 || NP == 3072;
}

template <typename T, int NP, int P>
constexpr int n_variants()
{
    if constexpr (!tune_size<T, NP>()) return 1;
    else if constexpr (P == 0) return sizeof(T) == 4 ? 4 : 3;
    else if constexpr (P == 1) return 3;
    else return sizeof(T) == 4 ? 3 : 2;
}

template <typename T, int NP, int P, int I>
constexpr int variant_bits()
{
    static_assert(I < n_variants<T, NP, P>(), "no such variant");
    if constexpr (I == 0) return 0;
    else if constexpr (P == 0)
        return I == 1 ? VAR_TW_A : (I == 2 && sizeof(T) == 4) ? VAR_E24 : VAR_FUSE_A;
    else if constexpr (P == 1) return I == 1 ? VAR_TW_B : VAR_PAIR;
    else return I == 1 ? VAR_TW_C : VAR_E24;
}

// f(std::integral_constant<int, bits>) for variant idx of pass P.
template <typename T, int NP, int P, int I = 0, typename F>
int with_variant(int idx, F &&f)
{
    if constexpr (I < n_variants<T, NP, P>()) {
        if (idx == I) return f(std::integral_constant<int, variant_bits<T, NP, P, I>()>{});
        return with_variant<T, NP, P, I + 1>(idx, f);
    } else {
        return S2D_ERR_INVALID_ARG;
    }
}

template <typename T, int NP>
int fused_steps_wave(s2d_ctx *c, int nsteps, double nu, double dt, double cfl)
{
    cpx<T> *Wn = (cpx<T> *)c->Wn, *Ws = (cpx<T> *)c->Ws;
    StageArgs a;
    a.nu = nu;
    a.dt = dt;
    a.cfl = cfl;
    a.dx = 2.0 * M_PI / (double)c->N;
    a.norm = (double)c->g.nxG * (double)c->g.nyG;
    a.inv_norm = 1.0 / a.norm;
    for (int st = 0; st < nsteps; ++st) {
        prof_next_step(c);
        for (int stage = 1; stage <= 3; ++stage) {
            const cpx<T> *src = stage == 1 ? Wn : Ws;
            cpx<T> *dst = stage == 3 ? Wn : Ws;
            a.stage = stage;
            {
                ProfScope ps(c, S2D_KERNEL_PASS_A);
                S2D_TRY((with_variant<T, NP, 0>(c->var[0], [&](auto v) {
                    return launch2_passA<T, NP, decltype(v)::value>(c, src, stage == 1);
                })));
            }
            {
                ProfScope ps(c, S2D_KERNEL_PASS_B);
                S2D_TRY((with_variant<T, NP, 1>(c->var[1], [&](auto v) {
                    return launch2_passB<T, NP, decltype(v)::value>(c, 1, stage == 1);
                })));
            }
            {
                ProfScope ps(c, S2D_KERNEL_PASS_C);
                S2D_TRY((with_variant<T, NP, 2>(c->var[2], [&](auto v) {
                    return launch2_passC<T, NP, decltype(v)::value>(c, src, dst, a);
                })));
            }
        }
    }
    return S2D_OK;
}

// Padded sizes with compiled lane-group FFT kernels; 0 when the generic
// kernels are used.
template <typename T>
int wave_np(const s2d_ctx *c)
{
    if (!c->wave_kernels) return 0;
    const int np = c->g.nxG;
    switch (np) {
    case 96: case 192: case 384: case 768: case 1536: case 3072: return np;
    }
    return 0;
}

#define S2D_WAVE_SWITCH(T, np, CALL)                            \
    switch (np) {                                               \
    case 96: { constexpr int NP = 96; CALL; } break;            \
    case 192: { constexpr int NP = 192; CALL; } break;          \
    case 384: { constexpr int NP = 384; CALL; } break;          \
    case 768: { constexpr int NP = 768; CALL; } break;          \
    case 1536: { constexpr int NP = 1536; CALL; } break;        \
    case 3072: { constexpr int NP = 3072; CALL; } break;        \
    }

template <typename T>
int fused_steps(s2d_ctx *c, int nsteps, double nu, double dt, double cfl)
{
    if (const int np = wave_np<T>(c)) {
        S2D_WAVE_SWITCH(T, np, return (fused_steps_wave<T, NP>(c, nsteps, nu, dt, cfl)))
    }
    const int EC = c->col.E, ER = c->rowf.E;
#define S2D_DISPATCH(ec, er) \
    if (EC == ec && ER == er) return fused_steps_E<T, ec, er>(c, nsteps, nu, dt, cfl);
    S2D_DISPATCH(12, 12)
    S2D_DISPATCH(6, 6)
    S2D_DISPATCH(8, 8)
    S2D_DISPATCH(4, 4)
    S2D_DISPATCH(3, 3)
    S2D_DISPATCH(2, 2)
#undef S2D_DISPATCH
    return fail(c, S2D_ERR_UNSUPPORTED, "fused step has no kernel for E=(%d,%d)", EC, ER);
}

template <typename T>
int fused_vmax(s2d_ctx *c, double *vmax)
{
    const int EC = c->col.E, ER = c->rowf.E;
    if (EC != ER) return fail(c, S2D_ERR_UNSUPPORTED, "no fused kernel for E=(%d,%d)", EC, ER);
    int r = S2D_ERR_UNSUPPORTED;
    const cpx<T> *Wn = (const cpx<T> *)c->Wn;
    if (const int np = wave_np<T>(c)) {
        S2D_WAVE_SWITCH(T, np, r = (launch2_passA<T, NP>(c, Wn, 1));
                        if (!r) r = (launch2_passB<T, NP>(c, 0, 1)))
    } else switch (EC) {
    case 12: r = launch_passA<T, 12>(c, Wn, 1); if (!r) r = launch_passB<T, 12>(c, 0, 1); break;
    case 8: r = launch_passA<T, 8>(c, Wn, 1); if (!r) r = launch_passB<T, 8>(c, 0, 1); break;
    case 6: r = launch_passA<T, 6>(c, Wn, 1); if (!r) r = launch_passB<T, 6>(c, 0, 1); break;
    case 4: r = launch_passA<T, 4>(c, Wn, 1); if (!r) r = launch_passB<T, 4>(c, 0, 1); break;
    case 3: r = launch_passA<T, 3>(c, Wn, 1); if (!r) r = launch_passB<T, 3>(c, 0, 1); break;
    case 2: r = launch_passA<T, 2>(c, Wn, 1); if (!r) r = launch_passB<T, 2>(c, 0, 1); break;
    default: return fail(c, S2D_ERR_UNSUPPORTED, "no fused kernel for E=%d", EC);
    }
    if (r) return r;
    DevScalars sc;
    S2D_HIP(c, hipMemcpyAsync(&sc, c->d_sc, sizeof(sc), hipMemcpyDeviceToHost, c->stream));
    // (pass B of the sizes with CFL slots wrote those; their maximum as in
    // pass C's vmax_of)
    std::vector<unsigned long long> vs(c->vslot ? VSLOTS * VSTRIDE : 0);
    if (c->vslot)
        S2D_HIP(c, hipMemcpyAsync(vs.data(), c->vslot, vs.size() * sizeof(unsigned long long),
                                  hipMemcpyDeviceToHost, c->stream));
    S2D_HIP(c, hipStreamSynchronize(c->stream));
    unsigned long long b = sc.vmax_bits;
    for (size_t i = 0; i < vs.size(); i += VSTRIDE) b = vs[i] > b ? vs[i] : b;
    std::memcpy(vmax, &b, sizeof(double));
    return S2D_OK;
}

// Dispatch a templated op on the context's precision.
#define S2D_DISPATCH_T(c, fn, ...) \
    ((c)->precision == S2D_FP64 ? fn<double>(c, ##__VA_ARGS__) : fn<float>(c, ##__VA_ARGS__))

int sync(s2d_ctx *c)
{
    S2D_HIP(c, hipStreamSynchronize(c->stream));
    return S2D_OK;
}

// W / PSI slots behind the compact state: refresh them.
int ensure_ref(s2d_ctx *c)
{
    if (c->fused && c->ref_stale) {
        S2D_TRY(S2D_DISPATCH_T(c, op_unpack));
        c->ref_stale = false;
    }
    return S2D_OK;
}

int ensure_compact(s2d_ctx *c)
{
    if (c->fused && c->compact_stale) {
        S2D_TRY(ensure_ref(c));
        S2D_TRY(S2D_DISPATCH_T(c, op_pack));
        c->compact_stale = false;
    }
    return S2D_OK;
}

void wrote(s2d_ctx *c, int f)
{
    if (f == S2D_FIELD_W) c->compact_stale = true;
}

// Diagnostics of the W slot (reference layout) with run r's time and step
// count; the caller has brought the W slot up to date.
int diag_of(s2d_ctx *c, int run, s2d_diag *out)
{
    const size_t n = nc_of(c);
    int blocks = grid_blocks(n, 256);
    if (blocks > c->n_red) blocks = c->n_red;
    if (c->precision == S2D_FP64)
        hipLaunchKernelGGL(k_diag_partial<double>, dim3(blocks), dim3(256), 0, c->stream,
                           F<double>(c, S2D_FIELD_W), c->g.cISize, c->g.nyG, c->d_kx, c->d_ky,
                           c->g.kxMG, c->d_red);
    else
        hipLaunchKernelGGL(k_diag_partial<float>, dim3(blocks), dim3(256), 0, c->stream,
                           F<float>(c, S2D_FIELD_W), c->g.cISize, c->g.nyG, c->d_kx, c->d_ky,
                           c->g.kxMG, c->d_red);
    S2D_LAUNCHED(c, "k_diag_partial");
    std::vector<double> part((size_t)blocks * 5);
    S2D_HIP(c, hipMemcpyAsync(part.data(), c->d_red, part.size() * sizeof(double),
                              hipMemcpyDeviceToHost, c->stream));
    DevScalars sc;
    S2D_HIP(c, hipMemcpyAsync(&sc, c->d_sc + run, sizeof(sc), hipMemcpyDeviceToHost, c->stream));
    S2D_TRY(sync(c));
    double a[5] = {0, 0, 0, 0, 0};
    for (int b = 0; b < blocks; ++b)
        for (int q = 0; q < 5; ++q) {
            double x = part[(size_t)b * 5 + q];
            if (q == 3)
                a[q] = x > a[q] ? x : a[q];
            else
                a[q] += x;
        }
    out->tke = a[0];
    out->enstrophy = a[1];
    out->palinstrophy = a[2];
    out->max_abs_w = a[3];
    out->time = sc.time;
    out->nadv = sc.nadv;
    if (a[4] > 0.0) return fail(c, S2D_ERR_NONFINITE, "non-finite vorticity in the state");
    return S2D_OK;
}

// ComputeSpectrum of the PSI slot (brought up to date by the caller):
// device shell sums, host mode counts and moments in shell order.
// The shell spectrum of the psi slot fa (fb < 0), or of the velocity
// spectra in slots fa (u) and fb (v).
int spectrum_of(s2d_ctx *c, double nu, double *ek, int *cnt, double *tke, double *eps,
                double *ediss, int fa, int fb)
{
    const s2d_grid &g = c->g;
    const int nb = s2d_spectrum_bins(&g);
    // the reduction buffer holds 6 n_red doubles (>= 3072: every shell count
    // up to N = 2048); larger grids take a buffer of their own
    double *d_ek = c->d_red, *own = nullptr;
    if ((size_t)nb > 6 * (size_t)c->n_red) {
        S2D_HIP(c, hipMalloc((void **)&own, sizeof(double) * nb));
        d_ek = own;
    }
    const int fv = fb < 0 ? fa : fb;
    if (c->precision == S2D_FP64) {
        if (fb < 0)
            hipLaunchKernelGGL((k_spectrum<double, false>), dim3(nb), dim3(256), 0, c->stream,
                               F<double>(c, fa), F<double>(c, fv), g.cISize, g.nyG, g.kxMG,
                               g.kyMG, c->xMaskBool, c->yMaskBool, d_ek);
        else
            hipLaunchKernelGGL((k_spectrum<double, true>), dim3(nb), dim3(256), 0, c->stream,
                               F<double>(c, fa), F<double>(c, fv), g.cISize, g.nyG, g.kxMG,
                               g.kyMG, c->xMaskBool, c->yMaskBool, d_ek);
    } else {
        if (fb < 0)
            hipLaunchKernelGGL((k_spectrum<float, false>), dim3(nb), dim3(256), 0, c->stream,
                               F<float>(c, fa), F<float>(c, fv), g.cISize, g.nyG, g.kxMG,
                               g.kyMG, c->xMaskBool, c->yMaskBool, d_ek);
        else
            hipLaunchKernelGGL((k_spectrum<float, true>), dim3(nb), dim3(256), 0, c->stream,
                               F<float>(c, fa), F<float>(c, fv), g.cISize, g.nyG, g.kxMG,
                               g.kyMG, c->xMaskBool, c->yMaskBool, d_ek);
    }
    const hipError_t el = hipGetLastError();
    std::vector<double> h((size_t)nb);
    const hipError_t ec =
        hipMemcpyAsync(h.data(), d_ek, sizeof(double) * nb, hipMemcpyDeviceToHost, c->stream);
    const hipError_t es = hipStreamSynchronize(c->stream);
    if (own) (void)hipFree(own);
    if (el != hipSuccess) return fail(c, S2D_ERR_HIP, "launch of k_spectrum failed");
    if (ec != hipSuccess || es != hipSuccess)
        return fail(c, S2D_ERR_HIP, "read-back of the spectrum failed");
    // mode counts of the padded spectrum, weights 1 / 2 (src/Analysis.F90:389-395)
    if (cnt) {
        for (int b = 0; b < nb; ++b) cnt[b] = 0;
        for (int j = 0; j < g.nyG; ++j)
            for (int i = 0; i < g.cISize; ++i) {
                const long long m = (long long)c->kx[i] * c->kx[i] + (long long)c->ky[j] * c->ky[j];
                long long b = (long long)std::sqrt((double)m);
                while (b * b > m) --b;
                if (m > b * b + b) ++b; // floor(sqrt(m) + 1/2)
                cnt[b] += (c->kx[i] == 0 || c->kx[i] == g.kxMG) ? 1 : 2;
            }
    }
    double t = 0.0, e = 0.0, z = 0.0; // src/Analysis.F90:439-448 (dk = 1)
    for (int b = 0; b < nb; ++b) {
        const double kb = (double)b;
        t = t + h[b];
        e = e + kb * kb * h[b];
        z = z + kb * kb * kb * kb * h[b];
        if (ek) ek[b] = h[b];
    }
    if (tke) *tke = t;
    if (eps) *eps = 2.0 * nu * e;
    if (ediss) *ediss = 2.0 * nu * z;
    return S2D_OK;
}

int check_ctx(const s2d_ctx *c)
{
    if (!c) return fail(nullptr, S2D_ERR_INVALID_ARG, "null context");
    return S2D_OK;
}

int check_fields(s2d_ctx *c, std::initializer_list<int> fs)
{
    for (int f : fs)
        if (!valid_field(f)) return fail(c, S2D_ERR_INVALID_ARG, "invalid field id %d", f);
    return S2D_OK;
}

int set_device(s2d_ctx *c)
{
    S2D_HIP(c, hipSetDevice(c->device));
    return S2D_OK;
}

// Closure coupling (s2d_closure_forcing / s2d_closure_steps) ------------
int check_closure(s2d_ctx *c, s2d_ctx *g, s2d_cnn *net)
{
    S2D_TRY(check_ctx(c));
    S2D_TRY(check_ctx(g));
    if (!net) return fail(c, S2D_ERR_INVALID_ARG, "null network");
    int nh = 0, nw = 0;
    if (s2d_cnn_grid(net, &nh, &nw) != S2D_OK) return fail(c, S2D_ERR_INVALID_ARG, "bad network");
    if (!c->fused || !c->wave_kernels || c->nruns != 1 || c->precision != S2D_FP32 ||
        g->precision != S2D_FP32 || g->g.dealias != S2D_NO_DEALIASING || g->g.nx != c->g.nx ||
        g->g.ny != c->g.ny || nh != g->g.nyG || nw != g->g.nxG || g->device != c->device)
        return fail(c, S2D_ERR_UNSUPPORTED,
                    "closure coupling: an FP32 fused run, an FP32 unpadded grid of its size and a "
                    "network of that grid");
    return S2D_OK;
}

// The closure forcing sequence on c's stream, no waits, in two halves around
// the network call. closure_inputs: (optionally run `run`'s compact state
// into W / PSI,) the cutoff sampling onto g, two c2r (the network's inputs in
// g's W and PSI). closure_output: g's TMP (the network's output) by r2c, the
// restriction back, the pack into run `run`'s forcing buffer, the shell guard.
int closure_restrict(s2d_ctx *c, s2d_ctx *src, int sf, s2d_ctx *dst, int df)
{
    const long long kc = c->g.nx / 2;
    const size_t n = nc_of(dst);
    hipLaunchKernelGGL((k_restrict<float, float>), dim3(grid_blocks(n, 256)), dim3(256), 0,
                       c->stream, F<float>(src, sf), src->g.cISize, src->g.nyG,
                       F<float>(dst, df), dst->g.cISize, dst->g.nyG, kc * kc, 0LL, 0LL);
    return hipGetLastError() == hipSuccess ? S2D_OK
                                           : fail(c, S2D_ERR_HIP, "launch of k_restrict failed");
}

int closure_inputs(s2d_ctx *c, s2d_ctx *g, int run, bool unpack)
{
    if (unpack) S2D_TRY(op_unpack<float>(c, run));
    hipStream_t gs = g->stream;
    g->stream = c->stream; // the whole sequence on one stream
    int r = closure_restrict(c, c, S2D_FIELD_W, g, S2D_FIELD_W);
    if (!r) r = closure_restrict(c, c, S2D_FIELD_PSI, g, S2D_FIELD_PSI);
    if (!r) r = op_c2r<float>(g, S2D_FIELD_W);
    if (!r) r = op_c2r<float>(g, S2D_FIELD_PSI);
    g->stream = gs;
    return r;
}

int closure_output(s2d_ctx *c, s2d_ctx *g, int run)
{
    hipStream_t gs = g->stream;
    g->stream = c->stream;
    int r = op_r2c<float>(g, S2D_FIELD_TMP);
    if (!r) r = closure_restrict(c, g, S2D_FIELD_TMP, c, S2D_FIELD_TMP);
    g->stream = gs;
    S2D_TRY(r);
    // the kept modes of TMP into the forcing buffer (as s2d_set_forcing)
    void *keep = c->Wn, *wslot = c->field[S2D_FIELD_W];
    c->Wn = c->forcing;
    c->field[S2D_FIELD_W] = c->field[S2D_FIELD_TMP];
    r = op_pack<float>(c, run);
    c->field[S2D_FIELD_W] = wslot;
    c->Wn = keep;
    if (!r && c->closure_guard) {
        const int nshell = (int)std::floor(std::sqrt(2.0) * (c->N / 2)) + 2;
        hipLaunchKernelGGL(k_shell_guard<float>, dim3(nshell), dim3(256), 0, c->stream,
                           (cpx<float> *)c->forcing + run * c->sW,
                           (const cpx<float> *)c->Wn + run * c->sW, c->N, c->K, c->KP, 0LL);
        if (hipGetLastError() != hipSuccess) r = fail(c, S2D_ERR_HIP, "launch of k_shell_guard failed");
    }
    return r;
}

int enqueue_closure_forcing(s2d_ctx *c, s2d_ctx *g, s2d_cnn *net, bool unpack)
{
    S2D_TRY(closure_inputs(c, g, 0, unpack));
    const int ld = g->g.rISize; // the real layout of the grid context
    if (s2d_cnn_forward_on(net, c->stream, 1, static_cast<const float *>(g->field[S2D_FIELD_W]),
                           static_cast<const float *>(g->field[S2D_FIELD_PSI]), 0, ld,
                           static_cast<float *>(g->field[S2D_FIELD_TMP]), 0, ld) != S2D_OK)
        return fail(c, S2D_ERR_HIP, "closure network: %s", s2d_cnn_last_error(net));
    return closure_output(c, g, 0);
}

void drop_closure_graphs(s2d_ctx *c)
{
    for (auto &gx : c->cl_graph) {
        if (gx) (void)hipGraphExecDestroy(gx);
        gx = nullptr;
    }
}

void drop_step_graph(s2d_ctx *c)
{
    if (c->step_graph) (void)hipGraphExecDestroy(c->step_graph);
    c->step_graph = nullptr;
}

// Multi-step calls of a single run with the compiled wave kernels replay
// one captured step (hipGraph): the step's 9 launches then go out without
// host round trips between them. The graph is re-captured when an argument
// baked into it changes (nu, dt, cfl, the forcing buffer, the schedule).
// Inside a timing session the sampled steps run eagerly with their events.
template <typename T>
int fused_steps_graphed(s2d_ctx *c, int nsteps, double nu, double dt, double cfl)
{
    const void *forcing = c->forcing_on ? c->forcing : nullptr;
    const bool fresh = c->step_graph && c->sg_nu == nu && c->sg_dt == dt && c->sg_cfl == cfl &&
                       c->sg_forcing == forcing && c->sg_var[0] == c->var[0] &&
                       c->sg_var[1] == c->var[1] && c->sg_var[2] == c->var[2];
    if (!fresh) {
        drop_step_graph(c);
        const bool prof = c->prof_on;
        c->prof_on = false; // no timing events inside the graph
        hipGraph_t g = nullptr;
        hipError_t eb = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
        int r = eb == hipSuccess ? fused_steps<T>(c, 1, nu, dt, cfl) : S2D_OK;
        const hipError_t ec = eb == hipSuccess ? hipStreamEndCapture(c->stream, &g) : eb;
        c->prof_on = prof;
        if (r != S2D_OK || ec != hipSuccess || !g) {
            if (g) (void)hipGraphDestroy(g);
            return r != S2D_OK ? r
                               : fail(c, S2D_ERR_HIP, "capture of the fused step failed: %s",
                                      hipGetErrorString(ec));
        }
        const hipError_t ei = hipGraphInstantiate(&c->step_graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            c->step_graph = nullptr;
            return fail(c, S2D_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(ei));
        }
        c->sg_nu = nu;
        c->sg_dt = dt;
        c->sg_cfl = cfl;
        c->sg_forcing = forcing;
        for (int k = 0; k < 3; ++k) c->sg_var[k] = c->var[k];
    }
    for (int st = 0; st < nsteps; ++st) {
        if (c->prof_on && c->prof_step % c->prof_every == 0) {
            S2D_TRY(fused_steps<T>(c, 1, nu, dt, cfl)); // a sampled step, with its events
        } else {
            if (c->prof_on) ++c->prof_step;
            S2D_HIP(c, hipGraphLaunch(c->step_graph, c->stream));
        }
    }
    return S2D_OK;
}


} // namespace

namespace {
int create_ctx(int nx, int ny, int dealias, int precision, int nruns, hipStream_t stream,
               s2d_ctx **out);
// the device address of s2d_diagnostics_noalloc's scratch (s2d_noalloc.inc)
int noalloc_setup(s2d_ctx *c);

// The seeded decaying-turbulence recipe (kept modes 0 < |k| <= kMax,
// E(k) ~ k^4 exp(-(k/k0)^2), |u_k|^2 = E/(pi k), phase 2 pi U(seed, kx, ky),
// Hermitian on kx = 0, TKE with the ComputeSpectrum weights scaled to ke).
// emit(i, j, re, im) receives every kept mode of the reference layout
// (i = kx, j = FFTW row), scaled, in row-major order, the order in which the
// TKE is summed: every caller gets the same bits.
template <typename Emit>
void turbulence_ic(const s2d_ctx *c, unsigned long long seed, double k0, double ke, Emit emit,
                   int ny)
{
    const double PI = std::acos(-1.0);
    struct Mode {
        int i, j;
        double re, im;
    };
    std::vector<Mode> modes;
    double tke = 0.0;
    const int kmax_i = c->g.cISize;
    for (int j = 0; j < ny; ++j) {
        const int ky = c->ky[j];
        for (int i = 0; i < kmax_i; ++i) {
            const int kx = c->kx[i];
            if (kx == 0 && ky == 0) continue;
            const double k2 = (double)kx * kx + (double)ky * ky;
            const double km = std::sqrt(k2);
            if (km > c->g.kMax + 1.0e-10) continue;
            int sy = ky;
            bool conj = false;
            if (kx == 0 && ky < 0) {
                sy = -ky;
                conj = true;
            }
            const double q = km / k0;
            const double E = k2 * k2 * std::exp(-q * q);
            const double au = std::sqrt(E / (PI * km)); // |u_k|
            const double th = 2.0 * PI * hash_uniform(seed, kx, sy);
            const double amp = au * km; // |w_k|
            double sn, cs; // glibc sincos: the same bits whichever compiler builds this
            ::sincos(th, &sn, &cs);
            modes.push_back({i, j, amp * cs, conj ? -(amp * sn) : amp * sn});
            const double wt = (kx == 0 || kx == c->g.kxMG) ? 1.0 : 2.0;
            tke += wt * 0.5 * au * au;
        }
    }
    const double s = (tke > 0.0) ? std::sqrt(ke / tke) : 0.0;
    for (const Mode &m : modes) emit(m.i, m.j, m.re * s, m.im * s);
}
}

// ================================================================ C ABI

extern "C" {

const char *s2d_last_error(const s2d_ctx *ctx) { return ctx ? ctx->err.c_str() : g_err.c_str(); }

int s2d_get_grid_size(int nx, int ny, int dealias, s2d_grid *g)
{
    if (!g) return fail(nullptr, S2D_ERR_INVALID_ARG, "null grid pointer");
    if (nx < 2 || ny < 2) return fail(nullptr, S2D_ERR_INVALID_ARG, "grid %dx%d too small", nx, ny);
    std::memset(g, 0, sizeof(*g));
    g->nx = nx;
    g->ny = ny;
    g->dealias = dealias;
    switch (dealias) {
    case S2D_NO_DEALIASING: // src/Spectral.F90:141-159
        g->nxG = nx;
        g->nyG = ny;
        break;
    case S2D_DEALIAS_3_2_RULE: // :160-179
        g->nxG = 3 * nx / 2;
        g->nyG = 3 * ny / 2;
        break;
    default: // :180-181 are empty in the reference
        return fail(nullptr, S2D_ERR_UNSUPPORTED,
                    "dealiasing method %d is not defined by the reference", dealias);
    }
    g->kxLG = 0;
    g->kyLG = -g->nyG / 2 + 1;
    g->kxMG = g->nxG / 2;
    g->kyMG = g->nyG / 2;
    g->kxLU = 0;
    g->kyLU = -ny / 2 + 1;
    g->kxMU = nx / 2;
    g->kyMU = ny / 2;
    g->kMax = (double)g->kxMU;
    g->rISize = 2 * (g->nxG / 2 + 1); // :189-190
    g->cISize = g->nxG / 2 + 1;
    // GetAllocSize with one process, src/Alloc.F90:88-121
    g->nxP = g->nxG;
    g->nyP = g->nyG;
    g->j1 = 1;
    g->j2 = g->nyG;
    g->localJ = 0;
    g->allocLocal = g->cISize * g->nyG;
    return S2D_OK;
}

int s2d_wavenumbers(const s2d_grid *g, int *kxG, int *kyG)
{
    if (!g || !kxG || !kyG) return fail(nullptr, S2D_ERR_INVALID_ARG, "null argument");
    for (int i = 0; i < g->cISize; ++i) kxG[i] = i;
    for (int i = 0; i < g->nyG; ++i) kyG[i] = (i > g->nyG / 2) ? i - g->nyG : i;
    return S2D_OK;
}

int s2d_create(int nx, int ny, int dealias, int precision, s2d_ctx **out)
{
    return create_ctx(nx, ny, dealias, precision, 1, nullptr, out);
}

} // extern "C"

namespace {

int create_ctx(int nx, int ny, int dealias, int precision, int nruns, hipStream_t stream,
               s2d_ctx **out)
{
    if (!out) return fail(nullptr, S2D_ERR_INVALID_ARG, "null output pointer");
    *out = nullptr;
    if (precision != S2D_FP64 && precision != S2D_FP32)
        return fail(nullptr, S2D_ERR_INVALID_ARG, "unknown precision %d", precision);
    if (nx % 2 || ny % 2)
        return fail(nullptr, S2D_ERR_UNSUPPORTED, "grid %dx%d: nx and ny must be even", nx, ny);
    s2d_grid g;
    S2D_TRY(s2d_get_grid_size(nx, ny, dealias, &g));
    if (g.nxG % 2)
        return fail(nullptr, S2D_ERR_UNSUPPORTED, "padded nx %d must be even", g.nxG);
    if (choose_elems(g.nxG / 2) == 0 || choose_elems(g.nyG) == 0 || choose_elems(g.nxG) == 0)
        return fail(nullptr, S2D_ERR_UNSUPPORTED,
                    "padded grid %dx%d: sizes must be 2^a 3^b with nxG/2 >= 2", g.nxG, g.nyG);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(nullptr, S2D_ERR_NO_DEVICE, "no GPU visible");

    if (nruns < 1) return fail(nullptr, S2D_ERR_INVALID_ARG, "run count must be positive");
    s2d_ctx *c = new s2d_ctx();
    c->g = g;
    c->precision = precision;
    c->nruns = nruns;
    c->celt = precision == S2D_FP64 ? 16 : 8;
    auto bail = [&](int code) {
        g_err = c->err;
        s2d_destroy(c);
        return code;
    };
    if (hipGetDevice(&c->device) != hipSuccess) {
        c->err = "hipGetDevice failed";
        return bail(S2D_ERR_HIP);
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->device) != hipSuccess) {
        c->err = "hipGetDeviceProperties failed";
        return bail(S2D_ERR_HIP);
    }
    c->max_lds = (int)prop.sharedMemPerBlock;
    if (noalloc_setup(c) != S2D_OK) return bail(S2D_ERR_HIP);
    {
        const char *e = std::getenv("S2D_GENERIC_KERNELS");
        c->wave_kernels = !(e && e[0] == '1');
    }
    if (stream) {
        c->stream = stream;
        c->own_stream = false;
    } else if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        c->err = "hipStreamCreate failed";
        return bail(S2D_ERR_HIP);
    }

    // wavenumbers, SetupMask (src/Spectral.F90:229-260), SetupTruncation (:272-313)
    c->kx.resize(g.cISize);
    c->ky.resize(g.nyG);
    s2d_wavenumbers(&g, c->kx.data(), c->ky.data());
    for (int i = 0; i < g.cISize; ++i)
        if (c->kx[i] == g.kxMG) {
            c->xMaskBool = 1;
            c->xMaskInd = i;
            break;
        }
    for (int j = 0; j < g.nyG; ++j)
        if (c->ky[j] == g.kyMG) {
            c->yMaskBool = 1;
            c->yMaskInd = j;
            break;
        }
    c->trunc_i1.assign(g.nyG, g.cISize);
    for (int j = 0; j < g.nyG; ++j) {
        const double ky = (double)c->ky[j];
        for (int i = 0; i < g.cISize; ++i) {
            const double kx = (double)c->kx[i];
            if (std::sqrt(kx * kx + ky * ky) > g.kMax + 1.0e-10) {
                c->trunc_i1[j] = i;
                break;
            }
        }
    }

    const size_t nc = (size_t)g.cISize * g.nyG;
    for (int f = 0; f < S2D_NFIELDS; ++f)
        if (hipMalloc(&c->field[f], nc * c->celt) != hipSuccess) {
            c->err = "device allocation of a field failed";
            return bail(S2D_ERR_ALLOC);
        }
    if (hipMalloc(&c->d_trunc, sizeof(int) * g.nyG) != hipSuccess ||
        hipMalloc(&c->d_kx, sizeof(int) * g.cISize) != hipSuccess ||
        hipMalloc(&c->d_ky, sizeof(int) * g.nyG) != hipSuccess ||
        hipMalloc(&c->d_sc, sizeof(DevScalars) * nruns) != hipSuccess ||
        hipMalloc(&c->d_nu, sizeof(double) * nruns) != hipSuccess ||
        (g.nxG >= VSLOT_MIN_NP &&
         (hipMalloc(&c->vslot, sizeof(unsigned long long) * VSLOTS * VSTRIDE * nruns) !=
              hipSuccess ||
          hipMemset(c->vslot, 0, sizeof(unsigned long long) * VSLOTS * VSTRIDE * nruns) !=
              hipSuccess))) {
        c->err = "device allocation of tables failed";
        return bail(S2D_ERR_ALLOC);
    }
    if (hipMemcpy(c->d_trunc, c->trunc_i1.data(), sizeof(int) * g.nyG, hipMemcpyHostToDevice) ||
        hipMemcpy(c->d_kx, c->kx.data(), sizeof(int) * g.cISize, hipMemcpyHostToDevice) ||
        hipMemcpy(c->d_ky, c->ky.data(), sizeof(int) * g.nyG, hipMemcpyHostToDevice)) {
        c->err = "upload of tables failed";
        return bail(S2D_ERR_HIP);
    }
    std::vector<DevScalars> sc0((size_t)nruns);
    std::memset(sc0.data(), 0, sizeof(DevScalars) * nruns);
    std::vector<double> nu0((size_t)nruns, 0.0);
    if (hipMemcpy(c->d_sc, sc0.data(), sizeof(DevScalars) * nruns, hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemcpy(c->d_nu, nu0.data(), sizeof(double) * nruns, hipMemcpyHostToDevice) !=
            hipSuccess) {
        c->err = "upload of scalars failed";
        return bail(S2D_ERR_HIP);
    }
    // SetupPlans, src/Spectral.F90:330-353
    int r;
    if ((r = make_devfft(c, g.nyG, &c->col)) || (r = make_devfft(c, g.nxG / 2, &c->rowh)) ||
        (r = make_devfft(c, g.nxG, &c->rowf)))
        return bail(r);
    c->tw_x = c->rowf.tw;

    // fused compact state (3/2 rule, square grids)
    // The fused kernels need: pass A = 3 column transforms per workgroup
    // (<= 1024 threads, 3 Np elements of LDS), pass B <= 512 threads.
    const bool fused_fits = 3 * c->col.plan.T <= 1024 && c->rowf.plan.T <= 512 &&
                            (size_t)3 * g.nyG * c->celt <= (size_t)c->max_lds &&
                            (size_t)g.nxG * c->celt <= (size_t)c->max_lds;
    if (dealias == S2D_DEALIAS_3_2_RULE && nx == ny && c->col.E == c->rowf.E && fused_fits) {
        c->fused = true;
        c->N = nx;
        c->K = nx / 2 + 1;
        c->KP = ((nx + 1) + 7) / 8 * 8;
        c->KG = (c->K + 15) / 16 * 16;
        c->sW = (long long)c->K * c->KP;  // compact column-major kept modes
        c->sG = (long long)g.nxG * c->KG; // G: Np rows of KG
        c->sH = (long long)c->K * g.nxG;  // H: K columns of Np
        const size_t ncomp = (size_t)c->sW * nruns * c->celt;
        const size_t nG = (size_t)c->sG * nruns * c->celt, nH = (size_t)c->sH * nruns * c->celt;
        if (hipMalloc(&c->Wn, ncomp) || hipMalloc(&c->Ws, ncomp)) {
            c->err = "device allocation of the compact state failed";
            return bail(S2D_ERR_ALLOC);
        }
        for (int f = 0; f < 3; ++f)
            if (hipMalloc(&c->G[f], nG)) {
                c->err = "device allocation of the mixed fields failed";
                return bail(S2D_ERR_ALLOC);
            }
        for (int f = 0; f < 2; ++f)
            if (hipMalloc(&c->H[f], nH)) {
                c->err = "device allocation of the mixed fields failed";
                return bail(S2D_ERR_ALLOC);
            }
    }
    if (nruns > 1 && !(c->fused && c->wave_kernels && c->g.nxG % 48 == 0 &&
                       (c->g.nxG / 48 & (c->g.nxG / 48 - 1)) == 0 && c->g.nxG <= 3072)) {
        c->err = "batched runs need a padded size with compiled kernels (N = 64..2048, powers of 2)";
        return bail(S2D_ERR_UNSUPPORTED);
    }
    c->n_red = 512 > nruns ? 512 : nruns;
    if (hipMalloc(&c->d_red, sizeof(double) * 6 * c->n_red)) {
        c->err = "device allocation of reduction buffers failed";
        return bail(S2D_ERR_ALLOC);
    }
    for (int f = 0; f < S2D_NFIELDS; ++f)
        if ((r = S2D_DISPATCH_T(c, op_fill, f))) return bail(r);
    if ((r = sync(c))) return bail(r);
    c->ref_stale = false;
    c->compact_stale = true;
    *out = c;
    return S2D_OK;
}

} // namespace

extern "C" {

int s2d_destroy(s2d_ctx *c)
{
    if (!c) return S2D_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (hipEvent_t e : c->prof_ev) (void)hipEventDestroy(e);
    drop_closure_graphs(c);
    drop_step_graph(c);
    for (int f = 0; f < S2D_NFIELDS; ++f) (void)hipFree(c->field[f]);
    void *ps[] = {c->d_trunc, c->d_kx, c->d_ky, c->d_sc, c->vslot, c->d_red, c->d_nu, c->forcing, c->col.tw,
                  c->rowh.tw, c->rowf.tw, c->Wn, c->Ws, c->G[0], c->G[1], c->G[2], c->H[0],
                  c->H[1]};
    for (void *p : ps) (void)hipFree(p);
    if (c->stream && c->own_stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return S2D_OK;
}

int s2d_get_grid(const s2d_ctx *c, s2d_grid *g)
{
    S2D_TRY(check_ctx(c));
    if (!g) return fail(const_cast<s2d_ctx *>(c), S2D_ERR_INVALID_ARG, "null grid pointer");
    *g = c->g;
    return S2D_OK;
}

int s2d_get_mask(const s2d_ctx *c, int *xb, int *xi, int *yb, int *yi)
{
    S2D_TRY(check_ctx(c));
    if (!xb || !xi || !yb || !yi)
        return fail(const_cast<s2d_ctx *>(c), S2D_ERR_INVALID_ARG, "null argument");
    *xb = c->xMaskBool;
    *xi = c->xMaskInd;
    *yb = c->yMaskBool;
    *yi = c->yMaskInd;
    return S2D_OK;
}

int s2d_get_truncation(const s2d_ctx *c, int *i1)
{
    S2D_TRY(check_ctx(c));
    if (!i1) return fail(const_cast<s2d_ctx *>(c), S2D_ERR_INVALID_ARG, "null argument");
    for (int j = 0; j < c->g.nyG; ++j) i1[j] = c->trunc_i1[j];
    return S2D_OK;
}

int s2d_set_field_host(s2d_ctx *c, int f, const double *buf)
{
    S2D_TRY(check_ctx(c));
    S2D_TRY(check_fields(c, {f}));
    if (!buf) return fail(c, S2D_ERR_INVALID_ARG, "null host buffer");
    S2D_TRY(set_device(c));
    S2D_TRY(ensure_ref(c));
    const size_t nd = 2 * nc_of(c);
    if (c->precision == S2D_FP64) {
        S2D_HIP(c, hipMemcpyAsync(c->field[f], buf, nd * sizeof(double), hipMemcpyHostToDevice,
                                  c->stream));
    } else {
        std::vector<float> tmp(nd);
        for (size_t k = 0; k < nd; ++k) tmp[k] = (float)buf[k];
        S2D_HIP(c, hipMemcpyAsync(c->field[f], tmp.data(), nd * sizeof(float),
                                  hipMemcpyHostToDevice, c->stream));
        S2D_TRY(sync(c));
    }
    wrote(c, f);
    return sync(c);
}

int s2d_get_field_host(s2d_ctx *c, int f, double *buf)
{
    S2D_TRY(check_ctx(c));
    S2D_TRY(check_fields(c, {f}));
    if (!buf) return fail(c, S2D_ERR_INVALID_ARG, "null host buffer");
    S2D_TRY(set_device(c));
    S2D_TRY(ensure_ref(c));
    const size_t nd = 2 * nc_of(c);
    if (c->precision == S2D_FP64) {
        S2D_HIP(c, hipMemcpyAsync(buf, c->field[f], nd * sizeof(double), hipMemcpyDeviceToHost,
                                  c->stream));
        return sync(c);
    }
    std::vector<float> tmp(nd);
    S2D_HIP(c, hipMemcpyAsync(tmp.data(), c->field[f], nd * sizeof(float), hipMemcpyDeviceToHost,
                              c->stream));
    S2D_TRY(sync(c));
    for (size_t k = 0; k < nd; ++k) buf[k] = (double)tmp[k];
    return S2D_OK;
}

int s2d_field_device_ptr(s2d_ctx *c, int f, void **ptr)
{
    S2D_TRY(check_ctx(c));
    S2D_TRY(check_fields(c, {f}));
    if (!ptr) return fail(c, S2D_ERR_INVALID_ARG, "null argument");
    *ptr = c->field[f];
    return S2D_OK;
}

int s2d_sync_fields(s2d_ctx *c)
{
    S2D_TRY(check_ctx(c));
    S2D_TRY(set_device(c));
    S2D_TRY(ensure_ref(c));
    return sync(c);
}

int s2d_zero_field(s2d_ctx *c, int f)
{
    S2D_TRY(check_ctx(c));
    S2D_TRY(check_fields(c, {f}));
    S2D_TRY(set_device(c));
    S2D_TRY(ensure_ref(c));
    S2D_TRY(S2D_DISPATCH_T(c, op_fill, f));
    wrote(c, f);
    return sync(c);
}

int s2d_copy_field(s2d_ctx *c, int src, int dst)
{
    S2D_TRY(check_ctx(c));
    S2D_TRY(check_fields(c, {src, dst}));
    S2D_TRY(set_device(c));
    S2D_TRY(ensure_ref(c));
    S2D_HIP(c, hipMemcpyAsync(c->field[dst], c->field[src], nc_of(c) * c->celt,
                              hipMemcpyDeviceToDevice, c->stream));
    wrote(c, dst);
    return sync(c);
}

#define S2D_OP_PROLOGUE(c, ...)              \
    S2D_TRY(check_ctx(c));                   \
    S2D_TRY(check_fields(c, {__VA_ARGS__})); \
    S2D_TRY(set_device(c));                  \
    S2D_TRY(ensure_ref(c))

int s2d_transform_r2c(s2d_ctx *c, int f)
{
    S2D_OP_PROLOGUE(c, f);
    S2D_TRY(S2D_DISPATCH_T(c, op_r2c, f));
    wrote(c, f);
    return sync(c);
}

int s2d_transform_c2r(s2d_ctx *c, int f)
{
    S2D_OP_PROLOGUE(c, f);
    S2D_TRY(S2D_DISPATCH_T(c, op_c2r, f));
    wrote(c, f);
    return sync(c);
}

int s2d_dudx(s2d_ctx *c, int in, int out)
{
    S2D_OP_PROLOGUE(c, in, out);
    S2D_TRY(S2D_DISPATCH_T(c, op_dudx, in, out));
    wrote(c, out);
    return sync(c);
}

int s2d_dudy(s2d_ctx *c, int in, int out)
{
    S2D_OP_PROLOGUE(c, in, out);
    S2D_TRY(S2D_DISPATCH_T(c, op_dudy, in, out));
    wrote(c, out);
    return sync(c);
}

int s2d_truncate(s2d_ctx *c, int f)
{
    S2D_OP_PROLOGUE(c, f);
    S2D_TRY(S2D_DISPATCH_T(c, op_truncate, f));
    wrote(c, f);
    return sync(c);
}

int s2d_compute_psi(s2d_ctx *c, int w, int psi)
{
    S2D_OP_PROLOGUE(c, w, psi);
    S2D_TRY(S2D_DISPATCH_T(c, op_psi, w, psi));
    wrote(c, psi);
    return sync(c);
}

int s2d_compute_velocity(s2d_ctx *c, int u, int v, int psi, int trans)
{
    S2D_OP_PROLOGUE(c, u, v, psi);
    S2D_TRY(S2D_DISPATCH_T(c, op_velocity, u, v, psi, trans));
    wrote(c, u);
    wrote(c, v);
    return sync(c);
}

int s2d_compute_vorticity(s2d_ctx *c, int u, int v, int w)
{
    S2D_OP_PROLOGUE(c, u, v, w);
    S2D_TRY(S2D_DISPATCH_T(c, op_vorticity, u, v, w));
    wrote(c, u);
    wrote(c, v);
    wrote(c, w);
    return sync(c);
}

int s2d_compute_rhs(s2d_ctx *c, double nu, int u, int v, int w, int psi, int dW)
{
    S2D_OP_PROLOGUE(c, u, v, w, psi, dW);
    S2D_TRY(S2D_DISPATCH_T(c, op_rhs, nu, u, v, w, psi, dW));
    wrote(c, u);
    wrote(c, v);
    wrote(c, w);
    wrote(c, dW);
    return sync(c);
}

int s2d_integrate_one_step(s2d_ctx *c, double nu, double dt)
{
    S2D_TRY(check_ctx(c));
    S2D_TRY(set_device(c));
    S2D_TRY(ensure_ref(c));
    S2D_TRY(S2D_DISPATCH_T(c, op_integrate, nu, dt));
    wrote(c, S2D_FIELD_W);
    return sync(c);
}

int s2d_rk3_steps(s2d_ctx *c, int nsteps, double nu, double dt, double cfl)
{
    S2D_TRY(check_ctx(c));
    if (nsteps < 0) return fail(c, S2D_ERR_INVALID_ARG, "negative step count");
    if (!(cfl > 0.0) && !(dt >= 0.0)) return fail(c, S2D_ERR_INVALID_ARG, "invalid dt");
    S2D_TRY(set_device(c));
    if (!c->fused) {
        for (int s = 0; s < nsteps; ++s) {
            double dts = dt;
            if (cfl > 0.0) S2D_TRY(s2d_cfl_dt(c, cfl, &dts));
            S2D_TRY(s2d_integrate_one_step(c, nu, dts));
        }
        return S2D_OK;
    }
    S2D_TRY(ensure_compact(c));
    if (nsteps == 0) return sync(c);
    const bool graphed = c->step_graph_on && c->nruns == 1 &&
                         (c->precision == S2D_FP64 ? wave_np<double>(c) : wave_np<float>(c)) != 0;
    if (graphed)
        S2D_TRY(S2D_DISPATCH_T(c, fused_steps_graphed, nsteps, nu, dt, cfl));
    else
        S2D_TRY(S2D_DISPATCH_T(c, fused_steps, nsteps, nu, dt, cfl));
    c->ref_stale = true;
    {
        const size_t n = (size_t)c->K * c->KP;
        if (c->precision == S2D_FP64)
            hipLaunchKernelGGL(k_check_finite<double>, dim3(grid_blocks(n, 256)), dim3(256), 0,
                               c->stream, (const cpx<double> *)c->Wn, n, (long long)n, c->d_sc);
        else
            hipLaunchKernelGGL(k_check_finite<float>, dim3(grid_blocks(n, 256)), dim3(256), 0,
                               c->stream, (const cpx<float> *)c->Wn, n, (long long)n, c->d_sc);
        S2D_LAUNCHED(c, "k_check_finite");
    }
    DevScalars sc;
    S2D_HIP(c, hipMemcpyAsync(&sc, c->d_sc, sizeof(sc), hipMemcpyDeviceToHost, c->stream));
    S2D_TRY(sync(c));
    double vmax;
    std::memcpy(&vmax, &sc.vmax_bits, sizeof(double));
    if (sc.bad || !(vmax == vmax) || vmax > 1.0e300)
        return fail(c, S2D_ERR_NONFINITE,
                    "non-finite values in the state (max |u|+|v| at the last step start = %g)",
                    vmax);
    // (dt = 0 is a run that has reached its end time)
    if (cfl > 0.0 && !(sc.dt > 0.0 && sc.dt < 1.0e300) && !(sc.t_end > 0.0 && sc.dt == 0.0))
        return fail(c, S2D_ERR_NONFINITE, "CFL time step collapsed (dt = %g)", sc.dt);
    return S2D_OK;
}

int s2d_rk3_step(s2d_ctx *c, double nu, double dt) { return s2d_rk3_steps(c, 1, nu, dt, 0.0); }

int s2d_cfl_dt(s2d_ctx *c, double cfl, double *dt)
{
    S2D_TRY(check_ctx(c));
    if (!dt) return fail(c, S2D_ERR_INVALID_ARG, "null argument");
    S2D_TRY(set_device(c));
    double vmax = 0.0;
    if (c->fused) {
        S2D_TRY(ensure_compact(c));
        S2D_TRY(S2D_DISPATCH_T(c, fused_vmax, &vmax));
    } else {
        S2D_TRY(ensure_ref(c));
        // physical u, v of PSI into the U and V slots (the reference's scratch)
        S2D_TRY(S2D_DISPATCH_T(c, op_velocity, S2D_FIELD_U, S2D_FIELD_V, S2D_FIELD_PSI, 1));
        S2D_TRY(sync(c));
        const size_t nd = 2 * nc_of(c);
        std::vector<double> u(nd), v(nd);
        S2D_TRY(s2d_get_field_host(c, S2D_FIELD_U, u.data()));
        S2D_TRY(s2d_get_field_host(c, S2D_FIELD_V, v.data()));
        for (int j = 0; j < c->g.nyG; ++j)
            for (int i = 0; i < c->g.nxG; ++i) {
                size_t k = (size_t)i + (size_t)j * c->g.rISize;
                double s = std::fabs(u[k]) + std::fabs(v[k]);
                if (s > vmax || s != s) vmax = s;
            }
    }
    if (!(vmax == vmax) || vmax > 1.0e300)
        return fail(c, S2D_ERR_NONFINITE, "non-finite velocity in the state");
    const double dx = 2.0 * M_PI / (double)c->g.nx;
    *dt = vmax > 0.0 ? cfl * dx / vmax : 0.0;
    return S2D_OK;
}

int s2d_diagnostics(s2d_ctx *c, s2d_diag *out)
{
    S2D_TRY(check_ctx(c));
    if (!out) return fail(c, S2D_ERR_INVALID_ARG, "null argument");
    S2D_TRY(set_device(c));
    S2D_TRY(ensure_ref(c));
    return diag_of(c, 0, out);
}

int s2d_spectrum_bins(const s2d_grid *g)
{
    if (!g) return -1;
    const double kMax = std::sqrt((double)g->kxMG * g->kxMG + (double)g->kyMG * g->kyMG);
    return (int)(kMax + 1.0) + 1; // src/Analysis.F90:326-335
}

int s2d_spectrum(s2d_ctx *c, double nu, double *ek, int *cnt, double *tke, double *eps,
                 double *ediss)
{
    S2D_TRY(check_ctx(c));
    S2D_TRY(set_device(c));
    S2D_TRY(ensure_ref(c));
    return spectrum_of(c, nu, ek, cnt, tke, eps, ediss, S2D_FIELD_PSI, -1);
}

int s2d_spectrum_uv(s2d_ctx *c, int u, int v, double nu, double *ek, int *cnt, double *tke,
                    double *eps, double *ediss)
{
    S2D_OP_PROLOGUE(c, u, v);
    return spectrum_of(c, nu, ek, cnt, tke, eps, ediss, u, v);
}

int s2d_restrict(s2d_ctx *src, int sf, s2d_ctx *dst, int df, double kcut)
{
    S2D_TRY(check_ctx(src));
    S2D_TRY(check_ctx(dst));
    S2D_TRY(check_fields(src, {sf}));
    S2D_TRY(check_fields(dst, {df}));
    if (!(kcut >= 0.0)) return fail(dst, S2D_ERR_INVALID_ARG, "kcut must be >= 0");
    if (src->device != dst->device)
        return fail(dst, S2D_ERR_UNSUPPORTED, "source and destination on different devices");
    S2D_TRY(set_device(src));
    S2D_TRY(ensure_ref(src));
    S2D_TRY(sync(src));
    S2D_TRY(ensure_ref(dst));
    const long long kc = (long long)std::floor(kcut + 1.0e-10);
    const size_t n = nc_of(dst);
    const int sc = src->g.cISize, sn = src->g.nyG, dc = dst->g.cISize, dn = dst->g.nyG;
    const dim3 grid(grid_blocks(n, 256)), blk(256);
    if (src->precision == S2D_FP64 && dst->precision == S2D_FP64)
        hipLaunchKernelGGL((k_restrict<double, double>), grid, blk, 0, dst->stream,
                           F<double>(src, sf), sc, sn, F<double>(dst, df), dc, dn, kc * kc,
                           0LL, 0LL);
    else if (src->precision == S2D_FP64)
        hipLaunchKernelGGL((k_restrict<double, float>), grid, blk, 0, dst->stream,
                           F<double>(src, sf), sc, sn, F<float>(dst, df), dc, dn, kc * kc,
                           0LL, 0LL);
    else if (dst->precision == S2D_FP64)
        hipLaunchKernelGGL((k_restrict<float, double>), grid, blk, 0, dst->stream,
                           F<float>(src, sf), sc, sn, F<double>(dst, df), dc, dn, kc * kc,
                           0LL, 0LL);
    else
        hipLaunchKernelGGL((k_restrict<float, float>), grid, blk, 0, dst->stream,
                           F<float>(src, sf), sc, sn, F<float>(dst, df), dc, dn, kc * kc,
                           0LL, 0LL);
    S2D_LAUNCHED(dst, "k_restrict");
    wrote(dst, df);
    return sync(dst);
}

int s2d_axpby(s2d_ctx *c, double a, int x, double b, int y)
{
    S2D_OP_PROLOGUE(c, x, y);
    const size_t n = nc_of(c);
    if (c->precision == S2D_FP64)
        hipLaunchKernelGGL(k_axpby<double>, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream,
                           F<double>(c, x), F<double>(c, y), a, b, n);
    else
        hipLaunchKernelGGL(k_axpby<float>, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream,
                           F<float>(c, x), F<float>(c, y), a, b, n);
    S2D_LAUNCHED(c, "k_axpby");
    wrote(c, y);
    return sync(c);
}

int s2d_set_forcing(s2d_ctx *c, int f)
{
    S2D_OP_PROLOGUE(c, f);
    if (!c->fused || !c->wave_kernels || c->nruns != 1)
        return fail(c, S2D_ERR_UNSUPPORTED,
                    "a forcing term needs the fused step of a single-run context");
    if (!c->forcing)
        S2D_HIP(c, hipMalloc(&c->forcing, (size_t)c->sW * c->celt));
    // the kept modes of slot f in the compact layout (the pack of W, applied
    // to slot f through a scratch swap of the compact pointer)
    void *keep = c->Wn;
    c->Wn = c->forcing;
    void *wslot = c->field[S2D_FIELD_W];
    c->field[S2D_FIELD_W] = c->field[f];
    const int r = S2D_DISPATCH_T(c, op_pack, 0);
    c->field[S2D_FIELD_W] = wslot;
    c->Wn = keep;
    S2D_TRY(r);
    c->forcing_on = true;
    c->cl_age = 0;
    return sync(c);
}

int s2d_closure_forcing(s2d_ctx *c, s2d_ctx *g, s2d_cnn *net)
{
    S2D_TRY(check_closure(c, g, net));
    S2D_TRY(set_device(c));
    S2D_TRY(ensure_ref(c));
    S2D_TRY(ensure_ref(g));
    S2D_TRY(sync(g)); // nothing of g's own stream may still run
    if (!c->forcing) S2D_HIP(c, hipMalloc(&c->forcing, (size_t)c->sW * c->celt));
    S2D_TRY(enqueue_closure_forcing(c, g, net, false));
    c->forcing_on = true;
    c->cl_age = 0;
    return sync(c);
}

int s2d_closure_steps(s2d_ctx *c, s2d_ctx *g, s2d_cnn *net, int nsteps, double nu, double dt,
                      double cfl, int every)
{
    S2D_TRY(check_closure(c, g, net));
    if (nsteps < 0 || every < 1) return fail(c, S2D_ERR_INVALID_ARG, "bad step count or period");
    if (!(cfl > 0.0) && !(dt >= 0.0)) return fail(c, S2D_ERR_INVALID_ARG, "invalid dt");
    S2D_TRY(set_device(c));
    S2D_TRY(ensure_compact(c));
    S2D_TRY(sync(g));
    if (!c->forcing) S2D_HIP(c, hipMalloc(&c->forcing, (size_t)c->sW * c->celt));
    c->forcing_on = true;
    if (cfl > 0.0) {
        const bool fresh =
            c->cl_net == net && c->cl_grid == g && c->cl_nu == nu && c->cl_cfl == cfl;
        if (!fresh) drop_closure_graphs(c);
        for (int k = 0; k < 2 && !c->cl_graph[k]; ++k) {
            // k = 0: forcing refresh + step; k = 1: step with the frozen forcing
            hipGraph_t gr = nullptr;
            S2D_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            int r = k == 0 ? enqueue_closure_forcing(c, g, net, true) : S2D_OK;
            if (!r) r = fused_steps<float>(c, 1, nu, 0.0, cfl);
            const hipError_t ec = hipStreamEndCapture(c->stream, &gr);
            if (r || ec != hipSuccess || !gr) {
                if (gr) (void)hipGraphDestroy(gr);
                return r ? r : fail(c, S2D_ERR_HIP, "capture of the closure step failed");
            }
            const hipError_t ei = hipGraphInstantiate(&c->cl_graph[k], gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            if (ei != hipSuccess) {
                c->cl_graph[k] = nullptr;
                return fail(c, S2D_ERR_HIP, "hipGraphInstantiate failed: %s",
                            hipGetErrorString(ei));
            }
        }
        c->cl_net = net;
        c->cl_grid = g;
        c->cl_nu = nu;
        c->cl_cfl = cfl;
    }
    // the refresh period runs on across calls and across CFL / fixed steps
    for (int st = 0; st < nsteps; ++st) {
        const bool refresh = c->cl_age >= every;
        if (cfl > 0.0) {
            S2D_HIP(c, hipGraphLaunch(c->cl_graph[refresh ? 0 : 1], c->stream));
        } else {
            if (refresh) S2D_TRY(enqueue_closure_forcing(c, g, net, true));
            S2D_TRY(fused_steps<float>(c, 1, nu, dt, 0.0));
        }
        c->cl_age = refresh ? 1 : c->cl_age + 1;
    }
    c->ref_stale = true;
    const size_t n = (size_t)c->K * c->KP;
    hipLaunchKernelGGL(k_check_finite<float>, dim3(grid_blocks(n, 256)), dim3(256), 0, c->stream,
                       (const cpx<float> *)c->Wn, n, (long long)n, c->d_sc);
    S2D_LAUNCHED(c, "k_check_finite");
    DevScalars sc;
    S2D_HIP(c, hipMemcpyAsync(&sc, c->d_sc, sizeof(sc), hipMemcpyDeviceToHost, c->stream));
    S2D_TRY(sync(c));
    double vmax;
    std::memcpy(&vmax, &sc.vmax_bits, sizeof(double));
    if (sc.bad || !(vmax == vmax) || vmax > 1.0e300)
        return fail(c, S2D_ERR_NONFINITE, "non-finite values in the state");
    return S2D_OK;
}

int s2d_set_closure_guard(s2d_ctx *c, int on)
{
    S2D_TRY(check_ctx(c));
    const bool g = on != 0;
    if (g != c->closure_guard) drop_closure_graphs(c);
    c->closure_guard = g;
    return S2D_OK;
}

int s2d_clear_forcing(s2d_ctx *c)
{
    S2D_TRY(check_ctx(c));
    c->forcing_on = false;
    c->cl_age = 1 << 30;
    return S2D_OK;
}

int s2d_get_time(s2d_ctx *c, double *time, long long *nadv, double *dt)
{
    S2D_TRY(check_ctx(c));
    S2D_TRY(set_device(c));
    DevScalars sc;
    S2D_HIP(c, hipMemcpyAsync(&sc, c->d_sc, sizeof(sc), hipMemcpyDeviceToHost, c->stream));
    S2D_TRY(sync(c));
    if (time) *time = sc.time;
    if (nadv) *nadv = sc.nadv;
    if (dt) *dt = sc.dt;
    return S2D_OK;
}

int s2d_set_end_time(s2d_ctx *c, double t_end)
{
    S2D_TRY(check_ctx(c));
    if (!(t_end < 1.0e300)) return fail(c, S2D_ERR_INVALID_ARG, "invalid end time");
    S2D_TRY(set_device(c));
    std::vector<DevScalars> sc((size_t)c->nruns);
    S2D_HIP(c, hipMemcpyAsync(sc.data(), c->d_sc, sizeof(DevScalars) * sc.size(),
                              hipMemcpyDeviceToHost, c->stream));
    S2D_TRY(sync(c));
    for (auto &s : sc) s.t_end = t_end > 0.0 ? t_end : 0.0;
    S2D_HIP(c, hipMemcpyAsync(c->d_sc, sc.data(), sizeof(DevScalars) * sc.size(),
                              hipMemcpyHostToDevice, c->stream));
    return sync(c);
}

int s2d_set_time(s2d_ctx *c, double time, long long nadv)
{
    S2D_TRY(check_ctx(c));
    S2D_TRY(set_device(c));
    DevScalars sc;
    S2D_HIP(c, hipMemcpyAsync(&sc, c->d_sc, sizeof(sc), hipMemcpyDeviceToHost, c->stream));
    S2D_TRY(sync(c));
    sc.time = time;
    sc.nadv = nadv;
    S2D_HIP(c, hipMemcpyAsync(c->d_sc, &sc, sizeof(sc), hipMemcpyHostToDevice, c->stream));
    return sync(c);
}

int s2d_set_turbulence(s2d_ctx *c, unsigned long long seed, double k0, double ke)
{
    S2D_TRY(check_ctx(c));
    if (!(k0 > 0.0) || !(ke >= 0.0)) return fail(c, S2D_ERR_INVALID_ARG, "k0 and ke must be positive");
    S2D_TRY(set_device(c));
    const int cI = c->g.cISize, ny = c->g.nyG;
    std::vector<double> w(2 * nc_of(c), 0.0);
    turbulence_ic(c, seed, k0, ke, [&](int i, int j, double re, double im) {
        const size_t k = 2 * ((size_t)i + (size_t)j * cI);
        w[k] = re;
        w[k + 1] = im;
    }, ny);
    S2D_TRY(s2d_set_field_host(c, S2D_FIELD_W, w.data()));
    S2D_TRY(S2D_DISPATCH_T(c, op_psi, S2D_FIELD_W, S2D_FIELD_PSI));
    return sync(c);
}

int s2d_set_taylor_green(s2d_ctx *c)
{
    S2D_TRY(check_ctx(c));
    S2D_TRY(set_device(c));
    const double PI = std::acos(-1.0);
    const int nxG = c->g.nxG, nyG = c->g.nyG, rI = c->g.rISize;
    const size_t nd = 2 * nc_of(c);
    std::vector<double> w(nd, 0.0), p(nd, 0.0);
    for (int j = 1; j <= nyG; ++j) // src/SetIC.F90:88-96 (1-based i, j)
        for (int i = 1; i <= nxG; ++i) {
            const double sx = std::sin(2.0 * PI * (double)i / (double)nxG);
            const double sy = std::sin(2.0 * PI * (double)j / (double)nyG);
            p[(size_t)(i - 1) + (size_t)(j - 1) * rI] = sx * sy;
            w[(size_t)(i - 1) + (size_t)(j - 1) * rI] = 2.0 * sx * sy;
        }
    S2D_TRY(s2d_set_field_host(c, S2D_FIELD_PSI, p.data()));
    S2D_TRY(s2d_set_field_host(c, S2D_FIELD_W, w.data()));
    S2D_TRY(s2d_transform_r2c(c, S2D_FIELD_PSI)); // :99-100
    return s2d_transform_r2c(c, S2D_FIELD_W);
}

int s2d_profile_begin(s2d_ctx *c, int max_launches)
{
    S2D_TRY(check_ctx(c));
    if (max_launches < 0 || max_launches > (1 << 20))
        return fail(c, S2D_ERR_INVALID_ARG, "max_launches out of range");
    S2D_TRY(set_device(c));
    const size_t need = 2 * (size_t)max_launches;
    while (c->prof_ev.size() < need) {
        hipEvent_t e;
        S2D_HIP(c, hipEventCreate(&e));
        c->prof_ev.push_back(e);
    }
    c->prof_kind.assign((size_t)max_launches, -1);
    c->prof_used = 0;
    c->prof_every = 1;
    c->prof_step = 0;
    c->prof_step_on = true;
    c->prof_on = true;
    return S2D_OK;
}

int s2d_profile_every(s2d_ctx *c, int every)
{
    S2D_TRY(check_ctx(c));
    if (every < 1) return fail(c, S2D_ERR_INVALID_ARG, "every must be >= 1");
    if (!c->prof_on) return fail(c, S2D_ERR_INVALID_ARG, "no timing session (s2d_profile_begin)");
    c->prof_every = every;
    return S2D_OK;
}

int s2d_profile_end(s2d_ctx *c, double *ms, long long *count)
{
    S2D_TRY(check_ctx(c));
    if (!ms || !count) return fail(c, S2D_ERR_INVALID_ARG, "null argument");
    S2D_TRY(set_device(c));
    S2D_TRY(sync(c));
    for (int k = 0; k < S2D_NKERNELS; ++k) {
        ms[k] = 0.0;
        count[k] = 0;
    }
    for (int u = 0; u < c->prof_used; ++u) {
        float t = 0.f;
        S2D_HIP(c, hipEventElapsedTime(&t, c->prof_ev[2 * u], c->prof_ev[2 * u + 1]));
        const int k = c->prof_kind[u];
        if (k >= 0 && k < S2D_NKERNELS) {
            ms[k] += (double)t;
            count[k] += 1;
        }
    }
    c->prof_on = false;
    c->prof_used = 0;
    c->prof_step_on = true;
    return S2D_OK;
}

} // extern "C"

#include "s2d_checkpoint.inc"
#include "s2d_ensemble.inc"
#include "s2d_noalloc.inc"
#include "s2d_tune.inc"
