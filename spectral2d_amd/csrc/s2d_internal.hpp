// s2d_internal.hpp -- context and shared declarations of libs2d.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "fft_block.hpp"
#include "s2d.h"

namespace s2d {

// Elements per thread for a length-n transform: the largest of
// 12, 8, 6, 4, 3, 2 that divides n and whose radices (a subset of
// {8, 4, 3, 2} dividing E) can factor n.
int choose_elems(int n);
// Radix plan for n with E elements per thread; false when n is not 2^a 3^b.
bool make_plan(int n, int E, FftPlan *p);

// A transform length with its plan and device twiddle table
// tw[e] = exp(-2 pi i e / n) (FP64 generation, stored as T).
struct DevFft {
    FftPlan plan{};
    int E = 0;
    void *tw = nullptr;
};

// Device scalars of a context (FP64 / integer, one small buffer).
struct DevScalars {
    unsigned long long vmax_bits; // max over the padded grid of |u|+|v| (double bits)
    double dt;                    // dt of the last step
    double time;                  // accumulated simulation time
    long long nadv;               // accumulated steps
    unsigned int bad;             // nonzero: a non-finite value was seen
    unsigned int pad32;
    double t_end;                 // > 0: every step's dt is clamped to end here
    double t_start;               // time at the start of the current step
    double pad;
};

} // namespace s2d

struct s2d_ctx {
    s2d_grid g{};
    int precision = S2D_FP64;
    size_t celt = 16; // bytes of one complex element on the device
    int device = 0;
    hipStream_t stream = nullptr;
    int max_lds = 65536;

    std::vector<int> kx, ky, trunc_i1;
    int xMaskBool = 0, xMaskInd = -1, yMaskBool = 0, yMaskInd = -1;

    void *field[S2D_NFIELDS] = {};
    int *d_trunc = nullptr; // per row: first truncated i (0-based)
    int *d_kx = nullptr;
    int *d_ky = nullptr;

    s2d::DevFft col;  // length nyG complex (columns)
    s2d::DevFft rowh; // length nxG/2 complex (rows, half-length real trick)
    s2d::DevFft rowf; // length nxG complex (rows, two-for-one in the fused pass)
    void *tw_x = nullptr; // exp(-2 pi i k / nxG), k = 0..nxG-1 (= rowf.tw)

    // a context may carry a batch of independent runs of one size (the
    // ensemble); run r's compact/G/H arrays start at base + r * stride and
    // its device scalars are d_sc[r]. The reference-layout slots serve one
    // run at a time.
    int nruns = 1;
    bool own_stream = true;
    long long sW = 0, sG = 0, sH = 0; // elements per run
    double *d_nu = nullptr;           // per-run viscosity (ensemble size classes)
    bool member = false;              // a size class of an ensemble
    void *forcing = nullptr;          // compact-layout RHS forcing (s2d_set_forcing)
    bool forcing_on = false;
    // captured closure steps (s2d_closure_steps): [0] forcing refresh + step,
    // [1] step; valid for the recorded grid, network, nu and cfl
    hipGraphExec_t cl_graph[2] = {nullptr, nullptr};
    void *cl_grid = nullptr, *cl_net = nullptr;
    double cl_nu = 0.0, cl_cfl = 0.0;
    bool closure_guard = false; // zero the closure forcing on backscattering shells
    // steps taken with the current forcing (s2d_closure_steps refreshes it
    // once this reaches its period); large: no forcing computed yet
    int cl_age = 1 << 30;

    // schedule of the fused step (s2d_set_schedule / s2d_tune): the variant
    // index of passes A, B, C among those compiled for the padded size, and
    // whether multi-step calls replay a captured step (hipGraph)
    int var[3] = {0, 0, 0};
    bool step_graph_on = true;
    hipGraphExec_t step_graph = nullptr;
    double sg_nu = 0.0, sg_dt = 0.0, sg_cfl = 0.0; // the captured step's arguments
    const void *sg_forcing = nullptr;              // its forcing buffer (or null)
    int sg_var[3] = {0, 0, 0};

    // fused compact state
    bool fused = false;
    bool wave_kernels = true; // compiled wave-FFT kernels for the hot sizes
    int N = 0, K = 0, KP = 0; // resolved size, kept kx count, compact column pitch
    int KG = 0;               // row pitch of the mixed G fields
    void *Wn = nullptr, *Ws = nullptr;
    void *G[3] = {}, *H[2] = {};
    s2d::DevScalars *d_sc = nullptr;
    // the CFL maximum of a plain context's row-pair pass B, in VSLOTS
    // slots per run, each on a line of its own (kernels_fused_v2.hpp)
    unsigned long long *vslot = nullptr;
    double *d_red = nullptr; // reduction partials
    int n_red = 0;
    bool ref_stale = false;     // W/PSI slots behind the compact state
    bool compact_stale = true;  // compact state behind the W slot

    // kernel timing (s2d_profile_begin / _end)
    bool prof_on = false;
    int prof_used = 0;
    int prof_every = 1;      // time one step in every prof_every
    long long prof_step = 0; // steps started since s2d_profile_begin
    bool prof_step_on = true;
    int prof_slot = -1; // event pair reserved for the next fused launch
    std::vector<hipEvent_t> prof_ev; // 2 per launch
    std::vector<int> prof_kind;

    std::string err;
};
