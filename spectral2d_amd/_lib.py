"""ctypes binding of libs2d.so (the C ABI declared in include/s2d.h).

The library is loaded from this package directory only. There is no
fallback: if libs2d.so is missing the import of the product API raises, so a
run can never silently compute on the CPU.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libs2d.so")

# status codes (include/s2d.h)
S2D_OK = 0
S2D_ERR_INVALID_ARG = 1
S2D_ERR_UNSUPPORTED = 2
S2D_ERR_HIP = 3
S2D_ERR_ALLOC = 4
S2D_ERR_NONFINITE = 5
S2D_ERR_NO_DEVICE = 6

NO_DEALIASING = 0
DEALIAS_3_2_RULE = 1
DEALIAS_2_3_RULE = 2
DEALIAS_GRID_SHIFT = 3
RK3_TVD_SHU = 1

FP64 = 0
FP32 = 1

FIELD_W, FIELD_PSI, FIELD_U, FIELD_V, FIELD_WS, FIELD_PSIS, FIELD_DW, FIELD_TMP = range(8)
NFIELDS = 8

# Every exported symbol of include/s2d.h.
EXPORTS = (
    "s2d_last_error", "s2d_get_grid_size", "s2d_wavenumbers", "s2d_create",
    "s2d_destroy", "s2d_get_grid", "s2d_get_mask", "s2d_get_truncation",
    "s2d_set_field_host", "s2d_get_field_host", "s2d_field_device_ptr",
    "s2d_sync_fields", "s2d_zero_field", "s2d_copy_field", "s2d_transform_r2c",
    "s2d_transform_c2r", "s2d_dudx", "s2d_dudy", "s2d_truncate",
    "s2d_compute_psi", "s2d_compute_velocity", "s2d_compute_vorticity",
    "s2d_compute_rhs", "s2d_integrate_one_step", "s2d_rk3_step", "s2d_rk3_steps",
    "s2d_cfl_dt", "s2d_diagnostics", "s2d_diagnostics_noalloc", "s2d_set_time",
    "s2d_set_end_time",
    "s2d_get_time", "s2d_profile_begin",
    "s2d_profile_every", "s2d_profile_end", "s2d_set_turbulence", "s2d_set_taylor_green",
    "s2d_spectrum_bins", "s2d_spectrum", "s2d_spectrum_uv", "s2d_ensemble_spectrum",
    "s2d_checkpoint_write", "s2d_checkpoint_read", "s2d_restrict", "s2d_axpby",
    "s2d_set_forcing", "s2d_clear_forcing",
    "s2d_cnn_last_error", "s2d_cnn_param_count", "s2d_cnn_create", "s2d_cnn_destroy",
    "s2d_cnn_set_params", "s2d_cnn_get_params", "s2d_cnn_set_scales", "s2d_cnn_forward",
    "s2d_cnn_gradient", "s2d_cnn_train_step", "s2d_cnn_set_dataset", "s2d_cnn_train_batch",
    "s2d_cnn_predict", "s2d_cnn_set_clip", "s2d_cnn_set_launch_width", "s2d_cnn_forward_on",
    "s2d_cnn_grid",
    "s2d_closure_forcing", "s2d_closure_steps", "s2d_set_closure_guard",
    "s2d_ensemble_last_error", "s2d_ensemble_create", "s2d_ensemble_destroy",
    "s2d_ensemble_size", "s2d_ensemble_run_grid", "s2d_ensemble_set_viscosity",
    "s2d_ensemble_set_end_time", "s2d_ensemble_closure_steps", "s2d_ensemble_clear_forcing",
    "s2d_ensemble_get_time",
    "s2d_ensemble_set_turbulence", "s2d_ensemble_set_vorticity", "s2d_ensemble_get_field",
    "s2d_ensemble_steps", "s2d_ensemble_diagnostics", "s2d_ensemble_set_schedule",
    "s2d_ensemble_profile_begin",
    "s2d_ensemble_profile_end",
    "s2d_schedule_variants", "s2d_get_schedule", "s2d_set_schedule", "s2d_tune",
    "s2d_ensemble_tune",
)

ENS_GRAPH = 1
ENS_UNIFIED = 0
ENS_PER_CLASS = 1
ENS_PER_CLASS_TUNED = 2
ENS_UNIFIED_FUSED_A = 3
ENS_STAGE_QUEUE = 4
TUNE_MEASURE = 1
TUNE_CACHE_ONLY = 2


class Grid(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("dealias", C.c_int),
        ("nxG", C.c_int), ("nyG", C.c_int),
        ("kxLG", C.c_int), ("kyLG", C.c_int), ("kxMG", C.c_int), ("kyMG", C.c_int),
        ("kxLU", C.c_int), ("kyLU", C.c_int), ("kxMU", C.c_int), ("kyMU", C.c_int),
        ("kMax", C.c_double),
        ("rISize", C.c_int), ("cISize", C.c_int),
        ("nxP", C.c_int), ("nyP", C.c_int), ("j1", C.c_int), ("j2", C.c_int),
        ("localJ", C.c_int), ("allocLocal", C.c_int),
    ]


class Diag(C.Structure):
    _fields_ = [
        ("tke", C.c_double), ("enstrophy", C.c_double), ("palinstrophy", C.c_double),
        ("max_abs_w", C.c_double), ("time", C.c_double), ("nadv", C.c_longlong),
    ]


class Schedule(C.Structure):
    _fields_ = [
        ("pass_a", C.c_int), ("pass_b", C.c_int), ("pass_c", C.c_int), ("step_graph", C.c_int),
        ("step_us", C.c_double), ("default_us", C.c_double),
        ("candidates", C.c_int), ("rejected", C.c_int), ("from_cache", C.c_int),
    ]


class S2DError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"libs2d error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    """Load libs2d.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m spectral2d_amd.build` "
            "(there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    vp, dp, ip = C.c_void_p, P(C.c_double), P(C.c_int)
    L.s2d_last_error.argtypes = [vp]
    L.s2d_last_error.restype = C.c_char_p
    L.s2d_get_grid_size.argtypes = [C.c_int, C.c_int, C.c_int, P(Grid)]
    L.s2d_wavenumbers.argtypes = [P(Grid), ip, ip]
    L.s2d_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, P(vp)]
    L.s2d_destroy.argtypes = [vp]
    L.s2d_get_grid.argtypes = [vp, P(Grid)]
    L.s2d_get_mask.argtypes = [vp, ip, ip, ip, ip]
    L.s2d_get_truncation.argtypes = [vp, ip]
    L.s2d_set_field_host.argtypes = [vp, C.c_int, dp]
    L.s2d_get_field_host.argtypes = [vp, C.c_int, dp]
    L.s2d_field_device_ptr.argtypes = [vp, C.c_int, P(vp)]
    L.s2d_sync_fields.argtypes = [vp]
    L.s2d_zero_field.argtypes = [vp, C.c_int]
    L.s2d_copy_field.argtypes = [vp, C.c_int, C.c_int]
    L.s2d_transform_r2c.argtypes = [vp, C.c_int]
    L.s2d_transform_c2r.argtypes = [vp, C.c_int]
    L.s2d_dudx.argtypes = [vp, C.c_int, C.c_int]
    L.s2d_dudy.argtypes = [vp, C.c_int, C.c_int]
    L.s2d_truncate.argtypes = [vp, C.c_int]
    L.s2d_compute_psi.argtypes = [vp, C.c_int, C.c_int]
    L.s2d_compute_velocity.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int]
    L.s2d_compute_vorticity.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    L.s2d_compute_rhs.argtypes = [vp, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
    L.s2d_integrate_one_step.argtypes = [vp, C.c_double, C.c_double]
    L.s2d_rk3_step.argtypes = [vp, C.c_double, C.c_double]
    L.s2d_rk3_steps.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_double]
    L.s2d_cfl_dt.argtypes = [vp, C.c_double, dp]
    L.s2d_diagnostics.argtypes = [vp, P(Diag)]
    L.s2d_set_time.argtypes = [vp, C.c_double, C.c_longlong]
    L.s2d_set_end_time.argtypes = [vp, C.c_double]
    L.s2d_get_time.argtypes = [vp, dp, P(C.c_longlong), dp]
    L.s2d_profile_begin.argtypes = [vp, C.c_int]
    L.s2d_profile_every.argtypes = [vp, C.c_int]
    L.s2d_profile_end.argtypes = [vp, dp, P(C.c_longlong)]
    L.s2d_set_turbulence.argtypes = [vp, C.c_ulonglong, C.c_double, C.c_double]
    L.s2d_set_taylor_green.argtypes = [vp]
    L.s2d_spectrum_bins.argtypes = [P(Grid)]
    L.s2d_restrict.argtypes = [vp, C.c_int, vp, C.c_int, C.c_double]
    L.s2d_axpby.argtypes = [vp, C.c_double, C.c_int, C.c_double, C.c_int]
    L.s2d_set_forcing.argtypes = [vp, C.c_int]
    L.s2d_clear_forcing.argtypes = [vp]
    L.s2d_checkpoint_write.argtypes = [vp, C.c_char_p]
    L.s2d_checkpoint_read.argtypes = [vp, C.c_char_p]
    L.s2d_spectrum.argtypes = [vp, C.c_double, dp, ip, dp, dp, dp]
    L.s2d_diagnostics_noalloc.argtypes = [vp, vp, vp, vp]
    L.s2d_spectrum_uv.argtypes = [vp, C.c_int, C.c_int, C.c_double, dp, ip, dp, dp, dp]
    L.s2d_ensemble_spectrum.argtypes = [vp, C.c_int, C.c_double, dp, ip, dp, dp, dp]
    L.s2d_ensemble_last_error.argtypes = [vp]
    L.s2d_ensemble_last_error.restype = C.c_char_p
    L.s2d_ensemble_create.argtypes = [C.c_int, ip, C.c_int, P(vp)]
    L.s2d_ensemble_destroy.argtypes = [vp]
    L.s2d_ensemble_size.argtypes = [vp, ip, ip]
    L.s2d_ensemble_run_grid.argtypes = [vp, C.c_int, P(Grid)]
    L.s2d_ensemble_set_viscosity.argtypes = [vp, dp]
    L.s2d_ensemble_set_end_time.argtypes = [vp, dp]
    L.s2d_ensemble_closure_steps.argtypes = [vp, vp, vp, C.c_int, C.c_double, C.c_double,
                                             C.c_int, C.c_int]
    L.s2d_ensemble_clear_forcing.argtypes = [vp]
    L.s2d_ensemble_get_time.argtypes = [vp, dp, C.POINTER(C.c_longlong)]
    L.s2d_ensemble_set_turbulence.argtypes = [vp, P(C.c_ulonglong), dp, dp]
    L.s2d_ensemble_set_vorticity.argtypes = [vp, C.c_int, dp]
    L.s2d_ensemble_get_field.argtypes = [vp, C.c_int, C.c_int, dp]
    L.s2d_ensemble_steps.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_int]
    L.s2d_ensemble_diagnostics.argtypes = [vp, P(Diag)]
    L.s2d_ensemble_set_schedule.argtypes = [vp, C.c_int]
    L.s2d_ensemble_profile_begin.argtypes = [vp, C.c_int]
    L.s2d_ensemble_profile_end.argtypes = [vp, dp, P(C.c_longlong)]
    fp, ll = P(C.c_float), C.c_longlong
    L.s2d_cnn_last_error.argtypes = [vp]
    L.s2d_cnn_last_error.restype = C.c_char_p
    L.s2d_cnn_param_count.argtypes = []
    L.s2d_cnn_create.argtypes = [C.c_int, C.c_int, C.c_int, P(vp)]
    L.s2d_cnn_destroy.argtypes = [vp]
    L.s2d_cnn_set_params.argtypes = [vp, fp]
    L.s2d_cnn_get_params.argtypes = [vp, fp]
    L.s2d_cnn_set_scales.argtypes = [vp, C.c_float, C.c_float, C.c_float]
    L.s2d_cnn_forward.argtypes = [vp, C.c_int, vp, vp, ll, C.c_int, vp, ll, C.c_int]
    L.s2d_cnn_gradient.argtypes = [vp, C.c_int, vp, vp, ll, C.c_int, vp, ll, C.c_int, dp, fp]
    L.s2d_cnn_train_step.argtypes = [vp, C.c_int, vp, vp, ll, C.c_int, vp, ll, C.c_int,
                                     C.c_float, dp]
    L.s2d_cnn_set_dataset.argtypes = [vp, ll, fp, fp, fp]
    L.s2d_cnn_train_batch.argtypes = [vp, C.c_int, ip, C.c_float, dp]
    L.s2d_cnn_predict.argtypes = [vp, ll, fp]
    L.s2d_cnn_set_clip.argtypes = [vp, C.c_int]
    L.s2d_cnn_set_launch_width.argtypes = [vp, C.c_int]
    L.s2d_cnn_forward_on.argtypes = [vp, vp, C.c_int, vp, vp, ll, C.c_int, vp, ll, C.c_int]
    L.s2d_cnn_grid.argtypes = [vp, ip, ip]
    L.s2d_closure_forcing.argtypes = [vp, vp, vp]
    L.s2d_set_closure_guard.argtypes = [vp, C.c_int]
    L.s2d_closure_steps.argtypes = [vp, vp, vp, C.c_int, C.c_double, C.c_double, C.c_double,
                                    C.c_int]
    L.s2d_schedule_variants.argtypes = [vp, ip]
    L.s2d_get_schedule.argtypes = [vp, P(Schedule)]
    L.s2d_set_schedule.argtypes = [vp, P(Schedule)]
    L.s2d_tune.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_char_p, C.c_int,
                           P(Schedule)]
    L.s2d_ensemble_tune.argtypes = [vp, C.c_double, C.c_double, C.c_char_p, C.c_int, ip, dp]
    for name in EXPORTS:
        if name == "s2d_cnn_param_count":
            L.s2d_cnn_param_count.restype = C.c_longlong
        elif name not in ("s2d_last_error", "s2d_ensemble_last_error", "s2d_cnn_last_error"):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def check(rc: int, ctx=None):
    if rc != S2D_OK:
        msg = lib().s2d_last_error(ctx)
        raise S2DError(rc, msg.decode() if msg else "")


def check_cnn(rc: int, net=None):
    if rc != S2D_OK:
        msg = lib().s2d_cnn_last_error(net)
        raise S2DError(rc, msg.decode() if msg else "")


def check_ens(rc: int, ens=None):
    if rc != S2D_OK:
        msg = lib().s2d_ensemble_last_error(ens)
        raise S2DError(rc, msg.decode() if msg else "")
