"""ctypes binding of the C-ABI in include/paratc_gpu.h (libparatc_gpu.so).

There is no fallback: if the HIP library is missing or cannot be loaded this
module raises, naming the file it looked for.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libparatc_gpu.so")

# error codes
PTC_OK = 0
PTC_E_ARG = -1
PTC_E_HIP = -2
PTC_E_ALLOC = -3
PTC_E_STATE = -4
PTC_E_UNSUPPORTED = -5
PTC_E_LIMIT = -6

# field ids (src/m_Variables.f90:10-34)
(F_UX, F_UY, F_UZ, F_PR, F_SCALAR, F_HISTX, F_HISTY, F_HISTZ, F_HISTC, F_ARR1, F_ARR2, F_HALO,
 F_ARRC, F_WORK1, F_WORK2, F_UY_YM, F_DUY_YM, NFIELDS) = range(18)

L_HALO, L_INTERIOR, L_PLANE = 0, 1, 2
PREC_FP64, PREC_FP32_REFINE = 0, 1
TRI_AUTO, TRI_THOMAS, TRI_PARTITIONED = 0, 1, 2
# dump_visu fields (include/paratc_gpu.h PTC_VISU_*)
(VISU_UX, VISU_UY, VISU_UZ, VISU_PR, VISU_WX, VISU_WY, VISU_WZ, VISU_WMAG, VISU_SCALAR, VISU_Q,
 VISU_LAMBDA2) = range(11)
(POISSON_CHUNK_ROWS, POISSON_SLAB_PLANES, POISSON_TRI_LINES, POISSON_FAST_ZTILE, POISSON_TRI_LANES,
 POISSON_TRI_COMPLEX, POISSON_STREAMING, POISSON_PERSIST) = 1, 2, 3, 4, 5, 6, 7, 8
RHS_SEGMENT_THREADS, RHS_ROWS, RHS_XCD_RUN, RHS_MIN_PLANES = 1, 2, 3, 4
S_DIVMAX_PRSRC, S_PRGRADAVER, S_SIMTIME, S_DT = 0, 1, 2, 3

# every symbol declared in include/paratc_gpu.h
SEED_CHANNEL, SEED_FRONT = 0, 1  # ptc_seed_fields kinds

EXPORTS = [
    "ptc_create", "ptc_destroy", "ptc_last_error", "ptc_abi_version", "ptc_upload", "ptc_download",
    "ptc_fill", "ptc_copy", "ptc_field_device_ptr", "ptc_synchronize", "ptc_get_stream", "ptc_get_metric",
    "ptc_pm_coe_update", "ptc_get_coe", "ptc_set_coe", "ptc_init_poisson_solver", "ptc_clc_ppe",
    "ptc_destroy_poisson_plan", "ptc_poisson_config", "ptc_poisson_set", "ptc_poisson_get_knob", "ptc_rhs_set", "ptc_rhs_get", "ptc_poisson_get_config", "ptc_poisson_search_log", "ptc_poisson_residual",
    "ptc_poisson_stage_times", "ptc_init_time_scheme", "ptc_clc_rhs_xyz", "ptc_clc_u1hat",
    "ptc_clc_u2hat", "ptc_clc_u3hat", "ptc_clc_scalar", "ptc_clc_prsrc", "ptc_clc_prsrc_noalloc", "ptc_fluid_vel_update",
    "ptc_pressure_update", "ptc_set_bc_and_update_halo", "ptc_set_bc_and_update_halo_for_prsrc",
    "ptc_set_bc_and_update_halo_pr", "ptc_set_bc_and_update_halo_scalar", "ptc_update_uy_ym",
    "ptc_calc_max_cfl", "ptc_check_divergence", "ptc_calc_vmax", "ptc_calc_ux_aver",
    "ptc_clc_mean_value", "ptc_channel_iterate", "ptc_get_scalar", "ptc_clc_stat", "ptc_get_stat", "ptc_init_spectra", "ptc_clc_spectra",
    "ptc_get_spectra", "ptc_visu_field", "ptc_step_graphs",
    "ptc_ensemble_create", "ptc_ensemble_destroy", "ptc_case_footprint", "ptc_ensemble_admit",
    "ptc_ensemble_step", "ptc_ensemble_sync", "ptc_ensemble_case", "ptc_ensemble_release", "ptc_ensemble_usage",
    "ptc_ensemble_latency", "ptc_seed_normal", "ptc_seed_fields",
]


class PtcGrid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("pad_", C.c_int),
                ("xlx", C.c_double), ("yly", C.c_double), ("zlz", C.c_double)]


class PtcParams(C.Structure):
    _fields_ = [
        ("flow_type", C.c_int), ("scalar", C.c_int), ("is_ux_const", C.c_int), ("ischeme", C.c_int),
        ("icfl", C.c_int), ("sbc_opt", C.c_int * 2), ("pad_", C.c_int),
        ("ubulk", C.c_double), ("xnu", C.c_double), ("ucrf", C.c_double), ("dtmax", C.c_double),
        ("cflc", C.c_double), ("gravity", C.c_double * 3),
        ("uxbc", C.c_double * 6), ("uybc", C.c_double * 6), ("uzbc", C.c_double * 6),
        ("sdiff", C.c_double), ("gev", C.c_double * 3), ("fvv", C.c_double * 3),
        ("sbc_val", C.c_double * 4), ("vel_limit", C.c_double), ("div_limit", C.c_double),
    ]


class PtcError(RuntimeError):
    """A nonzero return code; the reference aborts here (src/m_LogInfo.f90:164-170)."""

    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where}: code {code}: {msg}")
        self.code = code


_lib = None


def load(path: str = LIB_PATH):
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"HIP library not built: {path} is missing (run __graft_entry__.build())")
    L = C.CDLL(path)
    vp, ip, dp = C.c_void_p, C.c_int, C.POINTER(C.c_double)
    L.ptc_last_error.restype = C.c_char_p
    L.ptc_create.argtypes = [C.POINTER(PtcGrid), C.POINTER(PtcParams), dp, C.POINTER(vp)]
    L.ptc_destroy.argtypes = [vp]
    L.ptc_upload.argtypes = [vp, ip, dp, ip]
    L.ptc_download.argtypes = [vp, ip, dp, ip]
    L.ptc_fill.argtypes = [vp, ip, C.c_double]
    L.ptc_copy.argtypes = [vp, ip, ip]
    L.ptc_field_device_ptr.argtypes = [vp, ip, C.POINTER(vp), C.POINTER(C.c_longlong)]
    L.ptc_synchronize.argtypes = [vp]
    L.ptc_get_stream.argtypes = [vp, C.POINTER(vp)]
    L.ptc_get_metric.argtypes = [vp, ip, dp]
    L.ptc_pm_coe_update.argtypes = [vp, ip, C.c_double]
    L.ptc_get_coe.argtypes = [vp, dp]
    L.ptc_set_coe.argtypes = [vp, dp]
    L.ptc_init_poisson_solver.argtypes = [vp, ip, ip]
    L.ptc_clc_ppe.argtypes = [vp, ip, ip]
    L.ptc_destroy_poisson_plan.argtypes = [vp]
    L.ptc_poisson_config.argtypes = [vp, ip, ip, ip, ip]
    L.ptc_poisson_residual.argtypes = [vp, ip, ip, dp]
    L.ptc_poisson_get_config.argtypes = [vp, C.POINTER(ip)]
    L.ptc_poisson_search_log.argtypes = [vp, C.c_char_p, ip, C.POINTER(ip)]
    L.ptc_poisson_set.argtypes = [vp, ip, ip]
    L.ptc_poisson_get_knob.argtypes = [vp, ip, C.POINTER(ip)]
    L.ptc_rhs_set.argtypes = [vp, ip, ip]
    L.ptc_rhs_get.argtypes = [vp, ip, C.POINTER(ip)]
    L.ptc_poisson_stage_times.argtypes = [vp, ip, C.POINTER(C.c_float), C.POINTER(C.c_int)]
    L.ptc_init_time_scheme.argtypes = [vp]
    L.ptc_clc_rhs_xyz.argtypes = [vp] + [ip] * 13
    L.ptc_clc_u1hat.argtypes = [vp, ip, ip]
    L.ptc_clc_u2hat.argtypes = [vp, ip, ip, ip]
    L.ptc_clc_u3hat.argtypes = [vp, ip, ip]
    L.ptc_clc_scalar.argtypes = [vp, ip, ip]
    L.ptc_clc_prsrc.argtypes = [vp, ip, ip, ip, ip, dp]
    L.ptc_clc_prsrc_noalloc.argtypes = [vp, ip, ip, ip, ip, dp]
    L.ptc_fluid_vel_update.argtypes = [vp, ip, ip, ip, ip]
    L.ptc_pressure_update.argtypes = [vp, ip, ip]
    L.ptc_set_bc_and_update_halo.argtypes = [vp, ip, ip, ip, ip]
    L.ptc_set_bc_and_update_halo_for_prsrc.argtypes = [vp, ip, ip, ip, ip]
    L.ptc_set_bc_and_update_halo_pr.argtypes = [vp, ip]
    L.ptc_set_bc_and_update_halo_scalar.argtypes = [vp, ip]
    L.ptc_update_uy_ym.argtypes = [vp, ip, ip]
    L.ptc_calc_max_cfl.argtypes = [vp, ip, ip, ip, dp]
    L.ptc_check_divergence.argtypes = [vp, ip, ip, ip, dp]
    L.ptc_calc_vmax.argtypes = [vp, ip, ip, ip, dp]
    L.ptc_calc_ux_aver.argtypes = [vp, ip, dp]
    L.ptc_clc_mean_value.argtypes = [vp, ip, ip, dp]
    L.ptc_channel_iterate.argtypes = [vp, ip, dp]
    L.ptc_get_scalar.argtypes = [vp, ip, dp]
    L.ptc_clc_stat.argtypes = [vp, ip, ip, ip, ip, ip]
    L.ptc_step_graphs.argtypes = [vp, ip]
    L.ptc_get_stat.argtypes = [vp, dp, C.POINTER(ip), C.POINTER(ip), dp, ip]
    L.ptc_visu_field.argtypes = [vp, ip, ip, ip, ip, ip, vp, C.c_longlong]
    L.ptc_init_spectra.argtypes = [vp, ip, ip, ip, ip, C.POINTER(ip)]
    L.ptc_clc_spectra.argtypes = [vp, ip, ip, ip, ip, ip]
    L.ptc_get_spectra.argtypes = [vp, dp, dp, dp, C.POINTER(ip), ip]
    L.ptc_seed_normal.argtypes = [vp, ip, ip, ip, ip]
    L.ptc_seed_fields.argtypes = [vp, ip, ip, C.c_double, C.c_double, C.c_double]
    sz, ll = C.c_size_t, C.c_longlong
    L.ptc_ensemble_create.argtypes = [sz, C.POINTER(vp)]
    L.ptc_ensemble_destroy.argtypes = [vp]
    L.ptc_case_footprint.argtypes = [vp, C.POINTER(sz)]
    L.ptc_ensemble_admit.argtypes = [vp, vp, C.POINTER(ip)]
    L.ptc_ensemble_step.argtypes = [vp, ip]
    L.ptc_ensemble_sync.argtypes = [vp]
    L.ptc_ensemble_case.argtypes = [vp, ip, C.POINTER(vp), C.POINTER(ll)]
    L.ptc_ensemble_release.argtypes = [vp, ip, C.POINTER(vp)]
    L.ptc_ensemble_usage.argtypes = [vp, C.POINTER(sz), C.POINTER(sz), C.POINTER(ip)]
    L.ptc_ensemble_latency.argtypes = [vp, ip, dp, dp, C.POINTER(ll)]
    for name in EXPORTS:
        getattr(L, name).restype = getattr(L, name).restype if name == "ptc_last_error" else C.c_int
    L.ptc_last_error.restype = C.c_char_p
    _lib = L
    return L


def check(code: int, where: str):
    if code != PTC_OK:
        msg = _lib.ptc_last_error().decode() if _lib is not None else ""
        raise PtcError(code, where, msg)
    return code


def params_from_case(case, vel_limit: float = 0.0, div_limit: float = 0.0):
    case.validate()
    g = PtcGrid(case.nx, case.ny, case.nz, 0, case.xlx, case.yly, case.zlz)
    p = PtcParams()
    p.flow_type = case.flow_type
    p.scalar = int(bool(case.scalar))
    p.is_ux_const = int(bool(case.is_ux_const))
    p.ischeme = case.ischeme
    p.icfl = case.icfl
    p.sbc_opt[0], p.sbc_opt[1] = case.scalar_bc_option
    p.ubulk, p.xnu, p.ucrf, p.dtmax, p.cflc = case.ubulk, case.xnu, case.ucrf, case.dtmax, case.cflc
    for q in range(3):
        p.gravity[q] = case.gravity[q]
    xb, yb, zb = case.bc_values()
    for q in range(6):
        p.uxbc[q], p.uybc[q], p.uzbc[q] = xb[q], yb[q], zb[q]
    p.sdiff = case.sdiff
    gev, fvv = case.scalar_vectors()
    for q in range(3):
        p.gev[q], p.fvv[q] = gev[q], fvv[q]
    for q in range(4):
        p.sbc_val[q] = case.scalar_bc_values[q]
    p.vel_limit, p.div_limit = vel_limit, div_limit
    return g, p
