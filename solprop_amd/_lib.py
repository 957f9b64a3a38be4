"""ctypes binding of the C ABI declared in ``include/solprop_mi355x.h``.

The library is loaded from the package directory (built in-tree by
``solprop_amd.build``). There is no fallback: a missing library raises
:class:`LibraryMissing`, so no caller can run the hot path anywhere but on the device.
The library is not checked against its sources here; ``solprop_amd.build`` compiles
every listed source each time it runs and refuses a list naming a missing file.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libsolprop_mi355x.so"

SP_OK, SP_VALUE, SP_BACKEND, SP_CONVERGENCE, SP_STATE = 0, 1, 2, 3, 4
SP_MAX_TERMS = 4


class LibraryMissing(RuntimeError):
    """The HIP library is not built: run ``python -m solprop_amd.build``."""


class ScalarPathC(C.Structure):
    _fields_ = [
        ("kind", C.c_int),
        ("value", C.c_double),
        ("n_intervals", C.c_int),
        ("degree", C.c_int),
        ("edges", C.POINTER(C.c_double)),
        ("coefficients", C.POINTER(C.c_double)),
    ]


_D = C.c_double
_I = C.c_int
_PD = C.POINTER(C.c_double)
_PU8 = C.POINTER(C.c_uint8)
_PI64 = C.POINTER(C.c_int64)


class ProblemDescC(C.Structure):
    _fields_ = [
        ("kind", _I), ("n_t", _I), ("n_to", _I), ("n_modes", _I), ("n_nodes", _I),
        ("n_omega", _I),
        ("crop_index", _PI64), ("active_mask", _PU8), ("norm_mask", _PU8),
        ("omega_window", _PD), ("omega", _PD), ("time_window", _PD),
        ("time_window_derivative", _PD), ("dt_fine", _D), ("z_end", _D),
        ("n_phi", _I), ("phi_terms", _PD), ("phi_paths", ScalarPathC * SP_MAX_TERMS),
        ("phi_rate_paths", ScalarPathC * SP_MAX_TERMS),
        ("n_prefactor", _I), ("prefactor_terms", _PD),
        ("prefactor_paths", ScalarPathC * SP_MAX_TERMS),
        ("synthesis", _PD), ("projection", _PD),
        ("density", ScalarPathC), ("preionisation", ScalarPathC),
        ("has_kerr", _I), ("kerr_real", _I), ("kerr_coefficient", _D), ("kerr_self", _D),
        ("kerr_conjugate", _D),
        ("n_raman", _I), ("raman_undamped", _PD), ("raman_fraction", _D),
        ("raman_epsilon0", _D), ("raman_thg", _I),
        ("raman_rates", ScalarPathC * SP_MAX_TERMS),
        ("has_plasma", _I), ("charge_mass_ratio", _D), ("ionisation_potential", _D),
        ("log_field_min", _D), ("d_log_field", _D), ("field_min", _D), ("field_max", _D),
        ("rate_max", _D), ("rate_intervals", _I), ("rate_coefficients", _PD),
        ("n_taps", _I), ("taps", _PD),
        ("n_stages", _I), ("tableau_a", _PD), ("tableau_c", _PD), ("tableau_b", _PD),
        ("tableau_e", _PD), ("first_same_as_last", _I),
        ("noise_model", _I), ("noise", _PD),
        ("rtol", _D), ("atol", _D), ("safety", _D), ("max_factor", _D), ("min_factor", _D),
        ("beta1", _D), ("beta2", _D),
        ("lag", _I), ("reject_free", _I), ("estimator_order", _I), ("auto_init", _I),
        ("dz_min", _D), ("dz_max", _D), ("stretch", _D), ("dz_init", _D),
        ("method", _I), ("adams_order", _I), ("adams_variable_order", _I),
        ("adams_equal_approach", _I), ("adams_safety", _D), ("adams_max_growth", _D),
        ("adams_min_factor", _D),
        ("absorber_scheme", _I), ("absorber_rate", _PD), ("spatial_absorber", _PD),
        ("n_schedule", _I), ("schedule_edges", _PD),
        ("grid_nx", _I), ("grid_ny", _I),
        ("projection_coarse", _PD),
    ]


class MarchResultC(C.Structure):
    _fields_ = [
        ("z", _D), ("proposal", _D), ("done", _I), ("n_rows", _I),
        ("attempts", C.c_long), ("accepted", C.c_long), ("evaluations", C.c_long),
    ]


class CountersC(C.Structure):
    _fields_ = [
        ("trials", C.c_long), ("evaluations", C.c_long), ("accepted", C.c_long),
        ("graph_launches", C.c_long), ("device_bytes", C.c_long),
    ]


class RKScenarioC(C.Structure):
    _fields_ = [
        ("rtol", _D), ("atol", _D), ("safety", _D), ("max_factor", _D), ("min_factor", _D),
        ("beta1", _D), ("beta2", _D),
        ("lag", _I), ("reject_free", _I), ("estimator_order", _I),
        ("z0", _D), ("proposal", _D), ("dz_min", _D), ("dz_max", _D), ("z_end", _D),
        ("stretch", _D),
        ("auto_init", _I), ("interval_cap", _I), ("step_planes", _I),
        ("has_h0", _I), ("h0", _D),
        ("norms", _PD), ("n_norms", C.c_long),
        ("save_planes", _PD), ("n_planes", _I),
        ("breakpoints", _PD), ("n_breakpoints", _I),
        ("kinks", _PD), ("n_kinks", _I),
        ("targets", _PD), ("n_targets", _I),
        ("schedule", _PD), ("n_schedule", C.c_long),
        ("quiet_keys", C.POINTER(C.c_long)), ("quiet_values", C.POINTER(C.c_int)),
        ("n_quiet", _I),
        ("has_quiet_interval", _I), ("quiet_interval", _D),
        ("has_ceiling", _I), ("ceiling", _D),
        ("split_keys", C.POINTER(C.c_long)), ("split_values", _PD), ("n_split", _I),
    ]


class AdamsScenarioC(C.Structure):
    _fields_ = [
        ("order", _I), ("variable_order", _I), ("equal_approach", _I), ("interval_cap", _I),
        ("safety", _D), ("max_growth", _D), ("min_factor", _D), ("collar_rate", _D),
        ("max_calls", C.c_long),
        ("z0", _D), ("proposal", _D), ("dz_min", _D), ("dz_max", _D), ("z_end", _D),
        ("stretch", _D),
        ("norms", _PD), ("n_attempts", C.c_long),
        ("save_planes", _PD), ("n_planes", _I),
        ("breakpoints", _PD), ("n_breakpoints", _I),
        ("kinks", _PD), ("n_kinks", _I),
        ("targets", _PD), ("n_targets", _I),
        ("schedule", _PD), ("n_schedule", C.c_long),
        ("quiet_keys", C.POINTER(C.c_long)), ("quiet_values", C.POINTER(C.c_int)),
        ("n_quiet", _I),
        ("has_quiet_interval", _I), ("quiet_interval", _D),
    ]


class ScanTimingC(C.Structure):
    _fields_ = [
        ("rounds", C.c_long), ("device_ms", _D), ("host_ms", _D), ("device_bytes", C.c_long),
    ]


class StatisticsRequestC(C.Structure):
    _fields_ = [("flags", C.c_ulonglong), ("windows", _PD), ("n_windows", _I),
                ("on_axis", _PD), ("on_axis_rows", _I), ("intensity_factor", _D)]


class SessionPlanC(C.Structure):
    _fields_ = [
        ("planes", _PD), ("n_planes", _I), ("launch", _I), ("step_planes", C.c_long),
        ("statistics", C.c_ulonglong), ("windows", _PD), ("n_windows", _I),
        ("attempt_capacity", C.c_long), ("call_attempts", C.c_long), ("call_seconds", _D),
    ]


class SessionEventC(C.Structure):
    _fields_ = [
        ("kind", _I), ("observation", _I), ("z", _D), ("plane_index", C.c_long),
        ("has_stats", _I), ("ticket", _I), ("generation", C.c_ulonglong), ("outcome", _I),
        ("reason", _I), ("category", _I), ("stage", _I), ("floor_proposal", _D),
        ("message", C.c_char * 256),
    ]


class SessionBatchC(C.Structure):
    _fields_ = [
        ("sequence", C.c_ulonglong), ("event", SessionEventC), ("n_attempts", C.c_long),
        ("n_stats", _I),
    ]


class SessionStatusC(C.Structure):
    _fields_ = [
        ("z", _D), ("plane_index", C.c_long), ("attempts", C.c_ulonglong),
        ("rejected", C.c_ulonglong), ("state", _I), ("outcome", _I),
        ("generation", C.c_ulonglong), ("march_calls", C.c_ulonglong),
        ("batches", C.c_ulonglong), ("exports", C.c_ulonglong),
        ("export_bytes", C.c_ulonglong),
    ]


class FnoShapeC(C.Structure):
    _fields_ = [("n_params", _I), ("width", _I), ("modes", _I), ("layers", _I), ("hidden", _I),
                ("n", _I)]


_PF = C.POINTER(C.c_float)

# The exported entry points and their signatures (include/solprop_mi355x.h).
SIGNATURES = {
    "sp_workspace_create": (_I, [C.POINTER(ProblemDescC), C.POINTER(C.c_void_p), C.c_char_p,
                                 C.c_size_t]),
    "sp_workspace_destroy": (None, [C.c_void_p]),
    "sp_workspace_load": (_I, [C.c_void_p, _D, _PD, C.c_char_p, C.c_size_t]),
    "sp_workspace_trial": (_I, [C.c_void_p, _D, _D, _PD, C.c_char_p, C.c_size_t]),
    "sp_workspace_trial_field": (_I, [C.c_void_p, _PD, C.c_char_p, C.c_size_t]),
    "sp_workspace_state": (_I, [C.c_void_p, _PD, C.c_char_p, C.c_size_t]),
    "sp_workspace_advance": (_I, [C.c_void_p, _D, C.c_long, _PD, C.c_long,
                                  C.POINTER(MarchResultC), C.c_char_p, C.c_size_t]),
    "sp_workspace_export": (_I, [C.c_void_p, _PD, _I, _PD, C.c_char_p, C.c_size_t]),
    "sp_workspace_counters": (_I, [C.c_void_p, C.POINTER(CountersC)]),
    "sp_workspace_profile": (_I, [C.c_void_p, _I]),
    "sp_workspace_timing": (_I, [C.c_void_p, _PD, C.POINTER(C.c_long)]),
    "sp_testing_source": (_I, [C.c_void_p, C.c_double, _PD, _PD, _PD, _PD, C.c_char_p,
                               C.c_size_t]),
    "sp_testing_buffer": (_I, [C.c_void_p, C.c_char_p, _PD, C.c_long, C.POINTER(C.c_long),
                               C.c_char_p, C.c_size_t]),
    "sp_workspace_phases": (_I, [C.c_void_p, _PD, _I, C.POINTER(_I), C.c_char_p, C.c_size_t]),
    "sp_batch_advance": (_I, [C.POINTER(C.c_void_p), _I, _PD, C.c_long, _PD, C.c_long,
                              C.c_void_p, C.POINTER(_I), C.c_char_p, C.c_size_t]),
    "sp_scan_create": (_I, [C.POINTER(ProblemDescC), _I, C.POINTER(C.c_void_p), C.c_char_p,
                            C.c_size_t]),
    "sp_scan_destroy": (None, [C.c_void_p]),
    "sp_scan_load": (_I, [C.c_void_p, _I, _D, _PD, C.c_char_p, C.c_size_t]),
    "sp_scan_advance": (_I, [C.c_void_p, _PD, C.c_long, _PD, C.c_long, C.c_void_p,
                             C.POINTER(_I), C.c_char_p, C.c_size_t]),
    "sp_scan_state": (_I, [C.c_void_p, _I, _PD, C.c_char_p, C.c_size_t]),
    "sp_scan_timing_of": (_I, [C.c_void_p, C.POINTER(ScanTimingC)]),
    "sp_scan_schedule": (_I, [C.c_void_p, C.c_char_p, C.c_size_t]),
    "sp_workspace_bind_monitor": (_I, [C.c_void_p, _PD, _PD, C.c_char_p, C.c_size_t]),
    "sp_workspace_monitor_reading": (_I, [C.c_void_p, _PD, C.c_char_p, C.c_size_t]),
    "sp_workspace_statistics": (_I, [C.c_void_p, C.c_ulonglong, _PD, _I, _PD, _I, C.POINTER(_I),
                                     C.c_char_p, C.c_size_t]),
    "sp_workspace_statistics_request": (_I, [C.c_void_p, C.POINTER(StatisticsRequestC), _PD, _I,
                                             C.POINTER(_I), C.c_char_p, C.c_size_t]),
    "sp_rk_trace": (_I, [C.POINTER(RKScenarioC), C.c_char_p, C.c_size_t]),
    "sp_adams_trace": (_I, [C.POINTER(AdamsScenarioC), C.c_char_p, C.c_size_t]),
    "sp_testing_fft": (_I, [_PD, _PD, C.c_long, C.c_long, _I, C.c_char_p, C.c_size_t]),
    "sp_testing_gemm": (_I, [C.c_long, C.c_long, C.c_long, _PD, _PD, _PD, C.c_char_p,
                             C.c_size_t]),
    "sp_fft_noalloc": (_I, [C.c_void_p, C.c_void_p, C.c_long, C.c_long, _I, C.c_void_p,
                            C.c_char_p, C.c_size_t]),
    "sp_session_start": (_I, [C.c_void_p, _D, _PD, C.POINTER(SessionPlanC),
                              C.POINTER(C.c_void_p), C.c_char_p, C.c_size_t]),
    "sp_session_advance": (_I, [C.c_void_p, C.c_long, _D, C.POINTER(SessionBatchC), _PD,
                                C.c_long, _PD, _I, C.c_char_p, C.c_size_t]),
    "sp_session_acknowledge": (_I, [C.c_void_p, C.c_ulonglong, C.c_char_p, C.c_size_t]),
    "sp_session_request_stop": (None, [C.c_void_p]),
    "sp_session_status_of": (_I, [C.c_void_p, C.POINTER(SessionStatusC)]),
    "sp_session_export": (_I, [C.c_void_p, C.c_ulonglong, C.c_ulonglong, _PD, _I, _PD,
                               C.c_char_p, C.c_size_t]),
    "sp_session_fail_outside": (_I, [C.c_void_p, C.c_char_p, C.POINTER(SessionBatchC), _PD,
                                     C.c_long, C.c_char_p, C.c_size_t]),
    "sp_session_close": (None, [C.c_void_p]),
    "sp_fno_create": (_I, [C.POINTER(FnoShapeC), _I, C.POINTER(C.c_void_p), C.c_char_p,
                           C.c_size_t]),
    "sp_fno_destroy": (None, [C.c_void_p]),
    "sp_fno_parameter_count": (C.c_long, [C.c_void_p]),
    "sp_fno_set_parameters": (_I, [C.c_void_p, _PF, C.c_char_p, C.c_size_t]),
    "sp_fno_get_parameters": (_I, [C.c_void_p, _PF, C.c_char_p, C.c_size_t]),
    "sp_fno_forward": (_I, [C.c_void_p, _PF, _PF, _I, _PF, C.c_char_p, C.c_size_t]),
    "sp_fno_gradient": (_I, [C.c_void_p, _PF, _PF, _PF, _I, _PD, _PF, C.c_char_p,
                             C.c_size_t]),
    "sp_fno_load_dataset": (_I, [C.c_void_p, _PF, _PF, _PF, C.c_long, C.c_char_p,
                                 C.c_size_t]),
    "sp_fno_train": (_I, [C.c_void_p, C.POINTER(_I), C.c_long, _I, C.c_float, _PF, _PD,
                          C.c_char_p, C.c_size_t]),
    "sp_fno_reset_optimizer": (_I, [C.c_void_p]),
    "sp_testing_gemm_bf16": (_I, [C.c_long, C.c_long, C.c_long, _PF, _PF, _PF, C.c_char_p,
                                  C.c_size_t]),
    "sp_version": (C.c_char_p, []),
}

_LIB = None


def lib():
    """Load the library once; raise :class:`LibraryMissing` when it is not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise LibraryMissing(
            f"{LIB_PATH} is not built; run `python -m solprop_amd.build` (hipcc, gfx950)"
        )
    handle = C.CDLL(str(LIB_PATH))
    for name, (restype, argtypes) in SIGNATURES.items():
        fn = getattr(handle, name)
        fn.restype = restype
        fn.argtypes = argtypes
    _LIB = handle
    return handle


def exported_symbols() -> list[str]:
    """The entry points this binding expects the library to export."""
    return list(SIGNATURES)
