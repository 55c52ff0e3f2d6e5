"""ctypes mirror of include/antelope/sbe_c.h (structs and constants).

Kept in one place so the product binding (antelope_amd.solver) and the test
harness describe the boundary identically. Field order and types must match
the header; tests/test_abi.py checks sizes and offsets against a compiled
probe of the header.
"""
from __future__ import annotations

import ctypes as C

SBE_ABI_VERSION = 1

SBE_OK = 0
SBE_ERR_CONFIG = -1
SBE_ERR_MODEL = -2
SBE_ERR_OOM = -3
SBE_ERR_HIP = -4
SBE_ERR_NAN = -5
SBE_ERR_STATE = -6
SBE_ERR_RANGE = -7
SBE_ERR_UNSUPPORTED = -8
SBE_ERR_NODEVICE = -9

STATUS_NAMES = {
    0: "SBE_OK", -1: "SBE_ERR_CONFIG", -2: "SBE_ERR_MODEL", -3: "SBE_ERR_OOM", -4: "SBE_ERR_HIP",
    -5: "SBE_ERR_NAN", -6: "SBE_ERR_STATE", -7: "SBE_ERR_RANGE", -8: "SBE_ERR_UNSUPPORTED",
    -9: "SBE_ERR_NODEVICE",
}

SBE_ENV_GAUSS, SBE_ENV_GAUSS_CLIPPED, SBE_ENV_COS4, SBE_ENV_COS4_CLIPPED, SBE_ENV_SIN2 = range(5)
ENVELOPES = {
    "gauss": SBE_ENV_GAUSS,
    "gauss_clipped": SBE_ENV_GAUSS_CLIPPED,
    "cos4": SBE_ENV_COS4,
    "cos4_clipped": SBE_ENV_COS4_CLIPPED,
    "sin2": SBE_ENV_SIN2,
}

SBE_GAUGE_LENGTH_WANNIER, SBE_GAUGE_LENGTH_HAMILTONIAN, SBE_GAUGE_VELOCITY_HAMILTONIAN = range(3)
SBE_GRID_REFERENCE, SBE_GRID_MP = 0, 1
SBE_INIT_UNIFORM_VALENCE, SBE_INIT_FERMI_DIRAC = 0, 1
SBE_PREC_FP64, SBE_PREC_FP32, SBE_PREC_SPLIT16 = 0, 1, 2
SBE_RK4, SBE_RK5_FEHLBERG, SBE_RK_CASH_KARP = 4, 5, 45
SBE_EFIELD_REFERENCE, SBE_EFIELD_CONSISTENT = 0, 1
SBE_MODEL_HALDANE, SBE_MODEL_WANNIER = 1, 2
SBE_MODEL_WILSON_MASS, SBE_MODEL_SQUARE_SBAND, SBE_MODEL_KANE_MELE, SBE_MODEL_XYH_LATTICE, SBE_MODEL_TMDC = 3, 4, 5, 6, 7
SBE_MODEL_HALDANE_2L, SBE_MODEL_BI2SE3_SURF, SBE_MODEL_MULTI_WSM, SBE_MODEL_WDSM_TAAS = 8, 9, 10, 11
SBE_MODEL_NPARAMS = 24
SBE_BASIS_ORBITAL, SBE_BASIS_BAND = 0, 1
SBE_NOALLOC_SCRATCH_BYTES = 16 << 20
SBE_NOALLOC_K_ALIGN = 512

c_double_p = C.POINTER(C.c_double)
c_int32_p = C.POINTER(C.c_int32)


class sbe_pulse(C.Structure):
    _fields_ = [
        ("E0", C.c_double), ("w0", C.c_double), ("ellip", C.c_double), ("ncycles", C.c_double),
        ("cep", C.c_double), ("t0", C.c_double), ("phix", C.c_double),
        ("ncycles_clipped", C.c_double), ("thetaz", C.c_double), ("phiz", C.c_double),
        ("envelope", C.c_int32), ("reserved", C.c_int32),
    ]


class sbe_config(C.Structure):
    _fields_ = [
        ("abi_version", C.c_int32), ("gauge", C.c_int32), ("npoints", C.c_int32 * 3),
        ("rk_order", C.c_int32), ("dt", C.c_double), ("T1", C.c_double), ("T2", C.c_double),
        ("ksfactor", C.c_double * 3), ("inter_intra", C.c_int32), ("grid", C.c_int32),
        ("has_bz_axes", C.c_int32), ("has_bz_origin", C.c_int32), ("bz_axes", C.c_double * 9),
        ("bz_origin", C.c_double * 3), ("init_type", C.c_int32), ("precision", C.c_int32),
        ("fermi_energy", C.c_double), ("npulses", C.c_int32), ("efield_mode", C.c_int32),
        ("pulses", C.POINTER(sbe_pulse)), ("offset_before", C.c_double),
        ("offset_after", C.c_double), ("rtol", C.c_double), ("atol", C.c_double),
    ]


class sbe_model(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("nband", C.c_int32), ("nval", C.c_int32),
        ("is_dipole_zero", C.c_int32), ("a0", C.c_double), ("t1", C.c_double),
        ("t2", C.c_double), ("phi0", C.c_double), ("M0", C.c_double), ("nrpts", C.c_int32),
        ("reserved", C.c_int32), ("lattice", C.c_double * 9), ("rvec", c_int32_p),
        ("rweight", c_double_p), ("ham", c_double_p), ("pos", c_double_p),
        ("params", C.c_double * 24),
    ]


class sbe_info(C.Structure):
    _fields_ = [
        ("nt", C.c_int64), ("nk_total", C.c_int64), ("k_begin", C.c_int64), ("k_end", C.c_int64),
        ("nband", C.c_int32), ("nstage", C.c_int32), ("dt", C.c_double), ("atmin", C.c_double),
        ("atmax", C.c_double), ("dV", C.c_double), ("total_weight", C.c_double),
        ("steps_done", C.c_int64), ("device_bytes", C.c_int64),
    ]


class sbe_request_status(C.Structure):
    _fields_ = [("state", C.c_int32), ("status", C.c_int32), ("nt", C.c_int64), ("nband", C.c_int32),
                ("reserved", C.c_int32), ("footprint", C.c_int64), ("t_submit", C.c_double),
                ("t_admit", C.c_double), ("t_done", C.c_double)]


SBE_REQ_QUEUED, SBE_REQ_RUNNING, SBE_REQ_DONE, SBE_REQ_FAILED, SBE_REQ_CANCELLED = range(5)


class sbe_step_stats(C.Structure):
    _fields_ = [("blocks", C.c_int64), ("k_per_block", C.c_int64), ("accepted", C.c_int64),
                ("rejected", C.c_int64), ("max_accepted", C.c_int64), ("min_accepted", C.c_int64)]


class sbe_tuning(C.Structure):
    _fields_ = [("kpt", C.c_int32), ("chunk_steps", C.c_int32), ("nb_kernels", C.c_int32), ("nb_sums", C.c_int32),
                ("tuned", C.c_int32), ("candidates", C.c_int32), ("ms_per_step", C.c_double)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


# sbe_gen_matrix_fn: (user, k[3], out) -> None
sbe_gen_matrix_fn = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double))

# Exported C-ABI symbols with their ctypes signatures (argtypes, restype).
SIGNATURES = {
    "sbe_create": ([C.POINTER(sbe_config), C.POINTER(sbe_model), C.c_int64, C.c_int64,
                    C.POINTER(C.c_void_p)], C.c_int),
    "sbe_run": ([C.c_void_p, C.c_int64, C.c_int64], C.c_int),
    "sbe_synchronize": ([C.c_void_p], C.c_int),
    "sbe_is_idle": ([C.c_void_p], C.c_int),
    "sbe_get_series": ([C.c_void_p, c_double_p, c_double_p, c_double_p, c_double_p], C.c_int),
    "sbe_get_state": ([C.c_void_p, C.c_int32, c_double_p], C.c_int),
    "sbe_get_info": ([C.c_void_p, C.POINTER(sbe_info)], C.c_int),
    "sbe_query_footprint": ([C.POINTER(sbe_config), C.POINTER(sbe_model), C.c_int64, C.c_int64,
                             C.POINTER(C.c_size_t)], C.c_int),
    "sbe_laser_eval": ([C.POINTER(sbe_config), C.c_int64, c_double_p, c_double_p, c_double_p,
                        c_double_p, C.POINTER(C.c_int64)], C.c_int),
    "sbe_set_timing": ([C.c_void_p, C.c_int32], C.c_int),
    "sbe_get_step_stats": ([C.c_void_p, C.POINTER(sbe_step_stats)], C.c_int),
    "sbe_batch_get_step_stats": ([C.c_void_p, C.c_int32, C.POINTER(sbe_step_stats)], C.c_int),
    "sbe_get_timing": ([C.c_void_p, c_double_p, C.POINTER(C.c_int64)], C.c_int),
    "sbe_batch_create": ([C.c_int32, C.POINTER(sbe_config), C.POINTER(sbe_model), C.POINTER(C.c_void_p)], C.c_int),
    "sbe_batch_run": ([C.c_void_p, C.c_int64], C.c_int),
    "sbe_batch_synchronize": ([C.c_void_p], C.c_int),
    "sbe_batch_reset": ([C.c_void_p], C.c_int),
    "sbe_batch_get_info": ([C.c_void_p, C.c_int32, C.POINTER(sbe_info)], C.c_int),
    "sbe_batch_get_series": ([C.c_void_p, C.c_int32, c_double_p, c_double_p, c_double_p, c_double_p], C.c_int),
    "sbe_batch_get_timing": ([C.c_void_p, c_double_p, C.POINTER(C.c_int64)], C.c_int),
    "sbe_batch_last_error": ([C.c_void_p], C.c_char_p),
    "sbe_batch_destroy": ([C.c_void_p], None),
    "sbe_server_create": ([C.c_size_t, C.c_int32, C.c_int64, C.POINTER(C.c_void_p)], C.c_int),
    "sbe_server_submit": ([C.c_void_p, C.POINTER(sbe_config), C.POINTER(sbe_model), C.POINTER(C.c_int64)], C.c_int),
    "sbe_server_step": ([C.c_void_p], C.c_int),
    "sbe_server_poll": ([C.c_void_p, C.c_int64, C.POINTER(sbe_request_status)], C.c_int),
    "sbe_server_counts": ([C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_size_t),
                           C.POINTER(C.c_size_t)], C.c_int),
    "sbe_server_get_series": ([C.c_void_p, C.c_int64, c_double_p, c_double_p, c_double_p, c_double_p], C.c_int),
    "sbe_server_request_error": ([C.c_void_p, C.c_int64], C.c_char_p),
    "sbe_server_release": ([C.c_void_p, C.c_int64], C.c_int),
    "sbe_server_cancel": ([C.c_void_p, C.c_int64], C.c_int),
    "sbe_server_last_error": ([C.c_void_p], C.c_char_p),
    "sbe_server_destroy": ([C.c_void_p], None),
    "sbe_destroy": ([C.c_void_p], None),
    "sbe_run_haldane_noalloc": ([C.POINTER(sbe_config), C.POINTER(sbe_model), C.c_void_p, C.c_void_p, C.c_int64,
                                 C.c_int64, C.c_int64, C.c_int64, C.c_int32, C.c_void_p, C.c_void_p], C.c_int),
    "sbe_last_error": ([C.c_void_p], C.c_char_p),
    "sbe_autotune": ([C.POINTER(sbe_config), C.POINTER(sbe_model), C.c_int64, C.POINTER(sbe_tuning)], C.c_int),
    "sbe_tuning_get": ([C.POINTER(sbe_config), C.POINTER(sbe_model), C.POINTER(sbe_tuning)], C.c_int),
    "sbe_tuning_set": ([C.POINTER(sbe_config), C.POINTER(sbe_model), C.POINTER(sbe_tuning)], C.c_int),
    "sbe_tuning_clear": ([], None),
    "sbe_batch_autotune": ([C.c_int32, C.POINTER(sbe_config), C.POINTER(sbe_model), C.c_int64, C.POINTER(sbe_tuning)],
                           C.c_int),
    "sbe_sample_material": ([C.c_int32, C.c_int32, C.c_int32, c_double_p, c_double_p, C.POINTER(C.c_int32),
                             sbe_gen_matrix_fn, sbe_gen_matrix_fn, C.c_void_p, C.POINTER(sbe_model), c_double_p],
                            C.c_int),
    "sbe_free_sampled_material": ([C.POINTER(sbe_model)], None),
    "sbe_tmdc_params": ([C.c_char_p, C.c_int32, C.POINTER(sbe_model)], C.c_int),
    "sbe_hhg_points": ([C.c_int64], C.c_int64),
    "sbe_hhg_spectrum": ([C.c_int64, c_double_p, c_double_p, C.c_double, C.c_double, C.c_int32, c_double_p,
                          c_double_p, c_double_p, c_double_p, c_double_p], C.c_int),
    "sbe_write_hhg": ([C.c_char_p, C.c_int64, c_double_p, c_double_p, C.c_double, C.c_double, C.c_int32], C.c_int),
    "sbe_spectrum_last_error": ([], C.c_char_p),
}


def bind(lib: C.CDLL, signatures=SIGNATURES) -> C.CDLL:
    for name, (argtypes, restype) in signatures.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = restype
    return lib
