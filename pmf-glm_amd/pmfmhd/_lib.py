"""ctypes binding of libpmfmhd.so (the C ABI declared in include/pmfmhd.h)."""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpmfmhd.so")


class pmf_params(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("riemann", ctypes.c_int32), ("recon", ctypes.c_int32), ("bc", ctypes.c_int32 * 3),
                ("solver", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("dx", ctypes.c_double), ("gamma", ctypes.c_double), ("smallr", ctypes.c_double),
                ("pfl", ctypes.c_double), ("llf_dmin", ctypes.c_double), ("llf_pmin", ctypes.c_double)]


_lib = None


def lib():
    """Load the HIP library; raise if it has not been built (there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C pmf-glm_amd` "
                               "(or __graft_entry__.build()); pmfmhd has no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        P, i32, i64, d = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
        pd, pi64 = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64)
        pp = ctypes.POINTER(ctypes.c_void_p)
        sig = {
            "pmf_params_default": [ctypes.POINTER(pmf_params), i32, i32, i32, d],
            "pmf_create": [ctypes.POINTER(pmf_params), pp],
            "pmf_destroy": [P],
            "pmf_create_workspace": [ctypes.POINTER(pmf_params), pp],
            "pmf_step_into": [P, P, P, P, d, d, d, i32],
            "pmf_step_into_static": [ctypes.POINTER(pmf_params), P, P, P, d, d, d, i32, P, P],
            "pmf_import_mhd_raw": [P, ctypes.c_char_p],
            "pmf_advance": [P, i64, d, d, d, pd, pd],
            "pmf_export_mhd_raw": [P, ctypes.c_char_p],
            "pmf_max_wavespeed_of": [P, P, pd],
            "pmf_upload": [P, i32, P, i64],
            "pmf_download": [P, i32, P, i64],
            "pmf_device_ptr": [P, i32, pp],
            "pmf_scratch_ptr": [P, i32, pp],
            "pmf_ct_ptr": [P, i32, pp],
            "pmf_ct_download": [P, i32, P, i64],
            "pmf_set_stream": [P, P],
            "pmf_synchronize": [P],
            "pmf_max_wavespeed": [P, pd],
            "pmf_compute_dt": [P, d, pd, pd],
            "pmf_step_raw": [P, d, d, d, i32],
            "pmf_step": [P, d, d, d, i32],
            "pmf_evolve": [P, d, d, d, d, i32, pd, pi64],
            "pmf_max_divb": [P, pd],
            "pmf_divb_norm_fp64": [P, pd, pd],
            "pmf_conserved_totals_fp64": [P, pd],
            "pmf_max_divb_face": [P, pd],
            "pmf_ct_install_faces": [P, P, P, P],
            "pmf_ct_faces_from_potential": [P, P, P, P],
            "pmf_riemann_flux_batch": [P, P, P, i32, d, i64, P],
            "pmf_surrogate_create": [P, ctypes.c_uint64, pp],
            "pmf_surrogate_destroy": [P],
            "pmf_surrogate_get_weights": [P, P, i64],
            "pmf_surrogate_set_weights": [P, P, i64],
            "pmf_surrogate_reserve": [P, i64],
            "pmf_surrogate_dump": [P, P, d, d, d, d, ctypes.c_uint64, pi64],
            "pmf_surrogate_dataset": [P, i64, i64, P, P, P],
            "pmf_surrogate_upload": [P, P, P, P, i64],
            "pmf_surrogate_grad": [P, i64, i64, i32, P, pd],
            "pmf_surrogate_predict": [P, i64, i64, i32, P],
            "pmf_surrogate_train": [P, i32, i32, d, ctypes.c_uint64, i32, P],
            "pmf_surrogate_face_flux": [P, i64, i64, i32, d, P],
            "pmf_step_surrogate": [P, P, d, d, d, d, pi64],
            "pmf_set_chunk": [P, i32],
            "pmf_set_ct_kernel": [P, i32],
            "pmf_set_ct_workspace": [P, i32],
            "pmf_step_drag": [P, d, d, d, d, pd, i32],
            "pmf_buffer_alloc": [i64, pp],
            "pmf_species_to_fractions": [P, P, P, P, d],
            "pmf_advect_species": [P, P, P, P, P, d, d],
            "pmf_advect_passive_scalar": [P, P, P, d],
            "pmf_buffer_free": [P],
            "pmf_buffer_copy": [P, P, i64, i32],
            "pmf_step_radiation": [P, d, d, d, d, P, i32, i32],
            "pmf_step_sources_into": [P, P, P, P, P, d, d, d, d, pd, i32, i32],
            "pmf_autotune": [P, i32, ctypes.POINTER(ctypes.c_int32), pd],
            "pmf_ensemble_create": [ctypes.POINTER(pmf_params), i32, P, ctypes.POINTER(ctypes.c_void_p)],
            "pmf_ensemble_destroy": [P],
            "pmf_ensemble_upload": [P, i32, i32, P, i64],
            "pmf_ensemble_download": [P, i32, i32, P, i64],
            "pmf_ensemble_run": [P, i32, d, d, d, i32, pd],
            "pmf_ensemble_clock": [P, P, P, P],
            "pmf_ensemble_tune": [P, i32, pd],
            "pmf_ensemble_config": [P, P, P, P],
            "pmf_ensemble_set_config": [P, P, i32],
            "pmf_timing_enable": [P, i32],
            "pmf_timing_read": [P, pd, pi64],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        L.pmf_surrogate_nparams.restype = ctypes.c_int
        L.pmf_surrogate_nparams.argtypes = []
        L.pmf_last_error.restype = ctypes.c_char_p
        L.pmf_last_error.argtypes = []
        L.pmf_version.restype = ctypes.c_char_p
        L.pmf_version.argtypes = []
        _lib = L
    return _lib


class MHDError(RuntimeError):
    """Raised with the library's message (the reference's text for the same misuse)."""


def check(rc: int):
    if rc != 0:
        raise MHDError(lib().pmf_last_error().decode())
