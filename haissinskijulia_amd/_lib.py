"""ctypes binding of libhaissinski_mi355x.so (include/haissinski_mi355x.h).

The library is loaded from this package directory only.  There is no
fallback: if the library is missing or no gfx950 device is present, the
calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhaissinski_mi355x.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "haissinski_mi355x.h")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int64)

HM_OK = 0
HM_ERR_ARG = -1
HM_ERR_EMPTY_WAKE = -2
HM_ERR_NBINS = -3
HM_ERR_PIPE = -4
HM_ERR_DEVICE = -10
HM_ERR_OOM = -11
HM_ERR_LIMIT = -12
HM_ERR_BUSY = -13
HM_ERR_CANCELLED = -14


class HmParams(C.Structure):
    _fields_ = [
        ("E0", C.c_double), ("mass", C.c_double), ("voltage", C.c_double),
        ("harmonic", C.c_int64), ("radius", C.c_double), ("pipe_radius", C.c_double),
        ("alpha_c", C.c_double), ("phi_s", C.c_double), ("freq_rf", C.c_double),
        ("n_turns", C.c_int64), ("use_wakefield", C.c_int32), ("update_eta", C.c_int32),
        ("update_E0", C.c_int32), ("SR_damping", C.c_int32), ("use_excitation", C.c_int32),
    ]


class HmOptions(C.Structure):
    _fields_ = [
        ("normals", _dp), ("seed", C.c_uint64), ("turn_offset", C.c_int64),
        ("wake_table", _dp), ("delta_sigma", C.c_double), ("bin_half_width_sigmas", C.c_double),
        ("charge", C.c_double), ("n_resonators", C.c_int32), ("timing", C.c_int32),
        ("resonators", _dp), ("fft_block", C.c_int32), ("block_size", C.c_int32),
        ("monitor", _dp), ("serial", C.c_int32), ("fom_stats", C.c_int32),
        ("tile_pairs", C.c_int32), ("run_pairs", C.c_int32), ("bin_whole", C.c_int32),
    ]


class HmResult(C.Structure):
    _fields_ = [
        ("sigma_E", C.c_double), ("sigma_z", C.c_double), ("E0", C.c_double),
        ("turns_done", C.c_int64), ("drift_skipped", C.c_int32), ("pad_", C.c_int32),
        ("ms_total", C.c_double), ("ms_pass_avg", C.c_double), ("ms_conv_avg", C.c_double),
        ("pass_launches", C.c_int64), ("sigma_E0", C.c_double), ("sigma_z0", C.c_double),
        ("bin_half_width", C.c_double), ("mean_z", C.c_double), ("mean_dE", C.c_double),
        ("corr_z_dE", C.c_double),
    ]


class HmJob(C.Structure):
    _fields_ = [
        ("params", HmParams), ("opts", HmOptions), ("n", C.c_int64), ("nbins", C.c_int64),
        ("particle_seed", C.c_uint64), ("sigma_z", C.c_double), ("sigma_E", C.c_double),
        ("fail_after_turns", C.c_int32), ("fail_setup", C.c_int32),
    ]


class HmJobResult(C.Structure):
    _fields_ = [
        ("job_id", C.c_int64), ("status", C.c_int32), ("pad_", C.c_int32), ("result", HmResult),
        ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double),
        ("centroid_z", C.c_double), ("centroid_E", C.c_double),
    ]


class HaissinskiError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


_lib = None
_lock = threading.Lock()


def load() -> C.CDLL:
    """Load the built library (raises if it has not been built)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise HaissinskiError(HM_ERR_DEVICE,
                                  f"{LIB_PATH} is missing: build it with python -m haissinskijulia_amd.build")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        st = C.c_int32
        sig = {
            "hm_create": (st, [C.c_int32, C.POINTER(vp)]),
            "hm_destroy": (st, [vp]),
            "hm_last_error": (C.c_char_p, []),
            "hm_device_info": (st, [vp, C.c_char_p, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
            "hm_evolve": (st, [vp, _dp, _dp, C.c_int64, C.c_int64, C.POINTER(HmParams),
                               C.POINTER(HmOptions), C.POINTER(HmResult)]),
            "hm_bunch_create": (st, [vp, C.c_int64, C.POINTER(vp)]),
            "hm_bunch_destroy": (st, [vp]),
            "hm_bunch_upload": (st, [vp, _dp, _dp]),
            "hm_bunch_download": (st, [vp, _dp, _dp]),
            "hm_bunch_generate": (st, [vp, C.c_uint64, C.c_double, C.c_double, C.c_double, C.c_double]),
            "hm_bunch_generate_particles": (st, [vp, C.c_uint64, C.c_double, C.c_double, C.c_double,
                                                 C.c_double, C.POINTER(C.c_double),
                                                 C.POINTER(C.c_double)]),
            "hm_bunch_evolve": (st, [vp, C.c_int64, C.POINTER(HmParams), C.POINTER(HmOptions),
                                     C.POINTER(HmResult)]),
            "hm_bunch_device_ptrs": (st, [vp, C.POINTER(vp), C.POINTER(vp)]),
            "hm_evolve_device_noalloc": (st, [vp, vp, vp, C.c_int64, C.c_int64, C.POINTER(HmParams),
                                              C.POINTER(HmOptions), C.POINTER(HmResult)]),
            "hm_batch_evolve": (st, [vp, C.c_int32, C.POINTER(vp), _ip, C.POINTER(HmParams),
                                     C.POINTER(HmOptions), C.POINTER(HmResult)]),
            "hm_rf_kick": (st, [vp, C.c_double, C.c_double, C.c_double, C.c_double, _dp, _dp, C.c_int64]),
            "hm_quantum_excitation": (st, [vp, C.c_double, C.c_double, C.c_double, _dp, _dp, C.c_int64]),
            "hm_synchrotron_radiation": (st, [vp, C.c_double, C.c_double, _dp, C.c_int64]),
            "hm_drift": (st, [vp, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_int64, C.c_double,
                              C.c_double, _dp, _dp, C.c_int64, C.POINTER(C.c_int32)]),
            "hm_histogram": (st, [vp, _dp, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_int64, _ip]),
            "hm_wake_potential": (st, [vp, _ip, C.c_int64, _dp, _dp, C.c_double, C.c_int32, _dp]),
            "hm_interpolate_kick": (st, [vp, _dp, _dp, C.c_int64, _dp, C.c_int64, _dp]),
            "hm_global_std": (st, [vp, _dp, C.c_int64, _dp]),
            "hm_pair_stats": (st, [vp, _dp, _dp, C.c_int64, _dp]),
            "hm_normals": (st, [vp, C.c_uint64, C.c_int64, C.c_int64, C.c_int64, _dp]),
            "hm_bin_grid": (st, [C.c_double, C.c_int64, _dp, _dp, _dp]),
            "hm_stream_copy": (st, [vp, C.c_int64, C.c_int32, _dp, _dp]),
            "hm_server_create": (st, [vp, C.c_int64, C.c_int32, C.POINTER(vp)]),
            "hm_server_destroy": (st, [vp]),
            "hm_server_submit": (st, [vp, C.POINTER(HmJob), C.POINTER(C.c_int64)]),
            "hm_server_cancel": (st, [vp, C.c_int64]),
            "hm_server_tick": (st, [vp, C.c_int32, C.POINTER(C.c_int32)]),
            "hm_server_poll": (st, [vp, C.POINTER(HmJobResult), C.c_int32, C.POINTER(C.c_int32)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
        return L


def check(status: int) -> None:
    if status != HM_OK:
        msg = load().hm_last_error().decode(errors="replace")
        raise HaissinskiError(status, msg)


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray):
    assert a.dtype == np.int64 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def header_symbols() -> list[str]:
    """Names of the functions declared in include/haissinski_mi355x.h."""
    import re
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:hm_status|const char \*)\s*(hm_\w+)\s*\(", text, re.M)))
