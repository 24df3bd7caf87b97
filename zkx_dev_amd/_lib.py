"""ctypes binding of libzkx_radar.so (the C ABI declared in include/zkx_radar.h).

The library is loaded from zkx_dev_amd/lib/.  There is no fallback: if the HIP
library is missing, every operator raises.
"""
import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libzkx_radar.so")

ERRORS = {
    0: "SUCCESS", 1001: "INVALID_ARGUMENT", 1002: "INVALID_SHAPE", 1003: "INVALID_DTYPE",
    1004: "UNSUPPORTED_OPERATION", 2002: "CUDA_RUNTIME_ERROR", 2003: "OUT_OF_MEMORY",
    2099: "BACKEND_ERROR", 9001: "INTERNAL_ERROR",
}

P = ctypes.c_void_p
I = ctypes.c_int
D = ctypes.c_double
SZ = ctypes.c_size_t
LL = ctypes.c_longlong

SIGNATURES = {
    "zkx_last_error_message": (ctypes.c_char_p, []),
    "zkx_version": (I, []),
    "zkx_max_lds_fft_length": (I, []),
    "zkx_set_kernel_variant": (I, [ctypes.c_char_p, I]),
    "zkx_get_kernel_variant": (I, [ctypes.c_char_p]),
    "zkx_pulse_compression_c64": (I, [P, I, I, P, I, I, I, I, P, P, P]),
    "zkx_pc_template_prepare": (I, [P, I, I, I, I, P, P, P]),
    "zkx_pulse_compression_prepared_c64": (I, [P, I, I, P, I, P, P]),
    "zkx_pulse_doppler_c64": (I, [P, I, I, I, I, P, P, P, P]),
    "zkx_cfar_alpha": (I, [D, I, ctypes.POINTER(D)]),
    "zkx_ca_cfar_f32": (I, [P, I, P, P, P, D, P, P, P, P]),
    "zkx_ca_cfar_noalloc_max_cells": (LL, []),
    "zkx_ca_cfar_f32_noalloc": (I, [P, I, P, P, P, D, P, P, P]),
    "zkx_ca_cfar_window_f32": (I, [P, I, I, I, I, I, I, D, P, P, P]),
    "zkx_widen_c64_to_c128": (I, [P, P, SZ, P]),
    "zkx_widen_f32_to_f64": (I, [P, P, SZ, P]),
    "zkx_convert_to_f32": (I, [P, I, SZ, P, P]),
    "zkx_cfar_detect_typed": (I, [P, I, P, SZ, P, P]),
    "zkx_resample_poly_scale_taps": (I, [P, I, I, I, P, P]),
    "zkx_cfar_alpha_f32_device": (I, [ctypes.c_float, I, P, P]),
    "zkx_ambgfun_out_len": (LL, [I, I, I]),
    "zkx_ambgfun_c64": (I, [P, SZ, I, P, SZ, I, I, D, D, I, P, P]),
    "zkx_echo_source_c64": (I, [P, P, P, P, P, P]),
    "zkx_dwell_workspace_bytes": (SZ, [I, I, I]),
    "zkx_doppler_cfar_c64": (I, [P, I, I, I, I, I, I, I, I, D, P, P, P, P, P]),
    "zkx_dwell_batch_c64": (I, [P, I, I, I, P, I, I, I, I, I, D, I, P, P, P, P, P]),
    "zkx_upfirdn_out_len": (LL, [LL, LL, I, I]),
    "zkx_upfirdn_f32": (I, [P, I, P, I, I, I, I, I, P, P]),
    "zkx_upfirdn_c64": (I, [P, I, P, I, I, I, I, I, P, P]),
    "zkx_resample_poly_f32": (I, [P, I, P, I, I, I, I, I, P, P]),
    "zkx_resample_poly_c64": (I, [P, I, P, I, I, I, I, I, P, P]),
    "zkx_firfilter_f32": (I, [P, I, P, I, I, I, P, P, P, P]),
    "zkx_decimate_f32": (I, [P, I, P, I, I, I, I, I, P, P]),
    "zkx_channelize_poly_f32": (I, [P, LL, P, LL, I, P, P]),
    "zkx_correlate_out_len": (LL, [LL, LL, I]),
    "zkx_correlate_nd_out_shape": (I, [P, P, I, I, P, P]),
    "zkx_correlate_nd": (I, [P, P, P, P, I, I, I, P, P]),
    "zkx_fftconvolve_nd_shape": (I, [P, P, I, ctypes.c_uint, I, P]),
    "zkx_fftconvolve_nd": (I, [P, P, I, P, P, I, I, ctypes.c_uint, I, P, P]),
    "zkx_correlate_direct": (I, [P, LL, P, LL, I, I, P, P]),
    "zkx_correlate_fft": (I, [P, LL, P, LL, I, I, P, P]),
    "zkx_correlate2d_out_shape": (I, [LL, LL, LL, LL, I, P, P]),
    "zkx_correlate2d": (I, [P, LL, LL, P, LL, LL, I, I, I, ctypes.c_double, P, P]),
    "zkx_firwin_lowpass_hamming": (I, [I, ctypes.c_float, P]),
    "zkx_fftconvolve_out_len": (LL, [LL, LL, I]),
    "zkx_fftconvolve_c64": (I, [P, I, I, P, I, I, P, P]),
    "zkx_engine_create": (I, [ctypes.POINTER(P), I, I, I, I, I, I, I, D, I]),
    "zkx_engine_destroy": (I, [P]),
    "zkx_engine_add_template": (I, [P, P, I, ctypes.POINTER(I)]),
    "zkx_engine_set_options": (I, [P, I, I, I]),
    "zkx_engine_run": (I, [P, I, P, P, I, P, P, P, P]),
    "zkx_engine_plan": (I, [P, P, I, P, P]),
    "zkx_engine_autotune": (I, [P, I, P, P, I, P, P, P]),
    "zkx_engine_tuning_export": (I, [P, ctypes.c_char_p, I, ctypes.POINTER(I)]),
    "zkx_engine_tuning_import": (I, [P, ctypes.c_char_p]),
    "zkx_doppler_dft_mfma_c64": (I, [P, I, I, I, I, I, P, P, P]),
    "zkx_server_create": (I, [ctypes.POINTER(P), P, P, LL, LL, I, I, I]),
    "zkx_server_destroy": (I, [P]),
    "zkx_server_submit": (I, [P, LL, I, I, I, ctypes.POINTER(LL)]),
    "zkx_server_wait": (I, [P, LL, I, ctypes.POINTER(I)]),
    "zkx_server_result": (I, [P, LL, ctypes.POINTER(D), ctypes.POINTER(P), ctypes.POINTER(P)]),
    "zkx_server_release": (I, [P, LL]),
    "zkx_server_copy_result": (I, [P, LL, P, P, P]),
    "zkx_server_stats": (I, [P, P]),
    "zkx_server_hold": (I, [P, I]),
    "zkx_server_replay": (I, [P, P, P, P, P, P, I, P]),
}

_lib = None
_lock = threading.Lock()


INVALID_ARGUMENT, INVALID_SHAPE, INVALID_DTYPE = 1001, 1002, 1003
BACKEND_ERROR = 2099


class ZkxError(RuntimeError):
    """A non-zero return from the C ABI; .code is the reference ErrorCode."""

    def __init__(self, code, message):
        self.code = code
        self.name = ERRORS.get(code, str(code))
        super().__init__(f"{self.name} ({code}): {message}")


class ArgumentError(ZkxError, ValueError):
    """An invalid argument, shape or dtype (codes 1001-1003), rejected before
    any device work.  A ValueError, as the reference throws
    std::invalid_argument for these, that also carries the ErrorCode the C
    ABI reports for the same problem."""


def lib():
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.exists(LIB_PATH):
                    raise ImportError(
                        f"{LIB_PATH} is missing: build it with `python -m zkx_dev_amd.build`")
                L = ctypes.CDLL(LIB_PATH)
                for name, (res, args) in SIGNATURES.items():
                    fn = getattr(L, name)
                    fn.restype = res
                    fn.argtypes = args
                _lib = L
    return _lib


def check(code):
    if code != 0:
        msg = lib().zkx_last_error_message().decode(errors="replace")
        if code in (INVALID_ARGUMENT, INVALID_SHAPE, INVALID_DTYPE):
            raise ArgumentError(code, msg)
        raise ZkxError(code, msg)
