"""Loader of the native library (libsharpclaw_amd.so, the C ABI of
include/sharpclaw_amd.h).  There is no fallback: if the library is absent
or does not load, every entry point raises."""
import ctypes as C
import os

import numpy as np

from ._abi import ScBatchResult, ScParams, ScRequestStats, ScRunCtl, ScRunStats, ScServerStats

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsharpclaw_amd.so")

_lib = None


class ScError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"sharpclaw_amd error {code}: {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run python -m sharpclaw_amd.build")
    L = C.CDLL(LIB_PATH)
    dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")
    P = C.POINTER(ScParams)
    H = C.c_void_p
    dref = C.POINTER(C.c_double)
    sig = {
        "sc_last_error": ([], C.c_char_p),
        "sc_version": ([], C.c_int),
        "sc_create": ([P, C.POINTER(H)], C.c_int),
        "sc_destroy": ([H], C.c_int),
        "sc_set_q": ([H, dp], C.c_int),
        "sc_get_q": ([H, dp], C.c_int),
        "sc_set_aux": ([H, dp], C.c_int),
        "sc_output_begin": ([H], C.c_int),
        "sc_output_wait": ([H, dp], C.c_int),
        "sc_write_restart": ([H, C.c_char_p, C.c_double, C.c_double], C.c_int),
        "sc_read_restart": ([H, C.c_char_p, dref, dref], C.c_int),
        "sc_save_qold": ([H], C.c_int),
        "sc_restore_qold": ([H], C.c_int),
        "sc_step": ([H, C.c_double, C.c_double, C.c_void_p, dref], C.c_int),
        "sc_flux": ([H, C.c_void_p, C.c_void_p, C.c_double, C.c_double, dref], C.c_int),
        "sc_step_noalloc": ([P, C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_void_p, C.c_void_p],
                            C.c_int),
        "sc_ts": ([H, C.c_double, C.c_double, dref], C.c_int),
        "sc_advance": ([H, C.POINTER(ScRunCtl), C.c_double, C.c_double, dref, C.POINTER(ScRunStats)],
                       C.c_int),
        "sc_sums": ([H, dp], C.c_int),
        "sc_time_stage": ([H, C.c_int, C.c_double, dref], C.c_int),
        "sc_copy_bench": ([C.c_int, C.c_size_t, C.c_int, dref], C.c_int),
        "sc_stage_timing": ([H, C.c_int], C.c_int),
        "sc_stage_timing_get": ([H, dref, C.POINTER(C.c_longlong)], C.c_int),
        "sc_set_tile": ([H, C.c_int], C.c_int),
        "sc_get_tile": ([H, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
        "sc_rp": ([P, C.c_int, C.c_int, dp, dp, dp, dp, dp, dp], C.c_int),
        "sc_evec": ([P, C.c_int, C.c_int, dp, dp, dp], C.c_int),
        "sc_weno5": ([P, C.c_int, dp, dp, dp], C.c_int),
        "sc_weno5_char": ([P, C.c_int, dp, dp, dp, dp, dp], C.c_int),
        "sc_flux1": ([P, C.c_int, C.c_int, dp, dp, C.c_double, dref], C.c_int),
        "sc_batch_create": ([P, C.c_int, C.POINTER(C.c_int), dref, dref, C.POINTER(H)], C.c_int),
        "sc_batch_destroy": ([H], C.c_int),
        "sc_batch_set_q": ([H, C.c_int, dp], C.c_int),
        "sc_batch_get_q": ([H, C.c_int, dp], C.c_int),
        "sc_batch_run": ([H, C.POINTER(ScRunCtl), C.c_double, C.c_double, C.c_double,
                          C.POINTER(ScBatchResult), dref], C.c_int),
        "sc_batch_block": ([H, C.POINTER(C.c_int)], C.c_int),
        "sc_batch_tuning": ([H, C.c_int, C.POINTER(C.c_int), dref, C.POINTER(C.c_int)], C.c_int),
        "sc_batch_parts": ([H, C.POINTER(C.c_int)], C.c_int),
        "sc_server_create": ([C.c_int, C.c_size_t, C.c_int, C.POINTER(H)], C.c_int),
        "sc_server_submit": ([H, P, C.POINTER(ScRunCtl), C.c_double, C.c_double, C.c_double, dp,
                              C.POINTER(C.c_longlong)], C.c_int),
        "sc_server_poll": ([H, C.POINTER(C.c_longlong), C.c_int, C.POINTER(C.c_int)], C.c_int),
        "sc_server_wait": ([H, C.c_double, C.POINTER(C.c_int)], C.c_int),
        "sc_server_result": ([H, C.c_longlong, dp, C.POINTER(ScRequestStats)], C.c_int),
        "sc_server_drain": ([H, C.c_double], C.c_int),
        "sc_server_get_stats": ([H, C.POINTER(ScServerStats)], C.c_int),
        "sc_server_destroy": ([H], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise ScError(rc, lib().sc_last_error().decode())
    return rc
