"""ctypes binding of libsvm.so (include/svm.h).  The library is required: there is
no fallback path, and every entry point raises if it is missing."""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsvm.so")

# svm_status (include/svm.h)
OK, ERR_INVALID, ERR_NOT_FOUND, ERR_UNSUPPORTED, ERR_DEVICE, ERR_NO_MEMORY, ERR_JIT, ERR_BUSY, ERR_STATE = (
    0, -1, -2, -3, -4, -5, -6, -7, -8)
STATUS_NAMES = {0: "OK", -1: "INVALID", -2: "NOT_FOUND", -3: "UNSUPPORTED", -4: "DEVICE",
                -5: "NO_MEMORY", -6: "JIT", -7: "BUSY", -8: "STATE"}

# Every function declared in include/svm.h, with (restype, argtypes).
c_u8p = ctypes.POINTER(ctypes.c_uint8)
c_i32p = ctypes.POINTER(ctypes.c_int32)
c_i64p = ctypes.POINTER(ctypes.c_int64)
c_u64p = ctypes.POINTER(ctypes.c_uint64)
c_dblp = ctypes.POINTER(ctypes.c_double)
c_vp = ctypes.c_void_p


class SvmConfig(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("tuple_per_bulk", ctypes.c_uint64),
                ("hbm_budget_bytes", ctypes.c_uint64), ("num_streams", ctypes.c_int),
                ("verbose", ctypes.c_int)]


class SvmResultDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("ncols", ctypes.c_int32), ("rows", ctypes.c_int64),
                ("type", ctypes.c_int32 * 24), ("val_size", ctypes.c_int64 * 24),
                ("col", ctypes.c_void_p * 24), ("build_ms", ctypes.c_double),
                ("scan_ms", ctypes.c_double), ("finish_ms", ctypes.c_double)]


class SvmGenParams(ctypes.Structure):
    _fields_ = [("n_cust", ctypes.c_uint64), ("n_supp", ctypes.c_uint64),
                ("n_part", ctypes.c_uint64), ("zipf_s", ctypes.c_double)]


SIGNATURES = {
    "svm_init": (ctypes.c_int, [ctypes.POINTER(SvmConfig)]),
    "svm_shutdown": (ctypes.c_int, []),
    "svm_last_error": (ctypes.c_char_p, []),
    "svm_version": (ctypes.c_char_p, []),
    "svm_set_hybrid": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "svm_table_create": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, c_i32p, c_i64p, ctypes.c_uint64]),
    "svm_table_put_column": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, c_vp, ctypes.c_uint64]),
    "svm_table_get_column": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, c_vp, ctypes.c_uint64]),
    "svm_table_column_device": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(c_vp)]),
    "svm_table_info": (ctypes.c_int, [ctypes.c_char_p, c_u64p, ctypes.POINTER(ctypes.c_int)]),
    "svm_table_column_type": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, c_i32p, c_i64p]),
    "svm_table_view": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_uint64, ctypes.c_uint64]),
    "svm_table_drop": (ctypes.c_int, [ctypes.c_char_p]),
    "svm_table_exists": (ctypes.c_int, [ctypes.c_char_p]),
    "svm_gen_table": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64,
                                     ctypes.POINTER(SvmGenParams)]),
    "svm_gen_host": (ctypes.c_int, [ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(SvmGenParams),
                                    ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(c_vp), ctypes.c_int]),
    "svm_gen_schema": (ctypes.c_int, [ctypes.c_int, c_i32p, c_i64p]),
    "svm_pipeline_run": (ctypes.c_int, [c_u8p, ctypes.c_size_t, ctypes.POINTER(c_vp)]),
    "svm_pipeline_run_noalloc": (ctypes.c_int, [c_u8p, ctypes.c_size_t, ctypes.POINTER(c_vp), ctypes.c_int,
                                                ctypes.c_uint64, c_u64p]),
    "svm_result_kind": (ctypes.c_int, [c_vp]),
    "svm_result_rows": (ctypes.c_int64, [c_vp]),
    "svm_result_cols": (ctypes.c_int, [c_vp]),
    "svm_result_col_type": (ctypes.c_int, [c_vp, ctypes.c_int, c_i32p, c_i64p]),
    "svm_result_col": (c_vp, [c_vp, ctypes.c_int]),
    "svm_result_timing": (ctypes.c_int, [c_vp, c_dblp, c_dblp, c_dblp]),
    "svm_result_describe": (ctypes.c_int, [c_vp, ctypes.POINTER(SvmResultDesc)]),
    "svm_result_free": (None, [c_vp]),
    "svm_results_describe": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(c_vp), ctypes.POINTER(SvmResultDesc)]),
    "svm_results_copy": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(c_vp), c_vp, ctypes.c_uint64]),
    "svm_results_free": (None, [ctypes.c_int, ctypes.POINTER(c_vp)]),
    "svm_batch_run": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(c_u8p), ctypes.POINTER(ctypes.c_size_t),
                                     ctypes.POINTER(c_vp)]),
    "svm_server_start": (ctypes.c_int, [ctypes.c_int, ctypes.c_uint64, ctypes.c_int, ctypes.c_int]),
    "svm_server_submit": (ctypes.c_int, [ctypes.POINTER(c_u8p), ctypes.POINTER(ctypes.c_size_t), ctypes.c_int,
                                         c_u64p]),
    "svm_server_wait": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(c_vp)]),
    "svm_server_wait_timed": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(c_vp),
                                             ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "svm_server_stats": (ctypes.c_int, [c_u64p, c_u64p, c_u64p, c_u64p]),
    "svm_server_stop": (ctypes.c_int, []),
    "svm_debug_hash_drops": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_uint64, c_u64p]),
}

_lib = None


class SvmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{STATUS_NAMES.get(code, code)}] {msg}")
        self.code = code


def lib():
    """Load libsvm.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m shadowvm_amd.build` "
                               "(there is no CPU fallback for the data plane)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc < 0:
        raise SvmError(rc, lib().svm_last_error().decode(errors="replace"))
    return rc
