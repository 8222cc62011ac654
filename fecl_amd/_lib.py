"""Loads the in-tree HIP library (fecl_amd/libfecl_ldpc.so) and declares the C ABI.

There is no fallback: if the library is missing the import of any decoder call fails loudly.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfecl_ldpc.so")
_LIB = None

# (name, argtypes) of every function declared in include/fecl_ldpc.h
P, I32, I64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
SIGNATURES = {
    "fecl_last_error": [],
    "fecl_dvbs2_matrix": [I32, I32, I32, P, P, P],
    "fecl_gallager_matrix": [I64, I64, I64, ctypes.c_uint64, P, P, P],
    "fecl_ldpc_structure": [P, P, I64, I64, P, P, P],
    "fecl_encode": [P, P, I64, I64, P, I64, P],
    "fecl_ldpc_scaling_table": [P, I64, I64, P, P, P, P, I64, P, I64, P, P],
    "fecl_ldpc_create": [P, P, I64, I64, P, P],
    "fecl_ldpc_create_checks": [P, P, I64, I64, I64, P, P],
    "fecl_ldpc_destroy": [P],
    "fecl_ldpc_set_options": [P, P],
    "fecl_ldpc_get_options": [P, P, P, I64],
    "fecl_ldpc_sizes": [P, P, P, P, P],
    "fecl_ldpc_checks": [P, P, P, P],
    "fecl_ldpc_save": [P, I32, P, I64, P],
    "fecl_ldpc_load": [P, I64, P, P],
    "fecl_archive_save": [P, P, I64, I64, P, I32, P, I64, P],
    "fecl_archive_load": [P, I64, P, P, P, P, P, I64, P],
    "fecl_ldpc_encode": [P, P, I64, P],
    "fecl_ldpc_encode_device": [P, P, I64, P, P],
    "fecl_ldpc_puncturing": [I64, I64, P, I64, P, I64, P, I64, P],
    "fecl_permute_device": [P, I64, I64, P, P, I64, I32, P],
    "fecl_ldpc_check": [P, P, I64, P],
    "fecl_ldpc_decode": [P, P, I64, P, P],
    "fecl_ldpc_sodecode": [P, P, P, P, I64, P, P, P, P, P],
    "fecl_ldpc_decode_device": [P, P, I64, P, P, P],
    "fecl_ldpc_sodecode_device": [P, P, P, P, I64, P, P, P, P, P, P],
    "fecl_ldpc_decode_device_noalloc": [P, P, I64, P, P, P],
    "fecl_ldpc_profile": [P, I32],
    "fecl_ldpc_profile_read": [P, P, P],
    "fecl_ldpc_set_layout": [P, I32],
    "fecl_ldpc_get_layout": [P, P, P],
    "fecl_ldpc_set_fused_syndrome": [P, I32],
    "fecl_ldpc_set_pass_frames": [P, I64],
    "fecl_ldpc_set_compaction": [P, I32],
    "fecl_ldpc_set_cache_policy": [P, I32],
    "fecl_ldpc_set_vars_per_wave": [P, I32],
    "fecl_ldpc_get_vars_per_wave": [P, P],
    "fecl_ldpc_autotune": [P, I64, I32, P, P],
    "fecl_autotune_cache_config": [I32, ctypes.c_char_p],
    "fecl_autotune_cache_stats": [P, P, P, P, P],
    "fecl_autotune_cache_clear": [],
    "fecl_ldpc_graph_create": [P, P, I32, P, P, P, P],
    "fecl_ldpc_graph_launch": [P, P],
    "fecl_ldpc_graph_check": [P],
    "fecl_ldpc_graph_is_persistent": [P, P],
    "fecl_ldpc_set_graph_persistent": [P, I32],
    "fecl_ldpc_graph_destroy": [P],
    "fecl_ldpc_server_create": [P, I32, P, I64, P],
    "fecl_ldpc_server_slots": [P, I32, P],
    "fecl_ldpc_server_submit": [P, I32, P, I64, P, P, P],
    "fecl_ldpc_server_wait": [P, ctypes.c_uint64, ctypes.c_double, P, P],
    "fecl_ldpc_server_stats": [P, P, P, P],
    "fecl_ldpc_server_set_round_steps": [P, I32],
    "fecl_ldpc_server_destroy": [P],
    "fecl_ldpc_trim": [P],
}

FECL_OK, FECL_INVALID_ARGUMENT, FECL_OUT_OF_MEMORY, FECL_DEVICE_ERROR, FECL_NOT_SUPPORTED = range(5)


class FeclError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"fecl status {status}: {msg}")
        self.status = status


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run "
                               "`make -C fecl_amd/csrc`")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_char_p if name == "fecl_last_error" else ctypes.c_int
        _LIB = L
    return _LIB


def check(status):
    if status == FECL_OK:
        return
    msg = lib().fecl_last_error().decode()
    if status == FECL_INVALID_ARGUMENT:
        raise ValueError(msg)
    raise FeclError(status, msg)
