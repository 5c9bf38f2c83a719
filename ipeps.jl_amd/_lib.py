"""ctypes binding of libipeps_hip.so (the C ABI declared in include/ipeps_hip.h)."""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libipeps_hip.so")

F64, C128, MIXED = 0, 1, 2

_P = C.c_void_p
_I64P = C.POINTER(C.c_int64)
_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int)

# name -> (restype, argtypes); one table drives the binding
SIGNATURES = {
    "ipeps_version": (C.c_int, []),
    "ipeps_last_error": (C.c_char_p, []),
    "ipeps_ctx_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "ipeps_ctx_destroy": (C.c_int, [_P]),
    "ipeps_ctx_sync": (C.c_int, [_P]),
    "ipeps_ctx_mem": (C.c_int, [_P, _I64P, _I64P]),
    "ipeps_ctx_set_timing": (C.c_int, [_P, C.c_int]),
    "ipeps_ctx_set_timing_filter": (C.c_int, [_P, C.c_char_p]),
    "ipeps_ctx_kernel_stats": (C.c_int, [_P, C.c_char_p, C.c_int64, C.c_int]),
    "ipeps_tune_config": (C.c_int, [C.c_int, C.c_double]),
    "ipeps_tune_stats": (C.c_int, [_I64P]),
    "ipeps_tune_clear": (C.c_int, []),
    "ipeps_tune_save": (C.c_int, [C.c_char_p]),
    "ipeps_tune_load": (C.c_int, [C.c_char_p]),
    "ipeps_env_create": (C.c_int, [_P, C.c_int, _I64P, _P, C.POINTER(_P)]),
    "ipeps_env_destroy": (C.c_int, [_P]),
    "ipeps_env_set_option": (C.c_int, [_P, C.c_char_p, C.c_double]),
    "ipeps_env_get_draws": (C.c_int, [_P, _I64P]),
    "ipeps_env_projector_stats": (C.c_int, [_P, _DP, C.c_int]),
    "ipeps_env_copy_state": (C.c_int, [_P, _P]),
    "ipeps_env_tensor_dims": (C.c_int, [_P, C.c_int, _I64P, _IP]),
    "ipeps_env_get_tensor": (C.c_int, [_P, C.c_int, _P]),
    "ipeps_env_set_tensor": (C.c_int, [_P, C.c_int, C.c_int, _I64P, _P]),
    "ipeps_env_set_A": (C.c_int, [_P, _P]),
    "ipeps_rg_step": (C.c_int, [_P, C.c_int64, _DP, C.c_int64, _I64P]),
    "ipeps_run_ctm": (C.c_int, [_P, C.c_int64, C.c_int, C.c_int, C.c_double, C.c_int, _P, _P, C.c_int64,
                                _IP, _DP, _DP, C.c_int64, _I64P]),
    "ipeps_gs_energy": (C.c_int, [_P, _P, _P, C.c_int64, _DP, _DP]),
    "ipeps_gs_energy_4x4": (C.c_int, [_P, _I64P, _P, _P, _P, _P, _DP, _DP]),
    "ipeps_gs_norm": (C.c_int, [_P, _DP]),
    "ipeps_step_adjoint": (C.c_int, [_P, C.c_int64, _P, _P, C.c_int64, _DP, _P]),
    "ipeps_gs_energy_grad": (C.c_int, [_P, C.c_int, _I64P, _P, _P, _P, C.c_int64, C.c_int64, C.c_int, C.c_int,
                                       C.c_int, C.c_double, C.c_char_p, _DP, _P, _IP]),
    "ipeps_run_gs_vjp": (C.c_int, [_P, _P, _P, _P, C.c_int64, C.c_int64, C.c_int, C.c_int, C.c_double, C.c_double,
                                   _DP, _P, _IP]),
    "ipeps_es_grad": (C.c_int, [_P, _P, _P, C.c_int64, C.c_double, C.c_double, _P, C.c_int64, C.c_int, C.c_int,
                                C.c_int, C.c_double, _P, _P, _DP, _DP]),
    "ipeps_es_path_create": (C.c_int, [_P, C.c_int64, C.c_int, _P]),
    "ipeps_es_path_destroy": (C.c_int, [_P]),
    "ipeps_es_path_moves": (C.c_int, [_P, _I64P]),
    "ipeps_es_grad_path": (C.c_int, [_P, _P, _P, _P, C.c_int64, C.c_double, C.c_double, _P, C.c_int64, C.c_int,
                                     C.c_int, C.c_int, C.c_double, _P, _P, _DP, _DP]),
    "ipeps_tcon": (C.c_int, [_P, C.c_int, C.c_int, _I64P, _P, _IP, C.c_int, _I64P, _P, _IP, _I64P, _IP, _P,
                             C.c_int64]),
    "ipeps_absorb": (C.c_int, [_P, C.c_int, C.c_int, _I64P, _P, _IP, _I64P, _P, _IP, _P, _IP, _P, _P, _P, _P,
                               _P]),
    "ipeps_absorb_noalloc": (C.c_int, [_P, _I64P, _P, _IP, _I64P, _P, _IP, _P, _IP, _P, _P, _P, _P, _P]),
    "ipeps_gemm": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int64, C.c_int64,
                             C.c_int64, C.c_double, _P, C.c_int64, _P, C.c_int64, C.c_double, _P, C.c_int64]),
    "ipeps_svd": (C.c_int, [_P, C.c_int, C.c_int64, C.c_int64, _P, _P, _DP, _P]),
    "ipeps_orth": (C.c_int, [_P, C.c_int, C.c_int, C.c_int64, C.c_int64, _P, _P]),
    "ipeps_svd_thin": (C.c_int, [_P, C.c_int, C.c_int, C.c_double, C.c_int64, C.c_int64, _P, _P, _DP, _P, _IP]),
    "ipeps_svd_back": (C.c_int, [_P, C.c_int, C.c_int64, C.c_int64, _P, _DP, _P, _P, _DP, _P, C.c_double, _P]),
    "ipeps_correlation_spin": (C.c_int, [_P, _DP, _DP, C.c_int64, C.c_int, C.c_int, _DP, _DP, _DP]),
    "ipeps_tm_dims": (C.c_int, [_P, C.c_int, _I64P, _I64P]),
    "ipeps_tm_apply": (C.c_int, [_P, C.c_int, C.c_int64, _P, _P]),
    "ipeps_tm_dense": (C.c_int, [_P, C.c_int, _P]),
    "ipeps_batch_create": (C.c_int, [C.c_int, C.c_int, C.c_int, _I64P, C.POINTER(_P), _I64P, C.POINTER(_DP),
                                     C.POINTER(_DP), _I64P, C.c_int, C.c_char_p, C.POINTER(_P)]),
    "ipeps_batch_destroy": (C.c_int, [_P]),
    "ipeps_batch_run": (C.c_int, [_P, C.c_int, _DP, _DP]),
    "ipeps_batch_get_tensor": (C.c_int, [_P, C.c_int, C.c_int, _I64P, _IP, _P]),
    "ipeps_batch_stats": (C.c_int, [_P, _I64P]),
    "ipeps_env_run_graph": (C.c_int, [_P, C.c_int64, C.c_int, _I64P]),
}

_lib = None


class IpepsError(RuntimeError):
    pass


def lib():
    """Load the HIP library; raises if it has not been built (no CPU fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise IpepsError(f"{LIB_PATH} is missing: build it with `python ipeps.jl_amd/build.py`")
        l = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def check(status):
    if status != 0:
        msg = lib().ipeps_last_error().decode(errors="replace")
        raise IpepsError(f"ipeps_hip error {status}: {msg}")
