"""ctypes binding of libjsolver_amd.so (declarations: include/jsolver_amd.h).

The library is loaded on first use.  There is no CPU fallback: if the library
is missing or no GPU is visible, every compute entry point raises.
"""
import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libjsolver_amd.so")

JSV_FP64 = 0
JSV_MIXED = 1
JSV_STOP_DISTANCE = 0
JSV_STOP_REL_RESIDUAL = 1
JSV_TUNE_SLAB_BYTES = 0
JSV_TUNE_TWO_PASS_MIN = 1
JSV_TUNE_FUSED_MIN_M = 2
JSV_TUNE_ZCHUNK = 3
JSV_TUNE_SWEEP_ROWS = 4
JSV_TUNE_SWEEP_THREADS = 5
JSV_TUNE_GRAPHS = 6
JSV_TUNE_COARSE_M = 7
JSV_TUNE_SWEEP_VARIANT = 8
JSV_TUNE_FUSE_PROLONG = 9
JSV_TUNE_COARSE_LDS = 10
JSV_TUNE_SWEEP_PAIRS = 11
JSV_TUNE_COARSE_SOLO = 12
JSV_TUNE_COARSE_THREADS = 13
JSV_TUNE_FUSE_HALVES = 14
JSV_TUNE_DEVICE_STOP = 15
JSV_TUNE_COARSE_REG = 16
JSV_TUNE_ZIGZAG = 17
JSV_TUNE_FUSE_RESID = 18
JSV_BUF_PHI = 0
JSV_BUF_RHS = 1
JSV_NOALLOC_MAX_SEGMENTS = 1024


class JsvConfig(ctypes.Structure):
    _fields_ = [
        ("fixed", ctypes.c_int),
        ("epsilon", ctypes.c_double),
        ("gamma", ctypes.c_int),
        ("fcycle", ctypes.c_int),
        ("skip_pre", ctypes.c_int),
        ("nu1", ctypes.c_int),
        ("nu2", ctypes.c_int),
        ("max_iters", ctypes.c_int64),
        ("stop", ctypes.c_int),
        ("rtol", ctypes.c_double),
        ("inner_cycles", ctypes.c_int),
        ("check_every", ctypes.c_int),
        ("history", ctypes.POINTER(ctypes.c_double)),
        ("history_len", ctypes.c_int64),
    ]


class JsvResult(ctypes.Structure):
    _fields_ = [
        ("count", ctypes.c_int64),
        ("outer_steps", ctypes.c_int64),
        ("distance", ctypes.c_double),
        ("rel_residual", ctypes.c_double),
        ("nonfinite", ctypes.c_int),
    ]


_P = ctypes.c_void_p
_D = ctypes.c_double
_I = ctypes.c_int
_I64 = ctypes.c_int64
_DP = ctypes.POINTER(ctypes.c_double)

# name -> (restype, argtypes); every function declared in include/jsolver_amd.h
SIGNATURES = {
    "jsv_abi_version": (_I, []),
    "jsv_last_error": (ctypes.c_char_p, []),
    "jsv_plan_create": (_I, [_I, _I, _I, _I, _D, _D, _D, _I, ctypes.POINTER(_P)]),
    "jsv_plan_create_batched": (_I, [_I, _I, _I, _I, _I, _D, _D, _D, _I, ctypes.POINTER(_P)]),
    "jsv_plan_destroy": (_I, [_P]),
    "jsv_plan_levels": (_I, [_P, ctypes.POINTER(_I)]),
    "jsv_plan_level_shape": (_I, [_P, _I, ctypes.POINTER(_I)]),
    "jsv_plan_device_bytes": (_I, [_P, ctypes.POINTER(_I64)]),
    "jsv_plan_set_tuning": (_I, [_P, _I, _I64]),
    "jsv_plan_set_operator": (_I, [_P, _P, _D, _P, _P, _P, _DP, _P, _P]),
    "jsv_solve": (_I, [_P, _P, _P, _P, ctypes.POINTER(JsvConfig), ctypes.POINTER(JsvResult), _P]),
    "jsv_solve_batched": (_I, [_P, _P, _P, _P, ctypes.POINTER(JsvConfig), ctypes.POINTER(JsvResult),
                               ctypes.POINTER(_I64), _DP, _P]),
    "jsv_level_upload": (_I, [_P, _I, _I, _I, _P, _P]),
    "jsv_level_download": (_I, [_P, _I, _I, _I, _P, _P]),
    "jsv_smooth": (_I, [_P, _I, _I, _I, _P]),
    "jsv_restrict": (_I, [_P, _I, _I, _P]),
    "jsv_prolong_add": (_I, [_P, _I, _I, _I, _P]),
    "jsv_cycle": (_I, [_P, _I, ctypes.POINTER(JsvConfig), _I, _P]),
    "jsv_residual": (_I, [_P, _P, _P, _P, _DP, _P]),
    "jsv_max_abs_diff": (_I, [_P, _P, _I64, _DP, _P]),
    "jsv_max_abs_diff_noalloc": (_I, [_P, _P, _I64, _I64, _DP, _P]),
    "jsv_stream_bench": (_I, [_P, _I, _I64, _P, _I, _P]),
    "jsv_dirichlet_lift": (_I, [_P, _P, _P, _P, _P]),
    "jsv_do_bcs_2d": (_I, [_P, _I, _I, _I, _I, _P, _P, _P, _P, _P]),
    "jsv_do_bcs_3d": (_I, [_P, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P]),
    "jsv_forward_checkpoints": (_I, [_P, _P, _P, ctypes.POINTER(JsvConfig), _P, _P]),
    "jsv_adjoint": (_I, [_P, _P, _P, _P, ctypes.POINTER(JsvConfig), _P, _P, _P, _P, _P, _P, _P, _P]),
    "jsv_adjoint_full": (_I, [_P, _P, _P, _P, ctypes.POINTER(JsvConfig), _P, _P, _P, _P, _P, _P, _P, _P, _P,
                              _P]),
    "jsv_adjoint_device_bytes": (_I, [_P, ctypes.POINTER(_I64)]),
    "jsv_tangent": (_I, [_P, _P, _P, _P, _P, _P, ctypes.POINTER(JsvConfig), _P, _P, _P]),
}

_lock = threading.Lock()
_lib = None


class JsvError(RuntimeError):
    pass


def load(path=LIB_PATH):
    """Load the library (once) and attach signatures."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise JsvError(f"{path} not found: build it with `python -m jsolver_amd.build`")
        lib = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def check(status):
    if status != 0:
        msg = load().jsv_last_error()
        raise JsvError(msg.decode() if msg else f"jsolver_amd error {status}")


def lib_path_loaded():
    return LIB_PATH if _lib is not None else None
