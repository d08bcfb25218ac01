"""ctypes binding of libjoker.so (the C ABI declared in include/joker.h).

The library is the only compute path of this package: if it is missing, or
the process has no gfx950 device, every compute call raises -- there is no
CPU fallback.
"""
import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libjoker.so")
ABI_VERSION = 5

JK_OK = 0
JK_JITTER_REFERENCE = 0
JK_JITTER_PHYSICAL = 1
JK_KPRIOR_NORMAL = 0
JK_KPRIOR_FIXED_COMPANION_MASS = 1
JK_LMAX = 8

_lock = threading.RLock()
_lib = None
_inited = set()

_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64
_dp = ctypes.POINTER(ctypes.c_double)

# (name, restype, argtypes) of every symbol include/joker.h declares.
SIGNATURES = [
    ("jk_status_string", ctypes.c_char_p, [ctypes.c_int]),
    ("jk_abi_version", ctypes.c_int, []),
    ("jk_init", ctypes.c_int, [ctypes.c_int]),
    ("jk_release_scratch", ctypes.c_int, [ctypes.c_int]),
    ("jk_star_create", ctypes.c_int,
     [ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int,
      ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.POINTER(_vp)]),
    ("jk_star_create_cov", ctypes.c_int,
     [ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int,
      ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.POINTER(_vp)]),
    ("jk_star_destroy", ctypes.c_int, [_vp]),
    ("jk_lnl_f64", ctypes.c_int, [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("jk_lnl_mixed", ctypes.c_int, [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("jk_max_f64", ctypes.c_int, [_vp, _i64, _vp, _vp]),
    ("jk_polish_f64", ctypes.c_int,
     [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_double, _i64, _vp, _vp]),
    ("jk_reject", ctypes.c_int,
     [_vp, _i64, _u64, _u64, _u64, _u64, _u64, _i64, _vp, _vp, _vp, _vp]),
    ("jk_lnl_reject_f64", ctypes.c_int,
     [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _u64, _u64, _u64, _u64, _u64, _i64, _vp, _vp,
      _vp, _vp]),
    ("jk_lnl_reject_f64_noalloc", ctypes.c_int,
     [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _u64, _u64, _u64, _u64, _u64, _i64, _vp, _vp,
      _vp, _vp]),
    ("jk_debug_single", ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("jk_count_failures", ctypes.c_int, [_vp, _i64, _vp, _vp]),
    ("jk_reject_with_max", ctypes.c_int,
     [_vp, _i64, _vp, _u64, _u64, _u64, _u64, _u64, _i64, _vp, _vp, _vp]),
    ("jk_pcg64_uniform", ctypes.c_int, [_u64, _u64, _u64, _u64, _u64, _i64, _vp, _vp]),
    ("jk_batch_create", ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, ctypes.POINTER(_vp)]),
    ("jk_batch_destroy", ctypes.c_int, [_vp]),
    ("jk_batch_create_slots", ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                              ctypes.POINTER(_vp)]),
    ("jk_batch_set_star", ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp]),
    ("jk_batch_set_exact", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp]),
    ("jk_lnl_batch", ctypes.c_int,
     [_vp, ctypes.c_int, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int,
      ctypes.c_int, _vp, _vp]),
    ("jk_lnl_prune_batch", ctypes.c_int,
     [_vp, _vp, _vp, ctypes.c_int, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
      ctypes.c_int, ctypes.c_int, _vp, _vp]),
    ("jk_reject_batch", ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp]),
    ("jk_reject_batch_zero_check", ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("jk_polish_batch", ctypes.c_int,
     [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_double, _vp, _vp]),
    ("jk_settle_batch", ctypes.c_int,
     [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("jk_settle_reject_batch", ctypes.c_int,
     [_vp, ctypes.c_int, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
      _vp]),
    ("jk_settle_live_batch", ctypes.c_int,
     [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_double, _vp, _vp]),
    ("jk_settle_tol_batch", ctypes.c_int,
     [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_float, _vp, _vp]),
    ("jk_settle_pruned_batch", ctypes.c_int,
     [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_double,
      ctypes.c_double, _vp, _vp]),
    ("jk_seg_plan_create", ctypes.c_int,
     [_vp, ctypes.c_int, _vp, _i64, ctypes.POINTER(_vp)]),
    ("jk_seg_plan_destroy", ctypes.c_int, [_vp]),
    ("jk_seg_plan_set_streams", ctypes.c_int, [_vp, _vp, _vp]),
    ("jk_seg_plan_settle", ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("jk_seg_plan_reject", ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    ("jk_seg_plan_settle_reject", ctypes.c_int,
     [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("jk_prior_fill", ctypes.c_int,
     [_i64, _i64, _u64, ctypes.c_double, ctypes.c_double, _vp, ctypes.c_int, ctypes.c_double,
      ctypes.c_double, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("jk_linear_posterior_batch", ctypes.c_int,
     [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("jk_linear_posterior", ctypes.c_int,
     [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
]


class JokerLibError(RuntimeError):
    pass


def load(path=LIB_PATH):
    """Load libjoker.so and bind every exported symbol (no device needed)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise JokerLibError(
                f"libjoker.so not found at {path}: build it with "
                "`python -m thejoker_amd.build` (there is no CPU fallback)")
        lib = ctypes.CDLL(path)
        for name, res, args in SIGNATURES:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        if lib.jk_abi_version() != ABI_VERSION:
            raise JokerLibError("libjoker.so ABI version mismatch; rebuild it")
        _lib = lib
        return lib


def check(status, what="libjoker call"):
    if status != JK_OK:
        msg = load().jk_status_string(status).decode()
        raise JokerLibError(f"{what} failed: {msg} (status {status})")


def ensure_device(device=0):
    """Initialise libjoker on `device` (selects it as the current HIP device)."""
    lib = load()
    with _lock:
        if device not in _inited:
            check(lib.jk_init(int(device)), f"jk_init({device})")
            _inited.add(device)
        else:
            check(lib.jk_init(int(device)), f"jk_init({device})")
    return lib
