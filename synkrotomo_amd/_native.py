"""ctypes binding of libsynkro.so (include/synkro.h).

The library is the only compute path: if it is missing or the device is absent, the
functions here raise -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
import threading

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "lib", "libsynkro.so")

SK_OK, SK_EINVAL, SK_ETYPE, SK_EHIP, SK_ENOMEM = 0, 1, 2, 3, 4
SK_OPT_EXACT, SK_OPT_FP_VARIANT, SK_OPT_BP_VARIANT = 1, 2, 3
SK_MAT_FP, SK_MAT_BP, SK_MAT_ADJ = 0, 1, 2

_lib = None
_lock = threading.Lock()

_vp = ctypes.c_void_p
_i = ctypes.c_int
_f = ctypes.c_float
_f32p = ctypes.POINTER(ctypes.c_float)
_i64p = ctypes.POINTER(ctypes.c_int64)
_ip = ctypes.POINTER(ctypes.c_int)

# name -> (restype, argtypes); every symbol declared in include/synkro.h
SIGNATURES = {
    "sk_create": (_i, [_i, ctypes.POINTER(_vp)]),
    "sk_destroy": (None, [_vp]),
    "sk_last_error": (ctypes.c_char_p, [_vp]),
    "sk_set_option": (_i, [_vp, _i, _i]),
    "sk_get_option": (_i, [_vp, _i, _ip]),
    "sk_geom_create": (_i, [_vp, _f32p, _i, _f, _f, _i, _i, ctypes.POINTER(_vp)]),
    "sk_geom_destroy": (None, [_vp]),
    "sk_geom_query": (_i, [_vp, _ip, _ip, _ip, _ip, _ip]),
    "sk_lines_compute": (_i, [_f32p, _i, _f32p]),
    "sk_compare": (_i, [_vp, ctypes.c_int64, _vp, _vp, _f, _ip, _i64p, _f32p, _vp]),
    "sk_fp": (_i, [_vp, _vp, _i, _vp, _vp, _vp]),
    "sk_fp_noalloc": (_i, [_vp, _vp, _i, _vp, _vp, _vp]),
    "sk_bp": (_i, [_vp, _vp, _i, _vp, _vp, _vp]),
    "sk_bp_adjoint": (_i, [_vp, _vp, _i, _vp, _vp, _vp]),
    "sk_bp_transpose": (_i, [_vp, _vp, _i, _vp, _vp, _vp]),
    "sk_sirt": (_i, [_vp, _vp, _i, _vp, _vp, _i, _vp]),
    "sk_sirt_weights": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "sk_sirt_step": (_i, [_vp, _vp, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sk_sirt_batched": (_i, [_vp, _vp, _i, _i, _vp]),
    "sk_autotune": (_i, [_vp, _vp, _ip, _ip, _vp]),
    "sk_autotune_batch": (_i, [_vp, _vp, _i, _ip, _ip, _vp]),
    "sk_geom_set_variants": (_i, [_vp, _i, _i]),
    "sk_conv32": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sk_conv_in": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp]),
    "sk_conv_in2": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp]),
    "sk_conv_out": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sk_conv_out_dgrad": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _vp]),
    "sk_conv_in_dgrad": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp]),
    "sk_conv_in2_dgrad": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp]),
    "sk_conv_wgrad": (_i, [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp]),
    "sk_phantoms": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp]),
    "sk_sparse_build": (_i, [_vp, _vp, _i, ctypes.POINTER(_vp), _vp]),
    "sk_sparse_count": (_i, [_vp, _vp, _i, _i64p, _vp]),
    "sk_mat_info": (_i, [_vp, _i64p, _i64p, _i64p]),
    "sk_mat_arrays": (_i, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
    "sk_mat_copy": (_i, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "sk_mat_set_spmv_variant": (_i, [_vp, _i]),
    "sk_mat_spmv_variant": (_i, [_vp]),
    "sk_spmv": (_i, [_vp, _vp, _i, _vp, _vp, _vp]),
    "sk_mat_destroy": (None, [_vp]),
    "sk_normalize": (_i, [_vp, _i, _i, _i, _vp, _i, _vp, _i, _vp, _f, _f, _vp, _vp]),
    "sk_row_centroid": (_i, [_vp, ctypes.c_int64, _i, _vp, _vp, _vp]),
}


class SkProblem(ctypes.Structure):
    """sk_problem (include/synkro.h)."""
    _fields_ = [("geom", _vp), ("batch", _i), ("d_img", _vp), ("d_sino", _vp)]


class NativeMissing(RuntimeError):
    pass


def lib():
    """Load libsynkro.so once and declare its signatures."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise NativeMissing(
                f"{LIB_PATH} is not built; run `python -m synkrotomo_amd.build` "
                "(the HIP library is the only compute path)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int, ctx=None, what: str = "") -> None:
    """Map a C-ABI status to the Python exception the reference's wrapper would raise."""
    if rc == SK_OK:
        return
    msg = ""
    if ctx is not None:
        raw = lib().sk_last_error(ctx)
        msg = raw.decode() if raw else ""
    text = f"{what}: {msg}" if what else msg
    if rc == SK_EINVAL:
        raise ValueError(text)
    if rc == SK_ETYPE:
        raise TypeError(text)
    if rc == SK_ENOMEM:
        raise MemoryError(text)
    raise RuntimeError(text or f"libsynkro error {rc}")
