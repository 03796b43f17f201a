"""ctypes binding of liblubatch.so (include/lubatch.h, include/lubatch_util.h).

The library is built in-tree by ``python -m matrixinversion_amd.build`` into
``matrixinversion_amd/_lib/liblubatch.so``.  There is no fallback: if the
library is missing or cannot be loaded, ``lib()`` raises.
"""
from __future__ import annotations

import ctypes
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "liblubatch.so")
CMP_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "liblubref_rocsolver.so")

_vp = ctypes.c_void_p
_i = ctypes.c_int
_i64 = ctypes.c_int64
_sz = ctypes.c_size_t

# name -> (restype, argtypes); the C-ABI of include/lubatch.h first.
ABI = {
    "lu_batched_inverse": (_i, [_vp, _vp, _i, _i, _i]),
    "lu_batched_inverse_async": (_i, [_vp, _vp, _i, _i, _i, _vp]),
    "lu_batched_inverse_async64": (_i, [_vp, _vp, _i, _i64, _i, _vp]),
    "lu_batched_residual": (_i, [_vp, _vp, _i, _i64, ctypes.c_float, _vp, _vp, _vp]),
    "lu_ragged_plan_create": (_i, [ctypes.POINTER(_vp), _i64]),
    "lu_ragged_plan_destroy": (_i, [_vp]),
    "lu_batched_inverse_ragged_plan": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _i, _vp]),
    "lu_batched_inverse_ragged": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _i, _vp]),
    "lu_ragged_invalid_count": (_i, [_vp, ctypes.POINTER(_i)]),
    "lu_batched_inverse_ragged_noalloc": (_i, [_vp, _vp, _vp, _vp, _vp, _i64, _i, _vp]),
    "lu_tuning_set": (_i, [ctypes.c_char_p, ctypes.c_char_p]),
    "lu_batched_refine": (_i, [_vp, _vp, _i, _i64, _vp]),
    "lu_batched_inverse_lapack": (_i, [_vp, _vp, _vp, _i, _i64, _i, _vp]),
    "lu_perm_to_ipiv": (_i, [_vp, _i, _i64, _vp]),
    "lu_batched_inverse_refined": (_i, [_vp, _vp, _vp, _i, _i64, _i, _vp]),
    "lu_server_create": (_i, [ctypes.POINTER(_vp), _i64, _i64, _i64]),
    "lu_server_destroy": (_i, [_vp]),
    "lu_server_submit": (_i, [_vp, _vp, _vp, _i, _i, ctypes.POINTER(ctypes.c_uint64)]),
    "lu_server_progress": (_i, [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "lu_server_query": (_i, [_vp, ctypes.c_uint64, ctypes.POINTER(_i), ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                             ctypes.POINTER(ctypes.c_double)]),
    "lu_server_release": (_i, [_vp, ctypes.c_uint64]),
    "lu_server_poll": (_i, [_vp, ctypes.POINTER(ctypes.c_uint64), _i, ctypes.POINTER(_i)]),
    "lu_server_counters": (_i, [_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i)]),
    "lu_server_stats": (_i, [_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i64), ctypes.POINTER(_i64)]),
    "lu_tuning_get": (_i, [ctypes.c_char_p, ctypes.c_char_p]),
    "lu_tuning_set_lds_pad": (_i, [ctypes.c_char_p]),
    "lu_tuning_get_lds_pad": (_i, [ctypes.c_char_p]),
    "lu_status_string": (ctypes.c_char_p, [_i]),
    "lu_last_error": (ctypes.c_char_p, []),
    "lu_version": (ctypes.c_char_p, []),
    "lu_build_id": (ctypes.c_char_p, []),
}

UTIL = {
    "lubx_device_count": (_i, []),
    "lubx_device_info": (_i, [_i, ctypes.c_char_p, _i, ctypes.c_char_p, _i, ctypes.POINTER(_i),
                              ctypes.POINTER(_i), ctypes.POINTER(_sz), ctypes.POINTER(_i), ctypes.POINTER(_i),
                              ctypes.POINTER(_sz)]),
    "lubx_malloc": (_vp, [_sz]),
    "lubx_free": (_i, [_vp]),
    "lubx_host_alloc": (_vp, [_sz]),
    "lubx_host_free": (_i, [_vp]),
    "lubx_memcpy_htod": (_i, [_vp, _vp, _sz]),
    "lubx_memcpy_dtoh": (_i, [_vp, _vp, _sz]),
    "lubx_memcpy_dtod_async": (_i, [_vp, _vp, _sz, _vp]),
    "lubx_memset": (_i, [_vp, _i, _sz]),
    "lubx_device_sync": (_i, []),
    "lubx_stream_create": (_vp, []),
    "lubx_stream_destroy": (_i, [_vp]),
    "lubx_stream_sync": (_i, [_vp]),
    "lubx_event_create": (_vp, []),
    "lubx_event_destroy": (_i, [_vp]),
    "lubx_event_record": (_i, [_vp, _vp]),
    "lubx_event_sync": (_i, [_vp]),
    "lubx_event_elapsed_ms": (ctypes.c_float, [_vp, _vp]),
    "lubx_copy_f32": (_i, [_vp, _vp, _sz, _vp]),
    "lubx_serve_replay": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "lubx_gather_f32": (_i, [_vp, _vp, _vp, _i, _i64, _vp]),
    "lubx_scatter_f32": (_i, [_vp, _vp, _vp, _i, _i64, _vp]),
    "lubx_graph_capture_begin": (_i, [_vp]),
    "lubx_graph_capture_end": (_i, [_vp, ctypes.POINTER(_vp)]),
    "lubx_graph_launch": (_i, [_vp, _vp]),
    "lubx_graph_destroy": (_i, [_vp]),
}

_lib = None


class NativeError(RuntimeError):
    pass


def lib() -> ctypes.CDLL:
    """Load liblubatch.so once; raise if it is absent (no CPU fallback exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeError(f"{LIB_PATH} not found: build it with `python -m matrixinversion_amd.build`")
        l = ctypes.CDLL(LIB_PATH)
        for table in (ABI, UTIL):
            for name, (res, args) in table.items():
                fn = getattr(l, name)
                fn.restype = res
                fn.argtypes = args
        _lib = l
    return _lib


# include/lubatch_compare.h (benchmark-only rocSOLVER comparator, its own library)
COMPARE = {
    "lubref_rocsolver_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p]),
    "lubref_rocsolver_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "lubref_rocsolver_inverse": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.c_int, ctypes.c_int]),
}
_cmp = None


def cmp_lib() -> ctypes.CDLL:
    """Load the rocSOLVER comparator library (benchmarks and tests only)."""
    global _cmp
    if _cmp is None:
        if not os.path.exists(CMP_LIB_PATH):
            raise NativeError(f"{CMP_LIB_PATH} not found (built when rocSOLVER is installed)")
        l = ctypes.CDLL(CMP_LIB_PATH)
        for name, (res, args) in COMPARE.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _cmp = l
    return _cmp


def check(status: int, what: str = "") -> None:
    if status != 0:
        l = lib()
        msg = l.lu_status_string(status).decode()
        detail = l.lu_last_error().decode()
        raise NativeError(f"{what or 'liblubatch'} failed: status {status} ({msg}) {detail}".strip())
