"""Loader for the in-tree HIP library (libdsdhip.so) and its C signatures.

The product has no CPU fallback: if the library is missing or no GPU is
visible, calls fail loudly.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdsdhip.so")

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_i64p = C.POINTER(C.c_int64)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p


class DsdHipError(RuntimeError):
    pass


class NtfRow(C.Structure):
    _fields_ = [("a", C.c_double * 10), ("g", C.c_double * 10), ("order", C.c_int),
                ("freq", C.c_uint), ("name", C.c_char_p)]


_SIGS = {
    "dsdhip_version": ([], C.c_char_p),
    "dsdhip_ntf_count": ([], C.c_int),
    "dsdhip_ntf_get": ([C.c_int], C.POINTER(NtfRow)),
    "dsdhip_ntf_lookup": ([C.c_int, C.c_uint], C.c_int),
    "dsdhip_ntf_auto": ([C.c_uint], C.c_int),
    "gpu_hip_probe": ([], C.c_bool),
    "dsdhip_trellis_create": ([C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int], _vp),
    "dsdhip_trellis_destroy": ([_vp], None),
    "dsdhip_trellis_set_stream": ([_vp, _vp], C.c_int),
    "dsdhip_trellis_reset": ([_vp, C.c_int], C.c_int),
    "dsdhip_trellis_copy": ([_vp, C.c_int, C.c_int], C.c_int),
    "dsdhip_trellis_seed": ([_vp, C.c_int, _dp], C.c_int),
    "dsdhip_trellis_best_state": ([_vp, C.c_int, _dp], C.c_int),
    "dsdhip_trellis_process": ([_vp, C.c_int, _dp, _fp, C.c_size_t, C.POINTER(C.c_size_t)], C.c_int),
    "dsdhip_trellis_run_jobs": ([_vp, C.c_int, _ip, _ip, C.POINTER(_vp), _i64p, _i64p,
                                 C.POINTER(_vp), _dp, _i64p, _i64p, _ip], C.c_int),
    "dsdhip_trellis_force_literal": ([_vp, C.c_int], C.c_int),
}

_lib = None


def lib():
    """The loaded library (raises DsdHipError if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DsdHipError("libdsdhip.so is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def check(rc, what):
    if rc != 0:
        raise DsdHipError("%s failed (rc=%d)" % (what, rc))
