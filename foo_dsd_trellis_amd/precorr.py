"""PreCorr SDM on the GPU (dsdhip_precorr_*): greedy single-precision SDM
with the prediction-table correction, one sequential recurrence per
channel.  Mirrors precorr_context_init / precorr_process_block /
precorr_context_reset (include/precorr.h:40-52)."""
import ctypes as C

import numpy as np
import torch

from . import _lib

_vp = C.c_void_p


class PcState(C.Structure):
    """dsdhip_precorr_state_t (= gpu_precorr_state_t)"""
    _fields_ = [("state", C.c_float * 8), ("prev_y", C.c_float), ("history", C.c_int), ("phase", C.c_int),
                ("pad", C.c_float)]


def _sigs(L):
    if getattr(L, "_pc_sigs", False):
        return
    L.dsdhip_precorr_create.restype = _vp
    L.dsdhip_precorr_create.argtypes = [C.c_int, C.c_int, C.c_float]
    L.dsdhip_precorr_destroy.argtypes = [_vp]
    L.dsdhip_precorr_reset.argtypes = [_vp]
    L.dsdhip_precorr_table.argtypes = [_vp, C.POINTER(C.c_float)]
    L.dsdhip_precorr_get_state.argtypes = [_vp, C.c_int, C.POINTER(PcState)]
    L.dsdhip_precorr_set_state.argtypes = [_vp, C.c_int, C.POINTER(PcState)]
    L.dsdhip_precorr_process_dev.argtypes = [_vp, _vp, C.c_int64, C.c_size_t, _vp, C.c_int64]
    L.gpu_hip_precorr_process.argtypes = [_vp, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_size_t,
                                          C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int,
                                          C.POINTER(C.c_float), C.POINTER(PcState), C.POINTER(PcState)]
    L._pc_sigs = True


class PreCorr:
    def __init__(self, ntf_index, nch, state_limit=0.0):
        self.L = _lib.lib()
        _sigs(self.L)
        self.nch = nch
        self.h = self.L.dsdhip_precorr_create(ntf_index, nch, state_limit)
        if not self.h:
            raise _lib.DsdHipError("dsdhip_precorr_create failed")

    def table(self):
        t = np.zeros(2048, np.float32)
        _lib.check(self.L.dsdhip_precorr_table(self.h, t.ctypes.data_as(C.POINTER(C.c_float))), "table")
        return t

    def state(self, ch):
        s = PcState()
        _lib.check(self.L.dsdhip_precorr_get_state(self.h, ch, C.byref(s)), "get_state")
        return np.ctypeslib.as_array(s.state).copy(), int(s.history), int(s.phase), float(s.prev_y)

    def process(self, x):
        """cuda float64 (nch, n) -> packed bits int64 (nch, words)"""
        n = x.shape[1]
        out = torch.zeros((self.nch, (n + 63) // 64), dtype=torch.int64, device=x.device)
        _lib.check(self.L.dsdhip_precorr_process_dev(self.h, x.data_ptr(), x.stride(0), n, out.data_ptr(),
                                                     out.shape[1]), "dsdhip_precorr_process_dev")
        return out

    def reset(self):
        _lib.check(self.L.dsdhip_precorr_reset(self.h), "reset")

    def __del__(self):
        try:
            if self.h:
                self.L.dsdhip_precorr_destroy(self.h)
                self.h = None
        except Exception:
            pass


def seam_process(x, a, g, table, init=None):
    """gpu_hip_precorr_process on host arrays; returns (+-1 float32, final PcState)."""
    L = _lib.lib()
    _sigs(L)
    xf = np.ascontiguousarray(x, dtype=np.float32)
    out = np.zeros(xf.size, np.float32)
    af = np.ascontiguousarray(a, dtype=np.float32)
    gf = np.ascontiguousarray(g, dtype=np.float32)
    tf = np.ascontiguousarray(table, dtype=np.float32)
    fin = PcState()
    fp = C.POINTER(C.c_float)
    rc = L.gpu_hip_precorr_process(None, xf.ctypes.data_as(fp), out.ctypes.data_as(fp), xf.size,
                                   af.ctypes.data_as(fp), gf.ctypes.data_as(fp), af.size, tf.ctypes.data_as(fp),
                                   C.byref(init) if init is not None else None, C.byref(fin))
    if rc != 0:
        raise _lib.DsdHipError("gpu_hip_precorr_process failed")
    return out, fin
