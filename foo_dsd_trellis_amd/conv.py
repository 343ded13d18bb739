"""UPOLS convolution on the GPU (dsdhip_conv_* in libdsdhip.so): mirrors the
reference's conv_init / conv_load_ir / conv_process
(src/convolution.c:530-1041) for an IR given per channel at the signal
rate."""
import ctypes as C

import numpy as np

from . import _lib

_dp = C.POINTER(C.c_double)


def _sigs(L):
    if getattr(L, "_conv_sigs", False):
        return
    L.dsdhip_conv_create.restype = C.c_void_p
    L.dsdhip_conv_create.argtypes = [C.c_int, C.c_int, _dp, C.c_int64, C.c_int64]
    L.dsdhip_conv_destroy.argtypes = [C.c_void_p]
    L.dsdhip_conv_reset.argtypes = [C.c_void_p]
    L.dsdhip_conv_set_stream.argtypes = [C.c_void_p, C.c_void_p]
    L.dsdhip_conv_partitions.argtypes = [C.c_void_p]
    L.dsdhip_conv_process_dev.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_size_t, C.c_void_p,
                                          C.c_int64]
    L.dsdhip_conv_process_fifo_dev.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_size_t, C.c_void_p,
                                               C.c_int64]
    L.dsdhip_conv_fifo_state.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L._conv_sigs = True


class Upols:
    """Streaming convolution of nch channels with ir (nch, taps), partition P."""

    def __init__(self, ir, P=8192):
        self.lib = _lib.lib()
        _sigs(self.lib)
        ir = np.ascontiguousarray(np.atleast_2d(ir), dtype=np.float64)
        self.nch, self.taps = ir.shape
        self.P = P
        self.h = self.lib.dsdhip_conv_create(self.nch, P, ir.ctypes.data_as(_dp), self.taps, self.taps)
        if not self.h:
            raise _lib.DsdHipError("dsdhip_conv_create failed")

    @property
    def partitions(self):
        return self.lib.dsdhip_conv_partitions(self.h)

    def reset(self):
        _lib.check(self.lib.dsdhip_conv_reset(self.h), "dsdhip_conv_reset")

    def set_stream(self, stream):
        _lib.check(self.lib.dsdhip_conv_set_stream(self.h, stream), "dsdhip_conv_set_stream")

    def process_dev(self, d_in, in_stride, n, d_out, out_stride):
        """n a multiple of P."""
        _lib.check(self.lib.dsdhip_conv_process_dev(self.h, d_in, in_stride, n, d_out, out_stride),
                   "dsdhip_conv_process_dev")

    def process_fifo_dev(self, d_in, in_stride, n, d_out, out_stride):
        """Any n, with the reference's FIFO (conv_process_direct,
        src/convolution.c:893-958)."""
        _lib.check(self.lib.dsdhip_conv_process_fifo_dev(self.h, d_in, in_stride, n, d_out, out_stride),
                   "dsdhip_conv_process_fifo_dev")

    def fifo_state(self):
        """(pending input samples, produced samples waiting) per channel."""
        a, b = C.c_int64(0), C.c_int64(0)
        _lib.check(self.lib.dsdhip_conv_fifo_state(self.h, C.byref(a), C.byref(b)), "dsdhip_conv_fifo_state")
        return a.value, b.value

    def __del__(self):
        try:
            if self.h:
                self.lib.dsdhip_conv_destroy(self.h)
                self.h = None
        except Exception:
            pass
