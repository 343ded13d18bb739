"""The engine's pre-SDM conditioning on the GPU (dsdhip_cond_*): boxcar
and lowpass of same-rate DSD streams, the boxcar pre-smooth of PCM before
rate conversion, gain ramp and soft clip.  Mirrors the stages of
engine_process_block / engine_process_fir_gain (src/engine.c:478-899)."""
import ctypes as C

import numpy as np
import torch

from . import _lib

_vp = C.c_void_p


def _sigs(L):
    if getattr(L, "_cond_sigs", False):
        return
    L.dsdhip_lowpass_taps.argtypes = [C.c_uint, C.POINTER(C.c_double)]
    L.dsdhip_cond_create.restype = _vp
    L.dsdhip_cond_create.argtypes = [C.c_int, C.c_int, C.c_uint]
    L.dsdhip_cond_destroy.argtypes = [_vp]
    L.dsdhip_cond_reset.argtypes = [_vp]
    L.dsdhip_cond_set_stream.argtypes = [_vp, _vp]
    for f in ("dsdhip_cond_dsd_boxcar_dev", "dsdhip_cond_dsd_lowpass_dev"):
        getattr(L, f).argtypes = [_vp, _vp, C.c_int64, C.c_size_t, C.c_double, C.c_double, _vp, C.c_int64]
    L.dsdhip_cond_pcm_boxcar_dev.argtypes = [_vp, _vp, C.c_int64, C.c_size_t, _vp, C.c_int64]
    L.dsdhip_gain_ramp_dev.argtypes = [_vp, C.c_int64, C.c_int, C.c_size_t, C.c_double, C.c_double, _vp]
    L.dsdhip_soft_clip_dev.argtypes = [_vp, C.c_int64, C.c_int, C.c_size_t, _vp]
    L.dsdhip_f32_to_bits_dev.argtypes = [_vp, C.c_int64, C.c_int, C.c_size_t, _vp, C.c_int64, _vp]
    L._cond_sigs = True


def _L():
    L = _lib.lib()
    _sigs(L)
    return L


def lowpass_taps(rate):
    h = np.zeros(127)
    _L().dsdhip_lowpass_taps(rate, h.ctypes.data_as(C.POINTER(C.c_double)))
    return h


class Conditioner:
    """nch channels; boxcar taps (the engine's: 8 for rate conversion, 32/64/16
    same-rate by rate); lowpass_rate: design rate of the 127-tap lowpass or 0."""

    def __init__(self, nch, taps, lowpass_rate=0):
        self.L = _L()
        self.nch = nch
        self.h = self.L.dsdhip_cond_create(nch, taps, lowpass_rate)
        if not self.h:
            raise _lib.DsdHipError("dsdhip_cond_create failed")

    def _out(self, n, like):
        return torch.empty((self.nch, n), dtype=torch.float64, device=like.device)

    def dsd_boxcar(self, bits, n, gs, ge):
        out = self._out(n, bits)
        _lib.check(self.L.dsdhip_cond_dsd_boxcar_dev(self.h, bits.data_ptr(), bits.stride(0), n, gs, ge,
                                                     out.data_ptr(), n), "dsdhip_cond_dsd_boxcar_dev")
        return out

    def dsd_lowpass(self, bits, n, gs, ge):
        out = self._out(n, bits)
        _lib.check(self.L.dsdhip_cond_dsd_lowpass_dev(self.h, bits.data_ptr(), bits.stride(0), n, gs, ge,
                                                      out.data_ptr(), n), "dsdhip_cond_dsd_lowpass_dev")
        return out

    def pcm_boxcar(self, x):
        n = x.shape[1]
        out = self._out(n, x)
        _lib.check(self.L.dsdhip_cond_pcm_boxcar_dev(self.h, x.data_ptr(), x.stride(0), n, out.data_ptr(), n),
                   "dsdhip_cond_pcm_boxcar_dev")
        return out

    def reset(self):
        _lib.check(self.L.dsdhip_cond_reset(self.h), "dsdhip_cond_reset")

    def __del__(self):
        try:
            if self.h:
                self.L.dsdhip_cond_destroy(self.h)
                self.h = None
        except Exception:
            pass


def gain_ramp(x, gs, ge):
    """in place on cuda float64 (nch, n)"""
    _lib.check(_L().dsdhip_gain_ramp_dev(x.data_ptr(), x.stride(0), x.shape[0], x.shape[1], gs, ge, None),
               "dsdhip_gain_ramp_dev")
    return x


def soft_clip(x):
    _lib.check(_L().dsdhip_soft_clip_dev(x.data_ptr(), x.stride(0), x.shape[0], x.shape[1], None),
               "dsdhip_soft_clip_dev")
    return x


def to_bits(x):
    """cuda float32 (nch, n) of +-1 -> packed bits (x >= 0 -> 1)"""
    nch, n = x.shape
    out = torch.zeros((nch, (n + 63) // 64), dtype=torch.int64, device=x.device)
    _lib.check(_L().dsdhip_f32_to_bits_dev(x.data_ptr(), x.stride(0), nch, n, out.data_ptr(), out.shape[1], None),
               "dsdhip_f32_to_bits_dev")
    return out
