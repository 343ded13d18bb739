"""Streaming chain for one multichannel stream (dsdhip_stream_*): FIR
upsampling, optional UPOLS convolution, DAS, per chunk on the GPU.  Mirrors
engine_process_block (src/engine.c:478-729) with DAS as the SDM."""
import ctypes as C

import numpy as np
import torch

from . import _lib, fir, trellis

_dp = C.POINTER(C.c_double)


def _sigs(L):
    if getattr(L, "_stream_sigs", False):
        return
    L.dsdhip_stream_create.restype = C.c_void_p
    L.dsdhip_stream_create.argtypes = [C.c_uint, C.c_uint, C.c_double, C.c_int, C.c_size_t, _dp,
                                       C.c_int64, C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.c_int, C.c_double, C.c_size_t, C.c_int]
    L.dsdhip_stream_destroy.argtypes = [C.c_void_p]
    L.dsdhip_stream_set_stream.argtypes = [C.c_void_p, C.c_void_p]
    L.dsdhip_stream_reset.argtypes = [C.c_void_p]
    L.dsdhip_stream_ratio.argtypes = [C.c_void_p]
    L.dsdhip_stream_process_dev.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_size_t, C.c_void_p,
                                            C.c_int64, C.POINTER(C.c_size_t)]
    L.dsdhip_stream_copy_signal.restype = C.c_int64
    L.dsdhip_stream_copy_signal.argtypes = [C.c_void_p, _dp, C.c_int64]
    L._stream_sigs = True


class Stream:
    """fs_in PCM (float32, nch channels) -> fs_out DSD bits, chunk by chunk.
    ir: None or (nch, taps) float64 room impulse responses at fs_out."""

    def __init__(self, fs_in, fs_out, nch, chunk_frames, ir=None, P=8192, ntf_index=None, K=32, L=64,
                 depth=8, gain=fir.PCM_FIR_GAIN, state_limit=0.0, seg_len=65536, reposition=True):
        self.lib = _lib.lib()
        _sigs(self.lib)
        if ntf_index is None:
            ntf_index = trellis.ntf_auto(fs_out)
        self.nch, self.chunk_frames = nch, chunk_frames
        irp, ir_len = None, 0
        if ir is not None:
            self._ir = np.ascontiguousarray(np.atleast_2d(ir), dtype=np.float64)
            assert self._ir.shape[0] == nch
            irp, ir_len = self._ir.ctypes.data_as(_dp), self._ir.shape[1]
        self.h = self.lib.dsdhip_stream_create(fs_in, fs_out, gain, nch, chunk_frames, irp, ir_len, ir_len,
                                               P, ntf_index, depth, K, L, state_limit, seg_len,
                                               1 if reposition else 0)
        if not self.h:
            raise _lib.DsdHipError("dsdhip_stream_create failed")
        self.R = self.lib.dsdhip_stream_ratio(self.h)
        self.words = (chunk_frames * self.R + 63) // 64
        self.bits = torch.zeros((nch, self.words), dtype=torch.int64, device="cuda")

    def process(self, pcm):
        """pcm: cuda float32 (nch, frames) with frames <= chunk_frames.
        Returns the DSD samples produced per channel; the packed bits are in
        self.bits (bit 1 = +1.0, little-endian within 64-bit words)."""
        assert pcm.dtype == torch.float32 and pcm.is_cuda and pcm.shape[0] == self.nch
        n = C.c_size_t(0)
        _lib.check(self.lib.dsdhip_stream_process_dev(self.h, pcm.data_ptr(), pcm.stride(0), pcm.shape[1],
                                                      self.bits.data_ptr(), self.words, C.byref(n)),
                   "dsdhip_stream_process_dev")
        return n.value

    def bits_numpy(self, count):
        w = self.bits.cpu().numpy().view(np.uint64)
        b = np.unpackbits(w.view(np.uint8), axis=1, bitorder="little")[:, :count]
        return np.where(b > 0, 1.0, -1.0).astype(np.float32)

    def signal(self):
        """SDM input of the last chunk (nch, n) float64, from the device."""
        buf = np.zeros((self.nch, self.chunk_frames * self.R))
        n = self.lib.dsdhip_stream_copy_signal(self.h, buf.ctypes.data_as(_dp), buf.shape[1])
        if n < 0:
            raise _lib.DsdHipError("dsdhip_stream_copy_signal failed")
        return buf[:, :n]

    def reset(self):
        _lib.check(self.lib.dsdhip_stream_reset(self.h), "dsdhip_stream_reset")

    def set_stream(self, stream):
        _lib.check(self.lib.dsdhip_stream_set_stream(self.h, stream), "dsdhip_stream_set_stream")

    def __del__(self):
        try:
            if self.h:
                self.lib.dsdhip_stream_destroy(self.h)
                self.h = None
        except Exception:
            pass
