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
    L.dsdhip_stream_create_cfg.restype = C.c_void_p
    L.dsdhip_stream_create_cfg.argtypes = [C.POINTER(StreamCfg)]
    L.dsdhip_stream_set_gain.argtypes = [C.c_void_p, C.c_double]
    L.dsdhip_stream_process_dsd_dev.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_size_t, C.c_void_p,
                                                C.c_int64, C.POINTER(C.c_size_t)]
    L.dsdhip_stream_copy_signal.restype = C.c_int64
    L.dsdhip_stream_copy_signal.argtypes = [C.c_void_p, _dp, C.c_int64]
    L.dsdhip_stream_last_timeline.argtypes = [C.c_void_p, _dp]
    L._stream_sigs = True


IN_PCM, IN_DSD = 0, 1


class StreamCfg(C.Structure):
    """dsdhip_stream_cfg_t"""
    _fields_ = [("fs_in", C.c_uint), ("fs_out", C.c_uint), ("nch", C.c_int), ("chunk_frames", C.c_size_t),
                ("input", C.c_int), ("fir_gain", C.c_double), ("gain", C.c_double), ("pcm_boxcar", C.c_int),
                ("soft_clip", C.c_int), ("dsd_boxcar", C.c_int), ("dsd_lowpass", C.c_int), ("ir", _dp),
                ("ir_len", C.c_int64), ("ir_stride", C.c_int64), ("P", C.c_int), ("ntf_index", C.c_int),
                ("depth", C.c_int), ("cands", C.c_int), ("lat", C.c_int), ("state_limit", C.c_double),
                ("seg_len", C.c_size_t), ("reposition", C.c_int), ("fma_o8", C.c_int), ("seeding", C.c_int)]


# DAS seeding of the chain: the reference's rule by output rate, or forced
SEED_BY_RATE, SEED_SEG0_FIRST, SEED_ALL_PARALLEL = 0, 1, 2


class Stream:
    """fs_in PCM (float32, nch channels) -> fs_out DSD bits, chunk by chunk.
    ir: None or (nch, taps) float64 room impulse responses at fs_out."""

    def __init__(self, fs_in, fs_out, nch, chunk_frames, ir=None, P=8192, ntf_index=None, K=32, L=64,
                 depth=8, gain=fir.PCM_FIR_GAIN, state_limit=0.0, seg_len=65536, reposition=True,
                 input=IN_PCM, volume=1.0, pcm_boxcar=0, soft_clip=False, dsd_boxcar=0, dsd_lowpass=False,
                 fma_o8=False, seeding=SEED_BY_RATE):
        """gain: the engine's fir_gain; volume: cfg->gain (set_gain ramps it);
        seeding: SEED_BY_RATE (seg0-first below DSD512, as the reference),
        SEED_SEG0_FIRST or SEED_ALL_PARALLEL."""
        self.lib = _lib.lib()
        _sigs(self.lib)
        if ntf_index is None:
            ntf_index = trellis.ntf_auto(fs_out)
        self.nch, self.chunk_frames = nch, chunk_frames
        c = StreamCfg()
        c.fs_in, c.fs_out, c.nch, c.chunk_frames, c.input = fs_in, fs_out, nch, chunk_frames, input
        c.fir_gain, c.gain = gain, volume
        c.pcm_boxcar, c.soft_clip, c.dsd_boxcar, c.dsd_lowpass = pcm_boxcar, int(soft_clip), dsd_boxcar, int(
            dsd_lowpass)
        if ir is not None:
            self._ir = np.ascontiguousarray(np.atleast_2d(ir), dtype=np.float64)
            assert self._ir.shape[0] == nch
            c.ir, c.ir_len, c.ir_stride = self._ir.ctypes.data_as(_dp), self._ir.shape[1], self._ir.shape[1]
        c.P, c.ntf_index, c.depth, c.cands, c.lat = P, ntf_index, depth, K, L
        c.state_limit, c.seg_len, c.reposition, c.fma_o8 = state_limit, seg_len, int(reposition), int(fma_o8)
        c.seeding = seeding
        self.seg0_first = seeding == SEED_SEG0_FIRST or (seeding == SEED_BY_RATE and fs_out < 22579200)
        self.h = self.lib.dsdhip_stream_create_cfg(C.byref(c))
        if not self.h:
            raise _lib.DsdHipError("dsdhip_stream_create_cfg failed")
        self.R = self.lib.dsdhip_stream_ratio(self.h)
        self.words = (chunk_frames * self.R + 63) // 64
        self.bits = torch.zeros((nch, self.words), dtype=torch.int64, device="cuda")

    def set_gain(self, volume):
        _lib.check(self.lib.dsdhip_stream_set_gain(self.h, volume), "dsdhip_stream_set_gain")

    def process_dsd(self, bits, n):
        """bits: cuda int64 (nch, words) packed DSD at fs_in, n samples per channel."""
        cnt = C.c_size_t(0)
        _lib.check(self.lib.dsdhip_stream_process_dsd_dev(self.h, bits.data_ptr(), bits.stride(0), n,
                                                          self.bits.data_ptr(), self.words, C.byref(cnt)),
                   "dsdhip_stream_process_dsd_dev")
        return cnt.value

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

    def last_timeline(self):
        """Host-side timeline of the last chunk, ms from the call's start:
        DAS started, seg0 done, estimate started, last segment published,
        estimate joined, trellis read, run done (0.0 where not passed)."""
        t = (C.c_double * 7)()
        _lib.check(self.lib.dsdhip_stream_last_timeline(self.h, t), "dsdhip_stream_last_timeline")
        return list(t)

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
