"""DSD wire formats on the device (dsdhip_dop_* / dsdhip_bytes_*): DoP and
raw MSB-first bytes to and from packed-bit streams.  Mirrors include/dop.h
(dop_detect, dop_unpack, dop_pack, dop_pack_i24, bits_unpack, bits_pack)."""
import ctypes as C

import numpy as np
import torch

from . import _lib

_vp = C.c_void_p


def _sigs(L):
    if getattr(L, "_dop_sigs", False):
        return
    L.dsdhip_dop_detect.argtypes = [C.POINTER(C.c_float), C.c_size_t, C.c_int]
    L.dsdhip_dop_unpack_dev.argtypes = [_vp, C.c_size_t, C.c_int, _vp, C.c_int64, _vp]
    L.dsdhip_dop_pack_dev.argtypes = [_vp, C.c_int64, C.c_int, C.c_size_t, _vp, _vp]
    L.dsdhip_dop_pack_i24_dev.argtypes = [_vp, C.c_int64, C.c_int, C.c_size_t, C.c_int, _vp, _vp]
    L.dsdhip_bytes_unpack_dev.argtypes = [_vp, C.c_int64, C.c_int, C.c_size_t, _vp, C.c_int64, _vp]
    L.dsdhip_bytes_pack_dev.argtypes = [_vp, C.c_int64, C.c_int, C.c_size_t, _vp, C.c_int64, _vp]
    L.dsdhip_bits_to_f32_dev.argtypes = [_vp, C.c_int64, C.c_int, C.c_size_t, _vp, C.c_int64, _vp]
    L._dop_sigs = True


def _L():
    L = _lib.lib()
    _sigs(L)
    return L


def _stream():
    return torch.cuda.current_stream().cuda_stream


def detect(pcm24, channels=1):
    """pcm24: host float32 frames, interleaved when channels > 1."""
    a = np.ascontiguousarray(pcm24, dtype=np.float32)
    return bool(_L().dsdhip_dop_detect(a.ctypes.data_as(C.POINTER(C.c_float)), a.size // channels, channels))


def words(n_bits):
    return (n_bits + 63) // 64


def unpack(pcm24):
    """cuda float32 (frames, nch) interleaved DoP -> int64 (nch, words) packed bits."""
    frames, nch = pcm24.shape
    src = pcm24.contiguous()
    out = torch.zeros((nch, words(frames * 16)), dtype=torch.int64, device=pcm24.device)
    _lib.check(_L().dsdhip_dop_unpack_dev(src.data_ptr(), frames, nch, out.data_ptr(), out.shape[1], _stream()),
               "dsdhip_dop_unpack_dev")
    return out


def pack(bits, bit_count):
    """packed bits (nch, words) -> cuda float32 (bit_count // 16, nch) DoP frames."""
    nch = bits.shape[0]
    out = torch.empty((bit_count // 16, nch), dtype=torch.float32, device=bits.device)
    _lib.check(_L().dsdhip_dop_pack_dev(bits.data_ptr(), bits.stride(0), nch, bit_count, out.data_ptr(), _stream()),
               "dsdhip_dop_pack_dev")
    return out


def pack_i24(bits, bit_count, phase=0):
    """packed bits -> (uint8 (frames, nch, 3), next marker phase)."""
    nch = bits.shape[0]
    out = torch.empty((bit_count // 16, nch, 3), dtype=torch.uint8, device=bits.device)
    r = _L().dsdhip_dop_pack_i24_dev(bits.data_ptr(), bits.stride(0), nch, bit_count, phase, out.data_ptr(),
                                     _stream())
    if r < 0:
        raise _lib.DsdHipError("dsdhip_dop_pack_i24_dev failed")
    return out, r


def bytes_unpack(src, n_bits):
    """uint8 (nch, nbytes) MSB-first -> packed bits."""
    nch = src.shape[0]
    out = torch.zeros((nch, words(n_bits)), dtype=torch.int64, device=src.device)
    _lib.check(_L().dsdhip_bytes_unpack_dev(src.data_ptr(), src.stride(0), nch, n_bits, out.data_ptr(), out.shape[1],
                                            _stream()), "dsdhip_bytes_unpack_dev")
    return out


def bytes_pack(bits, n_bits):
    nch = bits.shape[0]
    out = torch.empty((nch, (n_bits + 7) // 8), dtype=torch.uint8, device=bits.device)
    _lib.check(_L().dsdhip_bytes_pack_dev(bits.data_ptr(), bits.stride(0), nch, n_bits, out.data_ptr(), out.shape[1],
                                          _stream()), "dsdhip_bytes_pack_dev")
    return out


def to_f32(bits, n):
    """packed bits -> cuda float32 (nch, n) of +-1.0."""
    nch = bits.shape[0]
    out = torch.empty((nch, n), dtype=torch.float32, device=bits.device)
    _lib.check(_L().dsdhip_bits_to_f32_dev(bits.data_ptr(), bits.stride(0), nch, n, out.data_ptr(), n, _stream()),
               "dsdhip_bits_to_f32_dev")
    return out
