"""Host-side mirror of the reference's FIR chain (include/fir.h:22-79) on
the HIP backend: `FirChain(fs_in, fs_out, gain, nch)` ~ fir_chain_init_ex +
fir_chain_process_d for nch channels, delay lines persisting between calls.
"""
import ctypes as C

import numpy as np

from ._lib import DsdHipError, check, lib

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)

# DSD rates (include/dsd_types.h:21-30)
DSD64, DSD128, DSD256, DSD512 = 2822400, 5644800, 11289600, 22579200
DSD48_64, DSD48_128, DSD48_256, DSD48_512 = 3072000, 6144000, 12288000, 24576000

# Gain the engine applies after the FIR for PCM rate conversion paths absent
# from its path table: fir_gain_db_to_linear(-3) * 0.5f as a float, times the
# volume gain 1.0 (src/engine.c:206-210, include/dsd_types.h:453-459).
PCM_FIR_GAIN = float(np.float32(np.float32(0.708) * np.float32(0.5)))


def plan(fs_in, fs_out):
    nt = (C.c_int * 9)()
    st = lib().dsdhip_fir_plan(fs_in, fs_out, nt)
    if st < 1:
        raise DsdHipError("unsupported rate pair %d -> %d" % (fs_in, fs_out))
    return list(nt[:st])


def halfband_taps(ntaps=63):
    h = np.zeros(ntaps)
    lib().dsdhip_halfband_taps(h.ctypes.data_as(_dp), ntaps)
    return h


class FirChain:
    def __init__(self, fs_in, fs_out, gain=1.0, nch=1):
        self.h = lib().dsdhip_fir_create(fs_in, fs_out, gain, nch)
        if not self.h:
            raise DsdHipError("dsdhip_fir_create failed")
        self.nch = nch
        self.stages = lib().dsdhip_fir_num_stages(self.h)
        self.ratio = 1 << self.stages

    def close(self):
        if self.h:
            lib().dsdhip_fir_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, s):
        check(lib().dsdhip_fir_set_stream(self.h, s), "fir set_stream")

    def set_fused(self, on=True):
        """Adjacent 63-tap stages in one pass (default) or one pass per stage;
        identical outputs."""
        check(lib().dsdhip_fir_set_fused(self.h, 1 if on else 0), "fir set_fused")

    def reset(self):
        check(lib().dsdhip_fir_reset(self.h), "fir reset")

    def process(self, x):
        """x: (nch, n) float32 host array -> (nch, n*ratio) float64."""
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(self.nch, -1)
        n = x.shape[1]
        out = np.zeros((self.nch, n * self.ratio))
        check(lib().dsdhip_fir_process(self.h, x.ctypes.data_as(_fp), n, out.ctypes.data_as(_dp)),
              "fir process")
        return out

    def process_dev(self, d_in, in_stride, n, d_out, out_stride):
        check(lib().dsdhip_fir_process_dev(self.h, d_in, in_stride, n, d_out, out_stride),
              "fir process_dev")
