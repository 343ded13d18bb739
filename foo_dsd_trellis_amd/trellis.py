"""Host-side mirror of the reference's trellis API (include/trellis.h:61-118)
on the HIP backend.

`SdmPool` holds `n` device-resident trellis contexts sharing one parameter
set (NTF, depth B, candidates K, latency L, state limit), with the same
operations as the reference's sdm_* functions.  Input/output device buffers
are torch tensors (PyTorch is used only for device memory and streams).
"""
import ctypes as C

import numpy as np

from ._lib import DsdHipError, check, lib

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)


def ntf_lookup(ntf_id, rate):
    return lib().dsdhip_ntf_lookup(ntf_id, rate)


def ntf_auto(rate):
    """Index of the reference's auto-selected NTF for a DSD rate (src/ntf.c:839-855)."""
    return lib().dsdhip_ntf_auto(rate)


def ntf_row(index):
    r = lib().dsdhip_ntf_get(index)
    if not r:
        raise DsdHipError("no NTF %d" % index)
    r = r.contents
    return dict(a=np.array(r.a[:]), g=np.array(r.g[:]), order=r.order, freq=r.freq,
                name=r.name.decode())


class SdmPool:
    """n trellis contexts on the device (sdm_context_init for each)."""

    def __init__(self, ntf_index, depth, cands, lat, state_limit=0.0, n=1):
        self.h = lib().dsdhip_trellis_create(ntf_index, depth, cands, lat, state_limit, n)
        if not self.h:
            raise DsdHipError("dsdhip_trellis_create rejected the parameters or found no GPU")
        self.n = n
        self.lat = lat

    def close(self):
        if self.h:
            lib().dsdhip_trellis_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream_ptr):
        check(lib().dsdhip_trellis_set_stream(self.h, stream_ptr), "set_stream")

    def force_literal(self, on=True):
        check(lib().dsdhip_trellis_force_literal(self.h, 1 if on else 0), "force_literal")

    def fma_o8(self, on=True):
        """Order 8: the reference AVX2 build's 4-candidate FMA filter."""
        check(lib().dsdhip_trellis_fma_o8(self.h, 1 if on else 0), "fma_o8")

    def wave_only(self, on=True):
        """K <= 2: run each job on a whole wave instead of one lane."""
        check(lib().dsdhip_trellis_wave_only(self.h, 1 if on else 0), "wave_only")

    def reset(self, i):
        check(lib().dsdhip_trellis_reset(self.h, i), "reset")

    def copy(self, dst, src):
        check(lib().dsdhip_trellis_copy(self.h, dst, src), "copy")

    def seed(self, i, state):
        st = np.zeros(10)
        st[:len(state)] = state
        check(lib().dsdhip_trellis_seed(self.h, i, st.ctypes.data_as(_dp)), "seed")

    def best_state(self, i):
        out = np.zeros(10)
        check(lib().dsdhip_trellis_best_state(self.h, i, out.ctypes.data_as(_dp)), "best_state")
        return out

    def process(self, i, x):
        """sdm_process_block with host arrays: returns float32 +-1.0 outputs."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros(max(1, x.size), dtype=np.float32)
        cnt = C.c_size_t(0)
        check(lib().dsdhip_trellis_process(self.h, i, x.ctypes.data_as(_dp), out.ctypes.data_as(_fp),
                                           x.size, C.byref(cnt)), "process")
        return out[:cnt.value]

    def drain(self, i, max_out=1 << 20):
        """sdm_drain: up to max_out float32 +-1.0 outputs from zero input."""
        out = np.zeros(max(1, max_out), dtype=np.float32)
        cnt = C.c_size_t(0)
        check(lib().dsdhip_trellis_drain(self.h, i, out.ctypes.data_as(_fp), max_out, C.byref(cnt)), "drain")
        return out[:cnt.value]

    def run_jobs(self, ctx, d_in, n_in, d_out, keep_from=None, copy_src=None, seed=None, noalloc=False):
        """Batched launch on device buffers.

        ctx, n_in, keep_from, copy_src: int sequences (one per job);
        d_in / d_out: device addresses (ints) of float64 input and uint64
        packed-bit output; seed: (njobs, 10) float64, NaN row = no seed.
        noalloc: dsdhip_trellis_run_jobs_noalloc (job table in a static
        device array, nothing allocated; at most DSDHIP_NOALLOC_MAX_JOBS).
        Returns (kept, slow_steps) arrays.
        """
        nj = len(ctx)
        ci = (C.c_int * nj)(*ctx)
        cs = (C.c_int * nj)(*(copy_src if copy_src is not None else [-1] * nj))
        pin = (C.c_void_p * nj)(*d_in)
        pout = (C.c_void_p * nj)(*d_out)
        ni = (C.c_int64 * nj)(*n_in)
        kf = (C.c_int64 * nj)(*(keep_from if keep_from is not None else [0] * nj))
        kept = (C.c_int64 * nj)()
        slow = (C.c_int64 * nj)()
        err = C.c_int(0)
        sd = None
        if seed is not None:
            sd = np.ascontiguousarray(seed, dtype=np.float64).reshape(nj, 10)
        fn = lib().dsdhip_trellis_run_jobs_noalloc if noalloc else lib().dsdhip_trellis_run_jobs
        rc = fn(
            self.h, nj, ci, cs, pin, ni, kf, pout,
            sd.ctypes.data_as(_dp) if sd is not None else None, kept, slow, C.byref(err))
        if rc != 0:
            raise DsdHipError("trellis jobs failed (err=%d)" % err.value)
        return np.array(kept[:]), np.array(slow[:])


def unpack_bits(words, count):
    """Packed little-endian bits (uint64 words) -> float32 +-1.0 array."""
    b = np.unpackbits(np.asarray(words, dtype=np.uint64).view(np.uint8), bitorder="little")[:count]
    return np.where(b > 0, 1.0, -1.0).astype(np.float32)
