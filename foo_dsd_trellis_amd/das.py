"""Host-side mirror of the reference's DAS parallel trellis
(src/dsp_plugin.c:1314-2761) on the HIP backend."""
import ctypes as C

import numpy as np

from ._lib import DsdHipError, check, lib

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_i64p = C.POINTER(C.c_int64)


def _coeffs(ntf_index):
    row = lib().dsdhip_ntf_get(ntf_index)
    if not row:
        raise DsdHipError("no NTF filter %d" % ntf_index)
    r = row.contents
    return np.array(r.a[:], np.float64), np.array(r.g[:], np.float64), r.order


def estimate_boundaries(ntf_index, init, x, bnd, state_limit=0.0, levels=16, index_bias=0.0):
    """The estimation block of src/threadpool.c:345-366 on the host: the
    16-level greedy recurrence (src/trellis.c:1118-1187) from init over x,
    the state at each boundary offset in bnd -> ((nb, 10) states, blocks
    recomputed in the reference form).  index_bias is a validation switch."""
    a, g, order = _coeffs(ntf_index)
    init = np.zeros(10) if init is None else np.asarray(init, np.float64)
    init = np.concatenate([init, np.zeros(10 - init.size)]) if init.size < 10 else init[:10].copy()
    x = np.ascontiguousarray(x, dtype=np.float64)
    bnd = np.ascontiguousarray(bnd, dtype=np.int64)
    out = np.zeros((max(1, bnd.size), 10))
    redo = C.c_int64(0)
    check(lib().dsdhip_estimate_boundaries_ex(a.ctypes.data_as(_dp), g.ctypes.data_as(_dp), order,
                                              init.ctypes.data_as(_dp), x.ctypes.data_as(_dp), x.size,
                                              bnd.ctypes.data_as(_i64p), bnd.size, state_limit, levels,
                                              out.ctypes.data_as(_dp), index_bias, C.byref(redo)),
          "dsdhip_estimate_boundaries")
    return out[:bnd.size], redo.value


def estimate_boundaries_dev(ntf_index, init, d_x, stride, nch, count, bnd, state_limit=0.0, levels=16):
    """The same on one device lane per channel: d_x a device pointer to nch
    rows of `stride` doubles; init (nch, 10) -> (nch, nb, 10)."""
    a, g, order = _coeffs(ntf_index)
    init = np.ascontiguousarray(np.broadcast_to(np.asarray(init, np.float64), (nch, 10)))
    bnd = np.ascontiguousarray(bnd, dtype=np.int64)
    out = np.zeros((nch, bnd.size, 10))
    check(lib().dsdhip_estimate_boundaries_dev(a.ctypes.data_as(_dp), g.ctypes.data_as(_dp), order,
                                               init.ctypes.data_as(_dp), d_x, stride, nch, count,
                                               bnd.ctypes.data_as(_i64p), bnd.size, state_limit, levels,
                                               out.ctypes.data_as(_dp)), "dsdhip_estimate_boundaries_dev")
    return out


def table_index(a0, levels, v):
    """Quantizer index from the host estimator's exact interval table (-1:
    table not used for these parameters)."""
    return lib().dsdhip_est_table_index(a0, levels, v)


class Das:
    """nch persistent trellis contexts + DAS plan (segment length or count,
    seed mode seg0-first / all-parallel, boundary reposition)."""

    def __init__(self, ntf_index, depth, cands, lat, nch, seg_len=65536, nseg=0,
                 seg0_first=False, reposition=True, state_limit=0.0):
        self.h = lib().dsdhip_das_create(ntf_index, depth, cands, lat, state_limit, nch, seg_len,
                                         nseg, 1 if seg0_first else 0, 1 if reposition else 0)
        if not self.h:
            raise DsdHipError("dsdhip_das_create rejected the parameters")
        self.nch = nch

    def close(self):
        if self.h:
            lib().dsdhip_das_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, s):
        check(lib().dsdhip_das_set_stream(self.h, s), "das set_stream")

    def force_literal(self, on=True):
        check(lib().dsdhip_das_force_literal(self.h, 1 if on else 0), "force_literal")

    def fma_o8(self, on=True):
        """Order 8: the reference AVX2 build's 4-candidate FMA filter."""
        check(lib().dsdhip_das_fma_o8(self.h, 1 if on else 0), "fma_o8")

    def wave_only(self, on=True):
        """K <= 2: run each job on a whole wave instead of one lane."""
        check(lib().dsdhip_das_wave_only(self.h, 1 if on else 0), "wave_only")

    def nseg(self, F):
        return lib().dsdhip_das_plan(self.h, F, None, None)

    def run(self, x):
        """x: (nch, F) float64 host -> ((nch, F) float32 +-1, assembled length)."""
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(self.nch, -1)
        F = x.shape[1]
        out = np.zeros((self.nch, F), np.float32)
        cnt = C.c_size_t(0)
        check(lib().dsdhip_das_run(self.h, x.ctypes.data_as(_dp), F, out.ctypes.data_as(_fp),
                                   C.byref(cnt)), "das run")
        return out, cnt.value

    def run_dev(self, d_in, in_stride, F, d_out, out_words):
        cnt = C.c_size_t(0)
        check(lib().dsdhip_das_run_dev(self.h, d_in, in_stride, F, d_out, out_words, C.byref(cnt)),
              "das run_dev")
        return cnt.value

    def reset(self):
        check(lib().dsdhip_das_reset(self.h), "das reset")

    def timing(self, on=True):
        check(lib().dsdhip_das_timing(self.h, 1 if on else 0), "das timing")

    def timing_read(self, clear=True):
        ms = C.c_double(0)
        nl = C.c_int64(0)
        ns = C.c_int64(0)
        check(lib().dsdhip_das_timing_read(self.h, C.byref(ms), C.byref(nl), C.byref(ns),
                                           1 if clear else 0), "timing_read")
        return ms.value, nl.value, ns.value

    def estimator_device(self, on=True):
        """seg0-first: run the seeding estimate on one device lane per channel
        instead of host threads (timing only; same bits)."""
        check(lib().dsdhip_das_estimator_device(self.h, 1 if on else 0), "estimator_device")

    def batch_launch(self, on=True):
        """seg0-first: launch the estimated segments in batches instead of
        the host-fed persistent launch (same bits)."""
        check(lib().dsdhip_das_batch_launch(self.h, 1 if on else 0), "batch_launch")

    def feed_limit(self, seconds):
        """Longest wait of one wave of the host-fed launch for its segment;
        a wave that waits longer leaves and the run fails."""
        check(lib().dsdhip_das_feed_limit(self.h, float(seconds)), "feed_limit")

    def feed_waves(self, waves_per_cu):
        """Resident waves per CU of the host-fed launch (0: the kernel's own
        limit)."""
        check(lib().dsdhip_das_feed_waves(self.h, int(waves_per_cu)), "feed_waves")

    def last_timeline(self):
        """Host-side timeline of the last seg0-first run (ms from its start):
        seg0 done, estimate started, last segment published, estimate joined,
        trellis results read, run done."""
        t = (C.c_double * 6)()
        check(lib().dsdhip_das_last_timeline(self.h, t), "last_timeline")
        return [round(v, 3) for v in t]

    def last_estimate(self):
        """(longest estimator chain ms, samples estimated, redo blocks, batches)
        of the last seg0-first run."""
        ms = C.c_double(0)
        ns = C.c_int64(0)
        rd = C.c_int64(0)
        nb = C.c_int(0)
        check(lib().dsdhip_das_last_estimate(self.h, C.byref(ms), C.byref(ns), C.byref(rd), C.byref(nb)),
              "last_estimate")
        return ms.value, ns.value, rd.value, nb.value

    def last_info(self, max_pos=0):
        ns = C.c_int(0)
        sl = C.c_int64(0)
        pos = (C.c_int * max(1, max_pos))()
        check(lib().dsdhip_das_last_info(self.h, C.byref(ns), C.byref(sl), pos, max_pos), "last_info")
        return ns.value, sl.value, list(pos[:max_pos])
