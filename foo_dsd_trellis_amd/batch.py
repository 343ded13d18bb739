"""Batched DAS over many tracks of uneven length (gpu_hip_plan_* in
libdsdhip.so): one persistent trellis launch per chunk over every segment
of every channel and track, the chunk step captured in a hipGraph and
replayed, launch shape and segment length chosen by the autotuner.

    plan = BatchPlan(48000, 24576000, ntf_index, K=32, L=64, nch=6, chunk_frames=24000)
    plan.enqueue([pcm0, pcm1, ...])     # torch float32 (nch, frames) on the GPU
    bits = plan.run()                   # output bit counts per track
    plan.outputs[t]                     # torch int64 (nch, words): packed DSD bits
"""
import ctypes as C

import numpy as np

from . import _lib, fir

_i64p = C.POINTER(C.c_int64)


class Tune(C.Structure):
    _fields_ = [("seg_len", C.c_int), ("waves_per_group", C.c_int), ("waves_per_cu", C.c_int),
                ("use_graph", C.c_int), ("msps_measured", C.c_double)]


def _sigs(L):
    if getattr(L, "_plan_sigs", False):
        return
    L.gpu_hip_plan_create.restype = C.c_void_p
    L.gpu_hip_plan_create.argtypes = [C.c_uint, C.c_uint, C.c_double, C.c_int, C.c_int, C.c_int,
                                      C.c_int, C.c_double, C.c_int, C.c_size_t, C.c_int,
                                      C.POINTER(Tune)]
    L.gpu_hip_plan_destroy.argtypes = [C.c_void_p]
    L.gpu_hip_plan_set_stream.argtypes = [C.c_void_p, C.c_void_p]
    L.gpu_hip_plan_enqueue.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_void_p), _i64p, _i64p,
                                       C.POINTER(C.c_void_p), _i64p, _i64p]
    L.gpu_hip_plan_launch_graph.argtypes = [C.c_void_p, C.c_int, _i64p]
    L.gpu_hip_plan_get_tune.argtypes = [C.c_void_p, C.POINTER(Tune)]
    L.gpu_hip_plan_info.argtypes = [C.c_void_p, _i64p, _i64p, C.POINTER(C.c_int),
                                    C.POINTER(C.c_double)]
    L._plan_sigs = True


class BatchPlan:
    def __init__(self, fs_in, fs_out, ntf_index, K=32, L=64, depth=8, nch=2, chunk_frames=24000,
                 gain=fir.PCM_FIR_GAIN, state_limit=0.0, reposition=True, seg_len=0,
                 waves_per_group=0, waves_per_cu=0, use_graph=True):
        self.lib = _lib.lib()
        _sigs(self.lib)
        t = Tune(seg_len, waves_per_group, waves_per_cu, 1 if use_graph else 0, 0.0)
        self.h = self.lib.gpu_hip_plan_create(fs_in, fs_out, gain, ntf_index, depth, K, L,
                                              state_limit, nch, chunk_frames,
                                              1 if reposition else 0, C.byref(t))
        if not self.h:
            raise _lib.DsdHipError("gpu_hip_plan_create rejected the configuration")
        self.nch = nch
        self.ratio = fs_out // fs_in
        self.outputs = []
        self._keep = []

    def enqueue(self, pcms):
        """pcms: list of torch float32 CUDA tensors shaped (nch, frames)."""
        import torch
        n = len(pcms)
        pp = (C.c_void_p * n)()
        st = (C.c_int64 * n)()
        fr = (C.c_int64 * n)()
        po = (C.c_void_p * n)()
        os_ = (C.c_int64 * n)()
        ow = (C.c_int64 * n)()
        self.outputs = []
        self._keep = list(pcms)
        for i, x in enumerate(pcms):
            if x.dtype != torch.float32 or not x.is_cuda or x.dim() != 2 or x.shape[0] != self.nch:
                raise ValueError("track %d: expected float32 CUDA tensor (nch, frames)" % i)
            if x.stride(1) != 1:
                raise ValueError("track %d: frames must be contiguous" % i)
            words = (x.shape[1] * self.ratio + 63) // 64 + 2
            out = torch.zeros((self.nch, words), dtype=torch.int64, device=x.device)
            self.outputs.append(out)
            pp[i] = x.data_ptr()
            st[i] = x.stride(0)
            fr[i] = x.shape[1]
            po[i] = out.data_ptr()
            os_[i] = words
            ow[i] = words
        _lib.check(self.lib.gpu_hip_plan_enqueue(self.h, n, pp, st, fr, po, os_, ow),
                   "gpu_hip_plan_enqueue")
        self.ntracks = n

    def run(self, max_chunks=-1):
        """Runs chunk steps; returns the output bit count of every track."""
        bits = (C.c_int64 * self.ntracks)()
        rc = self.lib.gpu_hip_plan_launch_graph(self.h, max_chunks, bits)
        if rc < 0:
            raise _lib.DsdHipError("gpu_hip_plan_launch_graph failed (the batched launch is "
                                   "disabled in this build)")
        self.last_chunks = rc
        return np.array(list(bits), dtype=np.int64)

    def tune(self):
        t = Tune()
        self.lib.gpu_hip_plan_get_tune(self.h, C.byref(t))
        return {"seg_len": t.seg_len, "waves_per_group": t.waves_per_group,
                "waves_per_cu": t.waves_per_cu, "use_graph": bool(t.use_graph),
                "trial_msps": t.msps_measured}

    def info(self):
        a, b, s, ms = C.c_int64(), C.c_int64(), C.c_int(), C.c_double()
        self.lib.gpu_hip_plan_info(self.h, C.byref(a), C.byref(b), C.byref(s), C.byref(ms))
        return {"chunks_total": a.value, "chunks_done": b.value, "segments_per_chunk": s.value,
                "elapsed_ms": ms.value}

    def __del__(self):
        try:
            if self.h:
                self.lib.gpu_hip_plan_destroy(self.h)
                self.h = None
        except Exception:
            pass
