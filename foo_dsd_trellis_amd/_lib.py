"""Loader for the in-tree HIP library (libdsdhip.so) and its C signatures.

The product has no CPU fallback: if the library is missing or no GPU is
visible, calls fail loudly.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdsdhip.so")

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_i64p = C.POINTER(C.c_int64)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p


class DsdHipError(RuntimeError):
    pass


class NtfRow(C.Structure):
    _fields_ = [("a", C.c_double * 10), ("g", C.c_double * 10), ("order", C.c_int),
                ("freq", C.c_uint), ("name", C.c_char_p)]


_SIGS = {
    "dsdhip_version": ([], C.c_char_p),
    "dsdhip_ntf_count": ([], C.c_int),
    "dsdhip_ntf_get": ([C.c_int], C.POINTER(NtfRow)),
    "dsdhip_ntf_lookup": ([C.c_int, C.c_uint], C.c_int),
    "dsdhip_ntf_auto": ([C.c_uint], C.c_int),
    "gpu_hip_probe": ([], C.c_bool),
    "dsdhip_trellis_create": ([C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int], _vp),
    "dsdhip_trellis_destroy": ([_vp], None),
    "dsdhip_trellis_set_stream": ([_vp, _vp], C.c_int),
    "dsdhip_trellis_reset": ([_vp, C.c_int], C.c_int),
    "dsdhip_trellis_copy": ([_vp, C.c_int, C.c_int], C.c_int),
    "dsdhip_trellis_seed": ([_vp, C.c_int, _dp], C.c_int),
    "dsdhip_trellis_best_state": ([_vp, C.c_int, _dp], C.c_int),
    "dsdhip_trellis_process": ([_vp, C.c_int, _dp, _fp, C.c_size_t, C.POINTER(C.c_size_t)], C.c_int),
    "dsdhip_trellis_run_jobs": ([_vp, C.c_int, _ip, _ip, C.POINTER(_vp), _i64p, _i64p,
                                 C.POINTER(_vp), _dp, _i64p, _i64p, _ip], C.c_int),
    "dsdhip_trellis_run_jobs_noalloc": ([_vp, C.c_int, _ip, _ip, C.POINTER(_vp), _i64p, _i64p,
                                         C.POINTER(_vp), _dp, _i64p, _i64p, _ip], C.c_int),
    "dsdhip_trellis_force_literal": ([_vp, C.c_int], C.c_int),
    "dsdhip_trellis_fma_o8": ([_vp, C.c_int], C.c_int),
    "dsdhip_das_fma_o8": ([_vp, C.c_int], C.c_int),
    "dsdhip_trellis_wave_only": ([_vp, C.c_int], C.c_int),
    "dsdhip_trellis_drain": ([_vp, C.c_int, _fp, C.c_size_t, C.POINTER(C.c_size_t)], C.c_int),
    "dsdhip_trellis_state_distance": ([_vp, C.c_int, C.c_int, _dp], C.c_int),
    "dsdhip_estimate_state": ([_dp, _dp, C.c_int, _dp, _dp, C.c_size_t, C.c_double, C.c_int, _dp], C.c_int),
    "dsdhip_estimate_boundaries": ([_dp, _dp, C.c_int, _dp, _dp, C.c_size_t, _i64p, C.c_int, C.c_double,
                                    C.c_int, _dp], C.c_int),
    "dsdhip_estimate_boundaries_ex": ([_dp, _dp, C.c_int, _dp, _dp, C.c_size_t, _i64p, C.c_int, C.c_double,
                                       C.c_int, _dp, C.c_double, _i64p], C.c_int),
    "dsdhip_estimate_boundaries_dev": ([_dp, _dp, C.c_int, _dp, _vp, C.c_int64, C.c_int, C.c_size_t, _i64p,
                                        C.c_int, C.c_double, C.c_int, _dp], C.c_int),
    "dsdhip_est_table_index": ([C.c_double, C.c_int, C.c_double], C.c_int),
    "dsdhip_das_estimator_device": ([_vp, C.c_int], C.c_int),
    "dsdhip_das_batch_launch": ([_vp, C.c_int], C.c_int),
    "dsdhip_das_feed_limit": ([_vp, C.c_double], C.c_int),
    "dsdhip_das_feed_waves": ([_vp, C.c_int], C.c_int),
    "dsdhip_das_last_timeline": ([_vp, C.POINTER(C.c_double)], C.c_int),
    "dsdhip_stream_last_timeline": ([_vp, C.POINTER(C.c_double)], C.c_int),
    "dsdhip_das_last_estimate": ([_vp, _dp, _i64p, _i64p, _ip], C.c_int),
    "dsdhip_das_wave_only": ([_vp, C.c_int], C.c_int),
    "gpu_hip_fir_lowpass_setup": ([_vp, _fp, C.c_int], C.c_int),
    "gpu_hip_fir_lowpass_f64": ([_vp, _fp, _dp, C.c_size_t, C.c_double], C.c_int),
    "gpu_hip_fir_lowpass": ([_vp, _fp, _fp, C.c_size_t, C.c_float], C.c_int),
    "gpu_hip_boxcar_smooth": ([_vp, _fp, _fp, C.c_size_t, C.c_int, C.c_float], C.c_int),
    "gpu_hip_trellis_fma_o8": ([_vp, C.c_int], C.c_int),
    "dsdhip_fir_plan": ([C.c_uint, C.c_uint, _ip], C.c_int),
    "dsdhip_halfband_taps": ([_dp, C.c_int], None),
    "dsdhip_fir_create": ([C.c_uint, C.c_uint, C.c_double, C.c_int], _vp),
    "dsdhip_fir_destroy": ([_vp], None),
    "dsdhip_fir_set_stream": ([_vp, _vp], C.c_int),
    "dsdhip_fir_num_stages": ([_vp], C.c_int),
    "dsdhip_fir_reset": ([_vp], C.c_int),
    "dsdhip_fir_set_fused": ([_vp, C.c_int], C.c_int),
    "dsdhip_fir_process_dev": ([_vp, _vp, C.c_int64, C.c_size_t, _vp, C.c_int64], C.c_int),
    "dsdhip_fir_process": ([_vp, _fp, C.c_size_t, _dp], C.c_int),
    "dsdhip_das_create": ([C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_size_t,
                           C.c_int, C.c_int, C.c_int], _vp),
    "dsdhip_das_destroy": ([_vp], None),
    "dsdhip_das_set_stream": ([_vp, _vp], C.c_int),
    "dsdhip_das_force_literal": ([_vp, C.c_int], C.c_int),
    "dsdhip_das_plan": ([_vp, C.c_size_t, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)], C.c_int),
    "dsdhip_das_run_dev": ([_vp, _vp, C.c_int64, C.c_size_t, _vp, C.c_int64, C.POINTER(C.c_size_t)], C.c_int),
    "dsdhip_das_run": ([_vp, _dp, C.c_size_t, _fp, C.POINTER(C.c_size_t)], C.c_int),
    "dsdhip_das_last_info": ([_vp, _ip, _i64p, _ip, C.c_int], C.c_int),
    "dsdhip_trellis_timing": ([_vp, C.c_int], C.c_int),
    "dsdhip_trellis_timing_read": ([_vp, _dp, _i64p, _i64p, C.c_int], C.c_int),
    "dsdhip_das_reset": ([_vp], C.c_int),
    "dsdhip_das_timing": ([_vp, C.c_int], C.c_int),
    "dsdhip_das_timing_read": ([_vp, _dp, _i64p, _i64p, C.c_int], C.c_int),
    "gpu_hip_create": ([], _vp),
    "gpu_hip_destroy": ([_vp], None),
    "gpu_hip_reset_chunk": ([_vp], None),
    "gpu_hip_fir_setup": ([_vp, _fp, C.c_int, C.c_int, C.c_bool], C.c_int),
    "gpu_hip_fir_setup_rates": ([_vp, C.c_uint, C.c_uint, C.c_double], C.c_int),
    "gpu_hip_fir_chain": ([_vp, _fp, _fp, C.c_size_t, C.POINTER(C.c_size_t)], C.c_int),
    "gpu_hip_fir_chain_f64": ([_vp, _fp, _dp, C.c_size_t, C.POINTER(C.c_size_t)], C.c_int),
    "gpu_hip_fir_batch": ([_vp, _fp, _fp, C.c_size_t, C.c_int, C.POINTER(C.c_size_t)], C.c_int),
    "gpu_hip_trellis_setup": ([_vp, C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_double, C.c_int], C.c_int),
    "gpu_hip_trellis": ([_vp, _dp, _fp, C.c_size_t], C.c_int),
    "gpu_hip_trellis_das": ([_vp, _dp, _fp, C.c_size_t, C.c_int], C.c_int),
    "gpu_hip_das_config": ([_vp, C.c_size_t, C.c_int, C.c_int], C.c_int),
    "gpu_hip_das_rate": ([_vp, C.c_uint], C.c_int),
    "gpu_hip_trellis_das_ex": ([_vp, _dp, _fp, C.c_size_t, C.c_int, C.POINTER(C.c_size_t)], C.c_int),
    "gpu_hip_trellis_das_count": ([_vp], C.c_size_t),
    "dsdhip_lf_checkpoints": ([C.c_size_t, C.c_int], C.c_size_t),
    "dsdhip_lf_workspace_bytes": ([C.c_int, C.c_size_t, C.c_int, C.c_int, C.c_int], C.c_size_t),
    "dsdhip_lf_forward": ([_vp, _vp, C.c_int, C.c_double, _vp, C.c_int, C.c_size_t, C.c_int, _vp,
                           _vp, _vp], C.c_int),
    "dsdhip_lf_backward": ([_vp, _vp, C.c_int, C.c_double, _vp, C.c_int, C.c_size_t, C.c_int, _vp,
                            _vp, _vp, _vp, _vp, _vp, C.c_size_t, C.c_int, _vp], C.c_int),
    "dsdhip_lf_backward_ex": ([_vp, _vp, C.c_int, C.c_double, _vp, C.c_int, C.c_size_t, C.c_int, _vp,
                               _vp, _vp, _vp, _vp, _vp, C.c_size_t, C.c_int, C.c_int, _vp], C.c_int),
    "dsdhip_pe_workspace_bytes": ([C.c_int, C.c_size_t, C.c_int], C.c_size_t),
    "dsdhip_pe_forward": ([_vp, C.c_int, _vp, C.c_int, C.c_size_t, _vp, _vp], C.c_int),
    "dsdhip_pe_backward": ([_vp, C.c_int, C.c_size_t, C.c_int, _vp, _vp, _vp, C.c_size_t, _vp],
                           C.c_int),
}

_lib = None


def _torch_runtime_first():
    """PyTorch for ROCm carries its own HIP runtime (torch/lib/libamdhip64.so)
    beside the system runtime this library links, so a process that uses
    both holds two HIP runtimes.  Torch's is brought up first: started
    after the library's runtime had already run work on the device, torch's
    lazy initialisation has reported no GPU."""
    try:
        import torch
    except ImportError:
        return
    try:
        if torch.cuda.is_available():
            torch.cuda.init()
    except Exception:
        pass


def lib():
    """The loaded library (raises DsdHipError if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DsdHipError("libdsdhip.so is not built (run __graft_entry__.build())")
        _torch_runtime_first()
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def check(rc, what):
    if rc != 0:
        raise DsdHipError("%s failed (rc=%d)" % (what, rc))
