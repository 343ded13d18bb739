"""Helpers for the sdm_* surface tests: the library's sdm_* entry points on
raw reference-layout contexts, and the flattened carried state of a
context as the compiled reference reads it (oracle/ref_driver.c
ref_export).  Test infrastructure only."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SO = os.path.join(ROOT, "foo_dsd_trellis_amd", "libdsdhip.so")

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_vp = C.c_void_p

_lib = None
MB_LEVELS_OFFSET = 30292        # offsetof(sdm_context_t, mb_levels)


def ours():
    global _lib
    if _lib is None:
        L = C.CDLL(SO)
        L.sdm_context_init.argtypes = [_vp, _vp, C.c_int, C.c_int, C.c_int]
        L.sdm_process_block.argtypes = [_vp, _dp, _fp, C.c_size_t]
        L.sdm_process_block.restype = C.c_size_t
        L.sdm_drain.argtypes = [_vp, _fp, C.c_size_t]
        L.sdm_drain.restype = C.c_size_t
        L.sdm_context_copy_state.argtypes = [_vp, _vp]
        L.sdm_state_distance.argtypes = [_vp, _vp]
        L.sdm_state_distance.restype = C.c_double
        L.sdm_context_reset.argtypes = [_vp]
        L.sdm_context_free.argtypes = [_vp]
        L.sdm_estimate_state_multibit.argtypes = [_vp, _dp, _dp, C.c_size_t, C.c_double, C.c_int, _dp]
        L.sdm_estimate_state.argtypes = [_vp, _dp, _dp, C.c_size_t, C.c_double, _dp]
        L.ntf_get_by_index.argtypes = [C.c_int]
        L.ntf_get_by_index.restype = _vp
        L.ntf_get_filter.argtypes = [C.c_int, C.c_uint]
        L.ntf_get_filter.restype = _vp
        L.ntf_auto_select.argtypes = [C.c_uint]
        L.ntf_auto_select.restype = _vp
        L.ntf_auto_select_precorr.argtypes = [C.c_uint]
        L.ntf_auto_select_precorr.restype = _vp
        L.dsdhip_sdm_roundtrip.argtypes = [_vp]
        _lib = L
    return _lib


def process(h, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(max(x.size, 1), dtype=np.float32)
    n = ours().sdm_process_block(h, x.ctypes.data_as(_dp), out.ctypes.data_as(_fp), x.size)
    return out[:n]


def drain(h, max_out):
    out = np.zeros(max(max_out, 1), dtype=np.float32)
    n = ours().sdm_drain(h, out.ctypes.data_as(_fp), max_out)
    return out[:n]


def raw(ref, h):
    n = ref.lib.ref_ctx_size()
    buf = (C.c_ubyte * n)()
    ref.lib.ref_ctx_read(C.c_void_p(h), buf)
    return bytes(buf)


def export(ref, h, other=None):
    """Carried state of context h (compiled reference's view)."""
    L = ref.lib
    meta = (C.c_int * 8)()
    st = np.zeros(320)
    cost = np.zeros(32)
    path = np.zeros(32, dtype=np.uint32)
    nxt = np.zeros(32, dtype=np.uint8)
    slot = np.zeros(32, dtype=np.int32)
    hist = np.zeros(32 * 256, dtype=np.uint8)
    npath = np.zeros(128, dtype=np.uint32)
    ncost = np.zeros(128)
    nlink = np.zeros(128, dtype=np.int32)
    L.ref_export(C.c_void_p(h), C.c_void_p(other) if other else None, meta, st.ctypes.data_as(_dp),
                 cost.ctypes.data_as(_dp), path.ctypes.data_as(C.c_void_p), nxt.ctypes.data_as(C.c_void_p),
                 slot.ctypes.data_as(C.c_void_p), hist.ctypes.data_as(C.c_void_p),
                 npath.ctypes.data_as(C.c_void_p), ncost.ctypes.data_as(_dp), nlink.ctypes.data_as(C.c_void_p))
    m = list(meta)
    n, lat = m[0], m[6]
    hb = (lat + 7) // 8
    hbits = np.unpackbits(hist.reshape(32, 256)[:n, :hb], axis=1, bitorder="little")[:, :lat]
    return dict(num_cands=n, pos=m[1], pending=m[2], draining=m[3], idx=m[4], K=m[5], L=lat,
                hist_fnum=m[7], st=st.reshape(32, 10)[:n].copy(), cost=cost[:n].copy(), path=path[:n].copy(),
                next=nxt[:n].copy(), slot=slot[:n].copy(), hist=hbits, npath=npath.copy(),
                ncost=ncost.copy(), nlink=nlink.copy())


def assert_same_carried(a, b, order, what=""):
    """Everything a context carries into its next steps, bit for bit: the
    survivors (state, metric, path, traceback bit, node, history) and every
    node's path, metric and chain link."""
    for k in ("num_cands", "pos", "pending", "draining", "idx"):
        assert a[k] == b[k], (what, k, a[k], b[k])
    assert a["hist_fnum"] == b["hist_fnum"], (what, "hist_fnum")
    assert np.array_equal(a["st"][:, :order].view(np.uint64), b["st"][:, :order].view(np.uint64)), (what, "st")
    assert np.array_equal(a["cost"].view(np.uint64), b["cost"].view(np.uint64)), (what, "cost")
    for k in ("path", "next", "slot", "hist", "npath", "nlink"):
        assert np.array_equal(a[k], b[k]), (what, k)
    assert np.array_equal(a["ncost"].view(np.uint64), b["ncost"].view(np.uint64)), (what, "ncost")


def _pointer_offsets():
    offs = []
    for b in range(2):
        tb = b * 14592
        for j in range(128):
            offs += [tb + j * 112 + 96, tb + j * 112 + 104]      # parent, path_list
        offs += [tb + 14336 + 8 * i for i in range(32)]          # act[]
    offs += [29184 + 8 * i for i in range(128)]                  # path_hash[]
    return offs


_POFFS = _pointer_offsets()


def normalized(ref, h):
    """Raw bytes of context h with every pointer into h itself replaced by
    its offset (so two contexts at different addresses compare equal)."""
    b = bytearray(raw(ref, h))
    for o in _POFFS:
        v = int.from_bytes(b[o:o + 8], "little")
        if h <= v < h + 46752:
            b[o:o + 8] = (0x7ff0000000000000 | (v - h)).to_bytes(8, "little")
    return bytes(b)
