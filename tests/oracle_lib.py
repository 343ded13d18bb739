"""ctypes access to the CPU oracle (oracle/liboracle.so) and, where it was
built, the compiled reference trellis core (oracle/_ref/libdsdref.so).

Test infrastructure only: the product package never imports this module.
Both libraries expose the same flat API (prefix `ora_` / `ref_`), so a test
can run the oracle and the reference on identical inputs.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libdsdref.so")
REF_FMA_SO = os.path.join(ROOT, "oracle", "_ref", "libdsdref_fma.so")

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)


def _ptr(a, t):
    return a.ctypes.data_as(t)


class TrellisLib:
    """Wrapper over one of the two libraries (prefix 'ora' or 'ref')."""

    def __init__(self, path, prefix):
        self.lib = C.CDLL(path)
        self.p = prefix
        L = self.lib
        f = lambda n: getattr(L, "%s_%s" % (prefix, n))
        f("ntf_get").argtypes = [C.c_int, _dp, _dp, C.POINTER(C.c_int), C.POINTER(C.c_uint), C.c_char_p, C.c_int]
        f("ntf_lookup").argtypes = [C.c_int, C.c_uint]
        f("ntf_auto").argtypes = [C.c_uint]
        f("ctx_new").argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double]
        f("ctx_new").restype = C.c_void_p
        f("ctx_free").argtypes = [C.c_void_p]
        f("ctx_reset").argtypes = [C.c_void_p]
        f("process").argtypes = [C.c_void_p, _dp, _fp, C.c_size_t]
        f("process").restype = C.c_size_t
        f("drain").argtypes = [C.c_void_p, _fp, C.c_size_t]
        f("drain").restype = C.c_size_t
        f("copy_state").argtypes = [C.c_void_p, C.c_void_p]
        f("best_state").argtypes = [C.c_void_p, _dp]
        f("seed").argtypes = [C.c_void_p, _dp]
        f("counters").argtypes = [C.c_void_p, C.POINTER(C.c_ulonglong)]
        f("estimate").argtypes = [C.c_int, _dp, _dp, C.c_size_t, C.c_double, C.c_int, _dp]
        self._f = f

    def ntf_count(self):
        return self._f("ntf_count")()

    def ntf_get(self, i):
        a = np.zeros(10)
        g = np.zeros(10)
        order = C.c_int()
        freq = C.c_uint()
        name = C.create_string_buffer(32)
        rc = self._f("ntf_get")(i, _ptr(a, _dp), _ptr(g, _dp), C.byref(order), C.byref(freq), name, 32)
        if rc:
            return None
        return dict(a=a, g=g, order=order.value, freq=freq.value, name=name.value.decode())

    def ntf_lookup(self, ntf_id, rate):
        return self._f("ntf_lookup")(ntf_id, rate)

    def ntf_auto(self, rate):
        return self._f("ntf_auto")(rate)

    def ctx(self, filter_index, depth, cands, lat, state_limit=0.0):
        h = self._f("ctx_new")(filter_index, depth, cands, lat, state_limit)
        if not h:
            raise ValueError("context init rejected")
        return Ctx(self, h)

    def estimate(self, filter_index, init, x, state_limit=0.0, levels=16):
        init = np.ascontiguousarray(init, dtype=np.float64)
        if init.size < 10:
            init = np.concatenate([init, np.zeros(10 - init.size)])
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros(10)
        self._f("estimate")(filter_index, _ptr(init, _dp), _ptr(x, _dp), x.size, state_limit, levels, _ptr(out, _dp))
        return out


class Ctx:
    def __init__(self, lib, h):
        self.lib = lib
        self.h = h

    def __del__(self):
        try:
            self.lib._f("ctx_free")(self.h)
        except Exception:
            pass

    def process(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros(max(x.size, 1), dtype=np.float32)
        n = self.lib._f("process")(self.h, _ptr(x, _dp), _ptr(out, _fp), x.size)
        return out[:n]

    def drain(self, max_out=1 << 20):
        out = np.zeros(max_out, dtype=np.float32)
        n = self.lib._f("drain")(self.h, _ptr(out, _fp), max_out)
        return out[:n]

    def reset(self):
        self.lib._f("ctx_reset")(self.h)

    def copy_from(self, other):
        self.lib._f("copy_state")(self.h, other.h)

    def best_state(self):
        out = np.zeros(10)
        self.lib._f("best_state")(self.h, _ptr(out, _dp))
        return out

    def seed(self, st):
        st = np.ascontiguousarray(st, dtype=np.float64)
        if st.size < 10:
            st = np.concatenate([st, np.zeros(10 - st.size)])
        self.lib._f("seed")(self.h, _ptr(st, _dp))

    def counters(self):
        out = (C.c_ulonglong * 8)()
        self.lib._f("counters")(self.h, out)
        keys = ["num_cands", "pos", "pending", "idx", "conv_fail", "cands_collapse",
                "next_filter_drops", "total_children"]
        return dict(zip(keys, list(out)))


_cache = {}


def oracle():
    if "ora" not in _cache:
        _cache["ora"] = TrellisLib(ORACLE_SO, "ora")
    return _cache["ora"]


def reference():
    """The compiled reference, or None where its sources were not present."""
    if "ref" not in _cache:
        _cache["ref"] = TrellisLib(REF_SO, "ref") if os.path.exists(REF_SO) else None
    return _cache["ref"]


class Dop:
    """src/dop.c's API over either library: the oracle (or_*) or the
    compiled reference (dop_* / bits_*), float +-1.0 streams."""

    def __init__(self, lib_, prefix):
        L = lib_.lib
        n = (lambda s: "or_" + s) if prefix == "ora" else (lambda s: s)
        self.f = {k: getattr(L, n(k)) for k in ("dop_unpack", "dop_pack", "dop_pack_i24", "bits_unpack",
                                                  "bits_pack")}
        self.det1 = getattr(L, "or_dop_detect" if prefix == "ora" else "dop_detect_interleaved")
        self.f["dop_pack_i24"].restype = C.c_int
        self.det1.restype = C.c_int if prefix == "ora" else C.c_bool

    def detect(self, pcm, channels=1):
        a = np.ascontiguousarray(pcm, dtype=np.float32)
        return bool(self.det1(_ptr(a, _fp), C.c_size_t(a.size // channels), C.c_int(channels)))

    def unpack(self, pcm):
        a = np.ascontiguousarray(pcm, dtype=np.float32)
        out = np.zeros(a.size * 16, np.float32)
        self.f["dop_unpack"](_ptr(a, _fp), _ptr(out, _fp), C.c_size_t(a.size))
        return out

    def pack(self, bits):
        b = np.ascontiguousarray(bits, dtype=np.float32)
        out = np.zeros(b.size // 16, np.float32)
        self.f["dop_pack"](_ptr(b, _fp), _ptr(out, _fp), C.c_size_t(b.size))
        return out

    def pack_i24(self, bits, phase=0):
        b = np.ascontiguousarray(bits, dtype=np.float32)
        out = np.zeros((b.size // 16) * 3, np.uint8)
        r = self.f["dop_pack_i24"](_ptr(b, _fp), out.ctypes.data_as(C.c_void_p), C.c_size_t(b.size), C.c_int(phase))
        return out, int(r)

    def bits_unpack(self, src, n_bits):
        s = np.ascontiguousarray(src, dtype=np.uint8)
        out = np.zeros(n_bits, np.float32)
        self.f["bits_unpack"](s.ctypes.data_as(C.c_void_p), _ptr(out, _fp), C.c_size_t(n_bits))
        return out

    def bits_pack(self, src):
        a = np.ascontiguousarray(src, dtype=np.float32)
        out = np.zeros((a.size + 7) // 8, np.uint8)
        self.f["bits_pack"](_ptr(a, _fp), out.ctypes.data_as(C.c_void_p), C.c_size_t(a.size))
        return out


class OrCond(C.Structure):
    _fields_ = [("taps", C.c_int), ("ring", C.c_double * 64), ("pos", C.c_int), ("sum", C.c_double),
                ("lp_n", C.c_int), ("lp", C.c_double * 127), ("dly", C.c_double * 127)]


class Cond:
    """The engine's pre-SDM conditioning for one channel (or_cond.c)."""

    def __init__(self, lib_, taps, lowpass_rate=0):
        self.L = lib_.lib
        self.c = OrCond()
        self.L.or_cond_init(C.byref(self.c), C.c_int(taps), C.c_uint(lowpass_rate))

    def _run(self, fn, x, *g):
        a = np.ascontiguousarray(x, dtype=np.float32)
        out = np.zeros(a.size)
        args = [C.byref(self.c), _ptr(a, _fp), C.c_size_t(a.size)] + [C.c_double(v) for v in g] + [_ptr(out, _dp)]
        getattr(self.L, fn)(*args)
        return out

    def dsd_boxcar(self, x, gs, ge):
        return self._run("or_cond_dsd_boxcar", x, gs, ge)

    def dsd_lowpass(self, x, gs, ge):
        return self._run("or_cond_dsd_lowpass", x, gs, ge)

    def pcm_boxcar(self, x):
        return self._run("or_cond_pcm_boxcar", x)


def lowpass_taps(lib_, rate):
    h = np.zeros(127)
    lib_.lib.or_lowpass_taps(C.c_uint(rate), _ptr(h, _dp))
    return h


def gain_ramp(lib_, x, gs, ge):
    y = np.ascontiguousarray(x, dtype=np.float64).copy()
    lib_.lib.or_gain_ramp(_ptr(y, _dp), C.c_size_t(y.size), C.c_double(gs), C.c_double(ge))
    return y


def soft_clip(lib_, x):
    y = np.ascontiguousarray(x, dtype=np.float64).copy()
    lib_.lib.or_soft_clip(_ptr(y, _dp), C.c_size_t(y.size))
    return y


class OrPc(C.Structure):
    _fields_ = [("a", C.c_float * 8), ("g", C.c_float * 8), ("order", C.c_int), ("state_limit", C.c_float),
                ("table", C.c_float * 2048), ("state", C.c_float * 10), ("history", C.c_uint), ("phase", C.c_int),
                ("prev_y", C.c_float)]


class PreCorr:
    """precorr_context_t over the oracle ("ora") or the compiled reference ("ref")."""

    def __init__(self, lib_, prefix, filter_index, state_limit=0.0):
        self.L = lib_.lib
        self.p = prefix
        if prefix == "ora":
            self.c = OrPc()
            self.L.or_pc_init(C.byref(self.c), C.c_int(filter_index), C.c_float(state_limit))
        else:
            self.L.ref_pc_new.restype = C.c_void_p
            self.L.ref_pc_new.argtypes = [C.c_int, C.c_float]
            for f in ("ref_pc_free", "ref_pc_reset"):
                getattr(self.L, f).argtypes = [C.c_void_p]
            self.L.ref_pc_process.argtypes = [C.c_void_p, _dp, _fp, C.c_size_t]
            self.L.ref_pc_table.argtypes = [C.c_void_p, _fp]
            self.L.ref_pc_state.argtypes = [C.c_void_p, _fp, C.POINTER(C.c_int), C.POINTER(C.c_int), _fp]
            self.h = self.L.ref_pc_new(filter_index, state_limit)

    def process(self, x):
        a = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros(a.size, np.float32)
        if self.p == "ora":
            self.L.or_pc_process(C.byref(self.c), _ptr(a, _dp), _ptr(out, _fp), C.c_size_t(a.size))
        else:
            self.L.ref_pc_process(self.h, _ptr(a, _dp), _ptr(out, _fp), a.size)
        return out

    def table(self):
        if self.p == "ora":
            return np.ctypeslib.as_array(self.c.table).copy()
        t = np.zeros(2048, np.float32)
        self.L.ref_pc_table(self.h, _ptr(t, _fp))
        return t

    def state(self):
        """(state[8], history, phase, prev_y)"""
        if self.p == "ora":
            return (np.ctypeslib.as_array(self.c.state)[:8].copy(), int(self.c.history), int(self.c.phase),
                    float(self.c.prev_y))
        st = np.zeros(8, np.float32)
        h, ph, py = C.c_int(0), C.c_int(0), (C.c_float * 1)()
        self.L.ref_pc_state(self.h, _ptr(st, _fp), C.byref(h), C.byref(ph), py)
        return st, int(h.value), int(ph.value), float(py[0])

    def __del__(self):
        try:
            if self.p == "ref" and self.h:
                self.L.ref_pc_free(self.h)
        except Exception:
            pass


def reference_fma():
    """The reference compiled with its shipped AVX2/FMA path selected, or None."""
    if "ref_fma" not in _cache:
        _cache["ref_fma"] = TrellisLib(REF_FMA_SO, "ref") if os.path.exists(REF_FMA_SO) else None
    return _cache["ref_fma"]


class fma_o8:
    """Context manager: oracle contexts created inside use the order-8 FMA
    filter of the reference's AVX2 build."""

    def __init__(self, lib_):
        self.L = lib_.lib
        self.L.ora_set_fma_o8.argtypes = [C.c_int]

    def __enter__(self):
        self.L.ora_set_fma_o8(1)
        return self

    def __exit__(self, *a):
        self.L.ora_set_fma_o8(0)


def pack_bits(pm1):
    """+-1.0 float stream -> packed little-endian bits (1 for +1)."""
    b = (np.asarray(pm1) > 0).astype(np.uint8)
    return np.packbits(b, bitorder="little")


def das(lib_, filter_index, depth, cands, lat, x, seg_len=0, nseg=2, seg0_first=1, reposition=1,
        max_seg=20000):
    """Oracle DAS over x (nch, F) float64 -> (out (nch, F) float32, n, stitch, plan)."""
    L = lib_.lib
    L.ora_das.restype = C.c_long
    L.ora_das.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _dp, C.c_int, C.c_size_t, C.c_size_t,
                          C.c_int, C.c_int, C.c_int, _fp, C.POINTER(C.c_int), C.POINTER(C.c_long),
                          C.c_int]
    x = np.ascontiguousarray(x, dtype=np.float64)
    nch, F = x.shape
    out = np.zeros((nch, F), np.float32)
    sp = np.zeros(max_seg, np.int32)
    plan = np.zeros(2 * max_seg, np.int64)
    n = L.ora_das(filter_index, depth, cands, lat, _ptr(x, _dp), nch, F, seg_len, nseg, seg0_first,
                  reposition, _ptr(out, _fp), sp.ctypes.data_as(C.POINTER(C.c_int)),
                  plan.ctypes.data_as(C.POINTER(C.c_long)), max_seg)
    if n < 0:
        raise ValueError("oracle DAS rejected the plan")
    L.ora_das_nseg.argtypes = [C.c_size_t, C.c_size_t, C.c_int, C.c_int]
    S = L.ora_das_nseg(F, seg_len, nseg, lat)
    return out, int(n), sp[:S], plan[:2 * S].reshape(S, 2)


def fir_chain(lib_, fs_in, fs_out, x, gain=1.0, long_double=False):
    """Oracle FIR chain of one channel x (float32) -> float64."""
    L = lib_.lib
    fn = L.ora_fir_chain_ld if long_double else L.ora_fir_chain
    fn.restype = C.c_size_t
    fn.argtypes = [C.c_uint, C.c_uint, _fp, C.c_size_t, C.c_double, _dp]
    x = np.ascontiguousarray(x, dtype=np.float32)
    r = fs_out // fs_in
    out = np.zeros(x.size * r)
    n = fn(fs_in, fs_out, _ptr(x, _fp), x.size, gain, _ptr(out, _dp))
    return out[:n]


def fir_chain_d(lib_, fs_in, fs_out, x, gain=1.0):
    """Oracle FIR chain of one channel with float64 input."""
    L = lib_.lib
    L.or_fir_chain_d.restype = C.c_size_t
    L.or_fir_chain_d.argtypes = [C.c_uint, C.c_uint, _dp, C.c_size_t, C.c_double, _dp]
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(x.size * (fs_out // fs_in))
    n = L.or_fir_chain_d(fs_in, fs_out, _ptr(x, _dp), x.size, gain, _ptr(out, _dp))
    return out[:n]


def halfband(lib_, ntaps):
    h = np.zeros(ntaps)
    lib_.lib.ora_halfband.argtypes = [_dp, C.c_int]
    lib_.lib.ora_halfband(_ptr(h, _dp), ntaps)
    return h


def stitch(lib_, prev, cur, lat):
    L = lib_.lib
    L.ora_stitch.argtypes = [_fp, _fp, C.c_int, C.c_int, _dp, C.POINTER(C.c_int)]
    prev = np.ascontiguousarray(prev, np.float32)
    cur = np.ascontiguousarray(cur, np.float32)
    q = C.c_double(0)
    nc = C.c_int(0)
    pos = L.ora_stitch(_ptr(prev, _fp), _ptr(cur, _fp), prev.size, lat, C.byref(q), C.byref(nc))
    return pos, q.value, nc.value


def lf_forward(lib_, a, g, T, u):
    """Oracle relaxed loop filter: u (B, N) float64 -> y (B, N)."""
    L = lib_.lib
    L.or_lf_forward.argtypes = [_dp, _dp, C.c_int, C.c_double, _dp, C.c_int, C.c_size_t, _dp]
    a = np.ascontiguousarray(a, np.float64)
    g = np.ascontiguousarray(g, np.float64)
    u = np.ascontiguousarray(u, np.float64)
    y = np.zeros_like(u)
    L.or_lf_forward(_ptr(a, _dp), _ptr(g, _dp), a.size, T, _ptr(u, _dp), u.shape[0], u.shape[1],
                    _ptr(y, _dp))
    return y


def lf_backward(lib_, a, g, T, u, gy, trunc_every=0):
    """Oracle adjoint: returns (gu (B, N), ga (P,), gg (P,)); trunc_every > 0
    restarts the adjoint from zero every trunc_every samples."""
    L = lib_.lib
    L.or_lf_backward_tr.argtypes = [_dp, _dp, C.c_int, C.c_double, _dp, C.c_int, C.c_size_t, _dp, _dp,
                                    _dp, _dp, C.c_size_t]
    a = np.ascontiguousarray(a, np.float64)
    g = np.ascontiguousarray(g, np.float64)
    u = np.ascontiguousarray(u, np.float64)
    gy = np.ascontiguousarray(gy, np.float64)
    gu = np.zeros_like(u)
    ga = np.zeros(a.size)
    gg = np.zeros(a.size)
    L.or_lf_backward_tr(_ptr(a, _dp), _ptr(g, _dp), a.size, T, _ptr(u, _dp), u.shape[0], u.shape[1],
                        _ptr(gy, _dp), _ptr(gu, _dp), _ptr(ga, _dp), _ptr(gg, _dp), trunc_every)
    return gu, ga, gg


def pe_forward(lib_, w, p):
    L = lib_.lib
    L.or_pe_forward.argtypes = [_dp, C.c_int, _fp, C.c_int, C.c_size_t, _dp]
    w = np.ascontiguousarray(w, np.float64)
    p = np.ascontiguousarray(p, np.float32)
    u = np.zeros(p.shape)
    L.or_pe_forward(_ptr(w, _dp), w.size, _ptr(p, _fp), p.shape[0], p.shape[1], _ptr(u, _dp))
    return u


def pe_backward(lib_, ntaps, p, gu):
    L = lib_.lib
    L.or_pe_backward.argtypes = [C.c_int, _fp, C.c_int, C.c_size_t, _dp, _dp]
    p = np.ascontiguousarray(p, np.float32)
    gu = np.ascontiguousarray(gu, np.float64)
    gw = np.zeros(ntaps)
    L.or_pe_backward(ntaps, _ptr(p, _fp), p.shape[0], p.shape[1], _ptr(gu, _dp), _ptr(gw, _dp))
    return gw


def inband_weights(N, fs, f_hi=20000.0):
    """One-sided rFFT weights of the in-band error power (0..f_hi):
    L = sum_k w_k |E_k|^2 / N^2, w = 2 inside the band except DC/Nyquist (1)."""
    k = np.arange(N // 2 + 1)
    w = np.where(k * fs / N <= f_hi, 2.0, 0.0)
    w[0] = 1.0 if w[0] else 0.0
    if N % 2 == 0 and w[-1]:
        w[-1] = 1.0
    return w


class DasStream:
    """Oracle DAS over successive chunks with persistent contexts."""

    def __init__(self, lib_, filter_index, depth, cands, lat, nch):
        L = lib_.lib
        L.ora_das_stream_new.restype = C.c_void_p
        L.ora_das_stream_new.argtypes = [C.c_int] * 5
        L.ora_das_stream_chunk.restype = C.c_long
        L.ora_das_stream_chunk.argtypes = [C.c_void_p, _dp, C.c_size_t, C.c_size_t, C.c_int, C.c_int, _fp]
        L.ora_das_stream_free.argtypes = [C.c_void_p]
        self.L = L
        self.nch = nch
        self.h = L.ora_das_stream_new(filter_index, depth, cands, lat, nch)

    def chunk(self, x, seg_len, seg0_first=0, reposition=1, nseg=0):
        """seg_len > 0: segments of that length; seg_len == 0: nseg parts"""
        x = np.ascontiguousarray(x, dtype=np.float64)
        F = x.shape[1]
        out = np.zeros((self.nch, F), np.float32)
        self.L.ora_das_stream_chunk_n.restype = C.c_long
        self.L.ora_das_stream_chunk_n.argtypes = [C.c_void_p, _dp, C.c_size_t, C.c_size_t, C.c_int, C.c_int,
                                                  C.c_int, _fp]
        n = self.L.ora_das_stream_chunk_n(self.h, _ptr(x, _dp), F, seg_len, nseg, seg0_first, reposition,
                                          _ptr(out, _fp))
        return out[:, :n]

    def __del__(self):
        try:
            self.L.ora_das_stream_free(self.h)
        except Exception:
            pass
