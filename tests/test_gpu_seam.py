"""The gpu_compute.h seam (gpu_hip_*) with host buffers, as the plugin
calls it: one channel per call in channel order with gpu_hip_reset_chunk
per chunk (FIR chain, lowpass, boxcar), the trellis on one persistent
context, DAS over all channels, and the GPU_MIN_SAMPLES rejection.  Each
against the oracle or an exact numpy statement, bit for bit."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as ol
from foo_dsd_trellis_amd import fir as FIR
from foo_dsd_trellis_amd._lib import lib

pytestmark = pytest.mark.gpu

_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)


def fp(a):
    return a.ctypes.data_as(_fp)


def dp(a):
    return a.ctypes.data_as(_dp)


@pytest.fixture
def ctx(gpu):
    L = lib()
    h = L.gpu_hip_create()
    assert h
    yield h
    L.gpu_hip_destroy(h)


def test_probe_and_info(gpu):
    L = lib()
    assert L.gpu_hip_probe()


def test_fir_chain_per_channel_slots(ctx, oracle):
    """Two channels, two chunks each, one call per channel per chunk: every
    channel's output equals the oracle's chain over its whole input."""
    L = lib()
    fs, fo, nch, cf = 44100, 2822400, 2, 9000
    rng = np.random.default_rng(3)
    x = (0.3 * rng.standard_normal((nch, 2 * cf))).astype(np.float32)
    assert L.gpu_hip_fir_setup_rates(ctx, fs, fo, FIR.PCM_FIR_GAIN) == 0
    R = fo // fs
    got = np.zeros((nch, 2 * cf * R))
    for k in range(2):
        L.gpu_hip_reset_chunk(ctx)
        for c in range(nch):
            part = np.ascontiguousarray(x[c, k * cf:(k + 1) * cf])
            out = np.zeros(cf * R)
            n = C.c_size_t(0)
            assert L.gpu_hip_fir_chain_f64(ctx, fp(part), dp(out), cf, C.byref(n)) == 0
            assert n.value == cf * R
            got[c, k * cf * R:(k + 1) * cf * R] = out
    for c in range(nch):
        assert np.array_equal(got[c], ol.fir_chain(oracle, fs, fo, x[c], FIR.PCM_FIR_GAIN))
    small = np.zeros(100, np.float32)
    out = np.zeros(100 * R)
    assert L.gpu_hip_fir_chain_f64(ctx, fp(small), dp(out), 100, None) == -1


def test_trellis_and_das(ctx, oracle):
    L = lib()
    rate, K, Lat, B = 11289600, 16, 32, 8
    fi = oracle.ntf_auto(rate)
    f = oracle.ntf_get(fi)
    P = f["order"]
    a = np.ascontiguousarray(f["a"][:10])
    g = np.ascontiguousarray(f["g"][:10])
    assert L.gpu_hip_trellis_setup(ctx, K, P, Lat, dp(a), dp(g), 0.0, B) == 0
    rng = np.random.default_rng(4)
    x = np.clip(0.2 * rng.standard_normal(30000), -0.5, 0.5)
    ref = oracle.ctx(fi, B, K, Lat).process(x)
    outs = []
    for a0, b0 in [(0, 12000), (12000, 30000)]:
        part = np.ascontiguousarray(x[a0:b0])
        out = np.zeros(b0 - a0, np.float32)
        assert L.gpu_hip_trellis(ctx, dp(part), fp(out), b0 - a0) == 0
        outs.append(out)
    y = np.concatenate(outs)
    # the first Lat inputs fill the latency; outputs are packed at the front
    assert np.array_equal(y[:len(ref) - 0][:12000 - Lat], ref[:12000 - Lat])
    assert np.array_equal(outs[1], ref[12000 - Lat:30000 - Lat])

    # DAS over 2 channels, two chunks, against the oracle's persistent DAS
    nch, F, seg = 2, 70000, 16384
    assert L.gpu_hip_das_config(ctx, seg, 0, 1) == 0
    xs = np.clip(0.2 * rng.standard_normal((nch, 2 * F)), -0.5, 0.5)
    das = ol.DasStream(oracle, fi, B, K, Lat, nch)
    for k in range(2):
        part = np.ascontiguousarray(xs[:, k * F:(k + 1) * F])
        out = np.zeros((nch, F), np.float32)
        assert L.gpu_hip_trellis_das(ctx, dp(part), fp(out), F, nch) == 0
        want = das.chunk(part, seg, 0, 1)
        n = want.shape[1]
        assert np.array_equal(out[:, :n], want), k


def _lowpass_ref(bits_pm1, h, gain, hist):
    """sum_k h[k] s[i-k] in ascending k (zero before the stream's start),
    times gain; hist = the samples before this chunk."""
    full = np.concatenate([hist, bits_pm1])
    off = hist.size
    n = bits_pm1.size
    acc = np.zeros(n)
    for k in range(h.size):
        idx = np.arange(n) + off - k
        v = np.where(idx >= 0, full[np.clip(idx, 0, None)], 0.0)
        acc = acc + np.where(v == 0.0, 0.0, np.where(v > 0, h[k], -h[k]))
    return acc * gain


def test_lowpass_and_boxcar_slots(ctx, oracle):
    L = lib()
    nch, cf = 2, 9000
    rng = np.random.default_rng(6)
    s = np.where(rng.standard_normal((nch, 2 * cf)) >= 0, 1.0, -1.0).astype(np.float32)
    taps = np.zeros(127)
    assert L.dsdhip_lowpass_taps(5644800, dp(taps)) == 127
    tf = taps.astype(np.float32)
    assert L.gpu_hip_fir_lowpass_setup(ctx, fp(tf), 127) == 0
    h = tf.astype(np.float64)
    gain = 0.708
    for k in range(2):
        L.gpu_hip_reset_chunk(ctx)
        for c in range(nch):
            part = np.ascontiguousarray(s[c, k * cf:(k + 1) * cf])
            out = np.zeros(cf)
            assert L.gpu_hip_fir_lowpass_f64(ctx, fp(part), dp(out), cf, gain) == 0
            hist = s[c, :k * cf].astype(np.float64)
            want = _lowpass_ref(part.astype(np.float64), h, gain, hist)
            assert np.array_equal(out, want), (k, c)
    # boxcar: exact window sums of +-1 times 1/taps times gain, to float
    taps_b, g32 = 32, np.float32(0.708)
    for k in range(2):
        L.gpu_hip_reset_chunk(ctx)
        for c in range(nch):
            part = np.ascontiguousarray(s[c, k * cf:(k + 1) * cf])
            out = np.zeros(cf, np.float32)
            assert L.gpu_hip_boxcar_smooth(ctx, fp(part), fp(out), cf, taps_b, g32) == 0
            full = s[c, :(k + 1) * cf].astype(np.float64)
            cs = np.concatenate([[0.0], np.cumsum(full)])
            i = np.arange(k * cf, (k + 1) * cf)
            lo = np.maximum(i - taps_b + 1, 0)
            win = cs[i + 1] - cs[lo]
            want = ((win * (1.0 / taps_b)) * float(g32)).astype(np.float32)
            assert np.array_equal(out, want), (k, c)
