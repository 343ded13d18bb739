"""The engine's pre-SDM conditioning in the oracle (or_cond.c) pinned by
what the mathematics fixes: exact integer window sums for the DSD boxcar,
exact (fsum) window sums for the PCM boxcar within its rounding, an
independent numpy Kaiser design and a float64 numpy convolution for the
lowpass, closed-form properties of the ramp and the soft clip."""
import math

import numpy as np
import pytest

import oracle_lib as ol

DSD64, DSD128, DSD256, DSD512 = 2822400, 5644800, 11289600, 22579200


def chunks(x, sizes):
    out, i = [], 0
    for s in sizes:
        out.append(x[i:i + s])
        i += s
    return out


@pytest.mark.parametrize("taps", [8, 16, 32, 64, 1, 5])
def test_dsd_boxcar_exact_windows(oracle, taps):
    rng = np.random.default_rng(taps)
    n = 3000
    x = np.where(rng.random(n) < 0.6, 1.0, -1.0).astype(np.float32)
    c = ol.Cond(oracle, taps)
    gains = [(0.708, 0.708), (0.5, 0.9), (0.9, 0.1)]
    parts = chunks(x, [700, 1, 1299, 1000])
    s = np.concatenate([[0] * taps, np.where(x > 0, 1, -1)]).astype(np.int64)
    cs = np.cumsum(s)
    pos = 0
    for k, p in enumerate(parts):
        gs, ge = gains[k % 3]
        y = c.dsd_boxcar(p, gs, ge)
        m = len(p)
        idx = np.arange(pos, pos + m) + taps
        win = cs[idx] - cs[idx - taps]                      # exact window sums, zeros before the start
        gstep = (ge - gs) / (m - 1) if m > 1 else 0.0
        ref = np.array([(float(w) * (1.0 / taps)) * (gs + gstep * float(i)) for i, w in enumerate(win)])
        assert np.array_equal(y, ref), k
        pos += m


def test_pcm_boxcar_windows(oracle):
    rng = np.random.default_rng(2)
    n, taps = 5000, 8
    x = (0.3 * rng.standard_normal(n)).astype(np.float32)
    x[1000:1100] *= 1e-9                                   # widely spread magnitudes
    c = ol.Cond(oracle, taps)
    y = np.concatenate([c.pcm_boxcar(p) for p in chunks(x, [100, 3, 2897, 2000])])
    xd = np.concatenate([[0.0] * taps, x.astype(np.float64)])
    for i in range(n):
        exact = math.fsum(xd[i + 1:i + taps + 1]) / taps
        # a running sum: error bounded by the magnitudes it has carried
        assert abs(y[i] - exact) <= 64 * 2.0 ** -53 * (np.max(np.abs(xd[:i + taps + 1])) + 1e-300), i


def test_lowpass_design_independent(oracle):
    for rate in (DSD64, DSD128, DSD256, DSD512):
        h = ol.lowpass_taps(oracle, rate)
        fc = min(50000.0 / (rate / 2.0), 0.5)
        M = 126
        x = np.arange(127) - M / 2.0
        ref = 2 * fc * np.sinc(2 * fc * x) * np.kaiser(127, 10.0)
        ref /= ref.sum()
        assert np.max(np.abs(h - ref)) < 1e-14
        assert abs(h.sum() - 1.0) < 1e-14 and np.allclose(h, h[::-1], atol=1e-17)


def test_lowpass_filter(oracle):
    rng = np.random.default_rng(3)
    n = 4000
    x = np.where(rng.random(n) < 0.5, 1.0, -1.0).astype(np.float32)
    c = ol.Cond(oracle, 8, DSD128)
    y = np.concatenate([c.dsd_lowpass(p, 1.0, 1.0) for p in chunks(x, [127, 1, 2000, 1872])])
    ref = np.convolve(x.astype(np.float64), ol.lowpass_taps(oracle, DSD128))[:n]
    assert np.max(np.abs(y - ref)) < 1e-13
    c2 = ol.Cond(oracle, 8, DSD128)
    yr = c2.dsd_lowpass(x, 0.2, 0.8)
    step = (0.8 - 0.2) / (n - 1)
    assert np.array_equal(yr, ol.Cond(oracle, 8, DSD128).dsd_lowpass(x, 1.0, 1.0) * (0.2 + step * np.arange(n)))


def test_ramp_and_clip(oracle):
    xp = np.linspace(0.0, 1.5, 10001)
    x = np.concatenate([-xp[:0:-1], xp])
    r = ol.gain_ramp(oracle, x, 0.3, 0.7)
    assert r[0] == x[0] * 0.3 and r[-1] == x[-1] * (0.3 + ((0.7 - 0.3) / 20000) * 20000)
    y = ol.soft_clip(oracle, x)
    inside = np.abs(x) <= 0.95
    assert np.array_equal(y[inside], x[inside])
    assert np.all(np.abs(y) < 1.0) and np.array_equal(y, -y[::-1])
    assert np.all(np.diff(y) >= 0)
    out = ~inside
    ref = np.sign(x[out]) * (0.95 + 0.05 * np.tanh((np.abs(x[out]) - 0.95) * (1.0 / (1.0 - 0.95))))
    assert np.max(np.abs(y[out] - ref)) < 4e-16
