"""FIR chain oracle checked against closed forms and the reference's own
FIR tests (test/test_fir.c): taps, impulse response, DC gain, passband,
image rejection, and a long-double evaluation of the same sums."""
import numpy as np
import pytest

import oracle_lib as ol


def test_taps_closed_form(oracle):
    for T in (63, 127):
        h = ol.halfband(oracle, T)
        c = (T - 1) // 2
        assert abs(h.sum() - 1.0) < 1e-14
        assert np.allclose(h, h[::-1], atol=0, rtol=0)
        # half-band: odd offsets from the centre vanish (sinc at integers)
        assert np.max(np.abs(h[c + 2::2])) < 1e-15
        assert abs(h[c] - 0.5) < 1e-3
        # Kaiser beta=12 window times sinc(n/2), independent evaluation
        n = np.arange(T) - c
        w = np.i0(12.0 * np.sqrt(np.clip(1 - (n / c) ** 2, 0, None))) / np.i0(12.0)
        ref = np.sinc(n / 2.0) * w
        ref /= ref.sum()
        assert np.max(np.abs(h - ref)) < 1e-13


def test_impulse_response_one_stage(oracle):
    x = np.zeros(64, np.float32)
    x[0] = 1.0
    y = ol.fir_chain(oracle, 44100, 88200, x)
    h = ol.halfband(oracle, 63)
    assert np.array_equal(y[:63], h * 2.0)          # taps x2, phase-0 insertion
    assert np.all(y[63:] == 0)


def test_dc_gain_and_counts(oracle):
    x = np.ones(4000, np.float32)
    y = ol.fir_chain(oracle, 44100, 2822400, x)
    assert y.size == 4000 * 64
    assert abs(y[-1000:].mean() - 1.0) < 1e-7          # 6-stage DC ripple ~4e-9


def _goertzel(x, f, fs):
    w = 2 * np.pi * f / fs
    c = 2 * np.cos(w)
    s1 = s2 = 0.0
    for v in x.astype(np.float64):
        s0 = v + c * s1 - s2
        s2, s1 = s1, s0
    re, im = s1 - s2 * np.cos(w), s2 * np.sin(w)
    return (re * re + im * im) / (x.size * x.size)


def _tone(fs, n, f):
    k = int(f / (fs / n) + 0.5) * (fs / n)          # bin-exact frequency
    return k, (0.5 * np.sin(2 * np.pi * k * np.arange(n) / fs)).astype(np.float32)


@pytest.mark.parametrize("f", [1000.0, 5000.0, 10000.0, 15000.0, 20000.0])
def test_passband_flat(oracle, f):
    """test/test_fir.c:170-216: DSD64 -> DSD128, +-0.1 dB to 20 kHz."""
    fs = 2822400
    n = 4096
    k, x = _tone(fs, n, f)
    y = ol.fir_chain(oracle, fs, 2 * fs, x)
    g = 10 * np.log10(_goertzel(y, k, 2 * fs) / _goertzel(x, k, fs))
    assert abs(g) < 0.1


def test_image_rejection(oracle):
    """test/test_fir.c:218-254: the image at fs_in - f is below -80 dB."""
    fs = 2822400
    n = 4096
    k, x = _tone(fs, n, 10000.0)
    y = ol.fir_chain(oracle, fs, 2 * fs, x)
    ratio = _goertzel(y, fs - k, 2 * fs) / _goertzel(y, k, 2 * fs)
    assert 10 * np.log10(ratio) < -80


def test_audio_band_44k1(oracle):
    """First stage from 44.1 kHz: 1 kHz and 10 kHz flat to 0.1 dB."""
    fs = 44100
    n = 8192
    for f in (1000.0, 10000.0):
        k, x = _tone(fs, n, f)
        y = ol.fir_chain(oracle, fs, 4 * fs, x)
        g = 10 * np.log10(_goertzel(y, k, 4 * fs) / _goertzel(x, k, fs))
        assert abs(g) < 0.1, f


def test_double_vs_long_double(oracle):
    rng = np.random.default_rng(1)
    x = (0.5 * rng.standard_normal(3000)).astype(np.float32)
    y = ol.fir_chain(oracle, 44100, 11289600, x, 0.354)
    t = ol.fir_chain(oracle, 44100, 11289600, x, 0.354, long_double=True)
    assert np.max(np.abs(y - t)) < 1e-12
