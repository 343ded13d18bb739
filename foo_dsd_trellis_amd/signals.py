"""Seeded synthetic PCM inputs for the benchmark workloads (numpy PCG64).

All generators return float32 arrays shaped (channels, samples).
"""
import numpy as np


def sine_noise(n, fs=44100, nch=2, freq=997.0, amp=0.5012, noise=1e-5, seed=1):
    """997 Hz sine at -6 dBFS plus Gaussian noise at -100 dBFS."""
    rng = np.random.default_rng(seed)
    t = np.arange(n) / fs
    s = amp * np.sin(2 * np.pi * freq * t)
    return (s[None, :] + noise * rng.standard_normal((nch, n))).astype(np.float32)


def pink(n, rng):
    """Unit-RMS 1/f-shaped Gaussian noise (amplitude spectrum ~ 1/sqrt(f))."""
    w = rng.standard_normal(n)
    f = np.fft.rfft(w)
    k = np.arange(f.size, dtype=np.float64)
    k[0] = 1.0
    f /= np.sqrt(k)
    f[0] = 0.0
    y = np.fft.irfft(f, n)
    return y / np.sqrt(np.mean(y * y))


def pink_multitone(n, fs=44100, nch=2, seed=2, pink_rms=0.2512, tone_amp=0.1, ntones=7,
                   fmin=100.0, fmax=15000.0, clip=0.5):
    """Pink noise at -12 dBFS RMS clipped to +-0.5 plus 7 log-spaced tones
    100 Hz..15 kHz at -20 dBFS each."""
    rng = np.random.default_rng(seed)
    t = np.arange(n) / fs
    freqs = np.geomspace(fmin, fmax, ntones)
    phases = rng.uniform(0, 2 * np.pi, size=(nch, ntones))
    out = np.empty((nch, n), np.float32)
    for c in range(nch):
        p = np.clip(pink_rms * pink(n, rng), -clip, clip)
        tones = sum(tone_amp * np.sin(2 * np.pi * f * t + phases[c, i]) for i, f in enumerate(freqs))
        out[c] = (p + tones).astype(np.float32)
    return out


def bandlimited_gauss(n, fs=48000, nch=6, seed=3, rms=0.1995, fmax=20000.0):
    """Gaussian noise low-passed at 20 kHz (FFT brick wall), -14 dBFS RMS."""
    rng = np.random.default_rng(seed)
    out = np.empty((nch, n), np.float32)
    for c in range(nch):
        f = np.fft.rfft(rng.standard_normal(n))
        fr = np.fft.rfftfreq(n, 1.0 / fs)
        f[fr > fmax] = 0.0
        y = np.fft.irfft(f, n)
        out[c] = (rms * y / np.sqrt(np.mean(y * y))).astype(np.float32)
    return out


def uneven_tracks(ntracks=16, nch=6, fs=48000, lo_s=5.0, hi_s=120.0, seed=4):
    """Lengths (samples) of the batched workload's tracks, U[5, 120] s."""
    rng = np.random.default_rng(seed)
    return [int(round(s * fs)) for s in rng.uniform(lo_s, hi_s, size=ntracks)]


def multitone_noise(nrows, n, fs=5644800, seed=9, dbfs=-10.0, ntones=5, fmin=200.0, fmax=18000.0):
    """Rows of seeded multitone plus Gaussian noise at `dbfs` RMS (float32):
    equal-amplitude log-spaced tones carry half the power, noise the rest."""
    rng = np.random.default_rng(seed)
    rms = 10.0 ** (dbfs / 20.0)
    t = np.arange(n) / fs
    freqs = np.geomspace(fmin, fmax, ntones)
    amp = np.sqrt(2.0 * 0.5 * rms * rms / ntones)
    out = np.empty((nrows, n), np.float32)
    for r in range(nrows):
        ph = rng.uniform(0, 2 * np.pi, ntones)
        x = sum(amp * np.sin(2 * np.pi * f * t + ph[i]) for i, f in enumerate(freqs))
        x = x + np.sqrt(0.5) * rms * rng.standard_normal(n)
        out[r] = x.astype(np.float32)
    return out


def room_ir(taps=1 << 20, fs=11289600, nch=6, rt60=0.4, seed=5):
    """Seeded synthetic room impulse responses (nch, taps) float64:
    Gaussian noise with an exponential decay of 60 dB over rt60 seconds,
    scaled to unit energy per channel."""
    rng = np.random.default_rng(seed)
    t = np.arange(taps) / fs
    env = np.exp(-np.log(1000.0) * t / rt60)
    out = np.empty((nch, taps))
    for c in range(nch):
        h = rng.standard_normal(taps) * env
        out[c] = h / np.sqrt(np.sum(h * h))
    return out
