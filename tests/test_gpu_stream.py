"""Streaming chain (FIR -> UPOLS -> DAS per chunk) against the oracle: the
SDM input of every chunk against the oracle's FP64 FIR chain followed by an
FP64 linear convolution of the whole stream (the convolution runs on complex
FP32 spectra: relative error bound stated below), and the DSD bits against
the oracle's per-chunk DAS with persistent contexts run on the same SDM
input (bit-exact), seeded as the reference seeds at the output rate
(seg0-first below DSD512)."""
import numpy as np
import pytest
import torch

import oracle_lib as ol

pytestmark = pytest.mark.gpu

FS_IN, FS_OUT = 44100, 2822400          # x64, 6 half-band stages
K, L, DEPTH = 16, 32, 8
CF, P = 1024, 1024                      # 65,536 samples per chunk = 8 blocks of 8*P
SEG = 16384


def setup(seed=4, nch=2, frames=3 * CF + 512):
    from foo_dsd_trellis_amd import signals
    pcm = signals.bandlimited_gauss(frames, fs=FS_IN, nch=nch, seed=seed, fmax=18000.0)
    pcm[:, 1500:2100] *= 1e-3            # quiet stretch: boundary reposition
    ir = signals.room_ir(20000, fs=FS_OUT, nch=nch, rt60=0.002, seed=seed)
    ir *= 0.5
    return np.ascontiguousarray(pcm), ir


@pytest.mark.parametrize("with_ir", [True, False])
def test_stream_matches_oracle(gpu, oracle, with_ir):
    from foo_dsd_trellis_amd import fir, stream, trellis
    pcm, ir = setup()
    nch, frames = pcm.shape
    fi = trellis.ntf_auto(FS_OUT)
    S = stream.Stream(FS_IN, FS_OUT, nch, CF, ir=ir if with_ir else None, P=P, ntf_index=fi, K=K, L=L,
                      depth=DEPTH, seg_len=SEG)
    R = S.R
    assert R == FS_OUT // FS_IN
    y = np.stack([ol.fir_chain(oracle, FS_IN, FS_OUT, pcm[c], fir.PCM_FIR_GAIN) for c in range(nch)])
    if with_ir:
        y = np.stack([np.convolve(y[c], ir[c])[:y.shape[1]] for c in range(nch)])
    das = ol.DasStream(oracle, fi, DEPTH, K, L, nch)
    d_pcm = torch.from_numpy(pcm).cuda()
    total = 0
    for f0 in range(0, frames, CF):
        f1 = min(frames, f0 + CF)
        n = S.process(d_pcm[:, f0:f1])
        sig = S.signal()
        assert sig.shape == (nch, (f1 - f0) * R)
        ref = y[:, f0 * R:f1 * R]
        err = np.max(np.abs(sig - ref)) / np.max(np.abs(y))
        assert err < (2e-5 if with_ir else 1e-9), (f0, err)
        want = das.chunk(sig, SEG, int(S.seg0_first), 1)
        assert n == want.shape[1], (f0, n, want.shape)
        assert np.array_equal(S.bits_numpy(n), want), "chunk at frame %d" % f0
        # host timeline of the chunk: DAS start, then the points the run
        # passed (zero otherwise), in order, ending with the DAS return
        tl = S.last_timeline()
        assert len(tl) == 7 and 0.0 < tl[0] <= tl[6], tl
        passed = [t for t in tl if t > 0.0]
        assert passed == sorted(passed), tl
        total += n
    assert total > (frames * R) * 0.9


def ramp_pts(fir_gain, prev, vol):
    """src/engine.c:603-606"""
    ge = fir_gain * vol
    gs = fir_gain * prev
    return (ge if gs == 0.0 else gs), ge


def sdm1(x):
    """a first-order 1-bit modulator: a DSD-like +-1 stream for inputs"""
    y = np.empty(x.size, np.float32)
    acc = 0.0
    for i, v in enumerate(x):
        acc += v - (1.0 if acc >= 0 else -1.0) * 1.0
        y[i] = 1.0 if acc >= 0 else -1.0
    return y


def check_stream(S, das, sigs_ref, feed, nch):
    """feed(k) runs chunk k on the GPU; sigs_ref[k] is the oracle's SDM input"""
    total = 0
    for k, ref in enumerate(sigs_ref):
        n = feed(k)
        sig = S.signal()
        assert sig.shape == ref.shape, k
        bad = np.argwhere(sig != ref)
        assert bad.size == 0, (k, bad[:3], sig[tuple(bad[0])], ref[tuple(bad[0])])
        want = das.chunk(sig, SEG, int(S.seg0_first), 1)
        assert n == want.shape[1], k
        assert np.array_equal(S.bits_numpy(n), want), k
        total += n
    return total


@pytest.mark.parametrize("clip", [True, False])
def test_stream_pcm_engine_conditioning(gpu, oracle, clip):
    """engine_process_block's PCM chain: boxcar(8) pre-smooth, FIR, gain
    ramp across volume changes, soft clip, DAS."""
    from foo_dsd_trellis_amd import signals, stream, trellis
    nch, frames = 2, 3 * CF + 512
    pcm = 4.0 * signals.bandlimited_gauss(frames, fs=FS_IN, nch=nch, seed=8, fmax=2000.0)
    pcm = np.ascontiguousarray(np.clip(pcm, -1.0, 1.0).astype(np.float32))
    fg = float(np.float32(1.0))
    vols = [1.25, 0.9, 0.9, 1.1]
    fi = trellis.ntf_auto(FS_OUT)
    S = stream.Stream(FS_IN, FS_OUT, nch, CF, ntf_index=fi, K=K, L=L, depth=DEPTH, seg_len=SEG, gain=fg,
                      volume=vols[0], pcm_boxcar=8, soft_clip=clip)
    R = S.R
    y = []
    for c in range(nch):
        bx = ol.Cond(oracle, 8).pcm_boxcar(pcm[c])
        y.append(ol.fir_chain_d(oracle, FS_IN, FS_OUT, bx))
    y = np.stack(y)
    refs, prev = [], 0.0
    for k, f0 in enumerate(range(0, frames, CF)):
        f1 = min(frames, f0 + CF)
        gs, ge = ramp_pts(fg, prev, vols[k])
        prev = vols[k]
        part = np.stack([ol.gain_ramp(oracle, y[c, f0 * R:f1 * R], gs, ge) for c in range(nch)])
        if clip:
            part = np.stack([ol.soft_clip(oracle, part[c]) for c in range(nch)])
        refs.append(part)
    assert clip is False or np.any(np.abs(np.concatenate(refs, axis=1)) > 0.95)
    d_pcm = torch.from_numpy(pcm).cuda()

    def feed(k):
        S.set_gain(vols[k])
        f0 = k * CF
        return S.process(d_pcm[:, f0:min(frames, f0 + CF)])

    das = ol.DasStream(oracle, fi, DEPTH, K, L, nch)
    check_stream(S, das, refs, feed, nch)


@pytest.mark.parametrize("lowpass", [False, True])
def test_stream_dsd_same_rate(gpu, oracle, lowpass):
    """same-rate DSD re-encode: boxcar (the engine's 32 taps at DSD64) or
    the 127-tap lowpass with the gain ramp, then DAS."""
    from foo_dsd_trellis_amd import cond, stream, trellis
    rate, nch, cf = 2822400, 2, 32768
    t = np.arange(4 * cf + 4096) / rate
    x = np.stack([sdm1(0.5 * np.sin(2 * np.pi * (997 + 300 * c) * t)) for c in range(nch)])
    fg = float(np.float32(0.708))
    vols = [1.0, 1.0, 0.5, 0.8, 0.8]
    fi = trellis.ntf_auto(rate)
    S = stream.Stream(rate, rate, nch, cf, ntf_index=fi, K=K, L=L, depth=DEPTH, seg_len=SEG, gain=fg,
                      volume=vols[0], input=stream.IN_DSD, dsd_lowpass=lowpass)
    assert S.R == 1
    conds = [ol.Cond(oracle, 32, rate if lowpass else 0) for _ in range(nch)]
    refs, prev = [], 0.0
    n_total = x.shape[1]
    starts = list(range(0, n_total, cf))
    for k, s0 in enumerate(starts):
        s1 = min(n_total, s0 + cf)
        gs, ge = ramp_pts(fg, prev, vols[k])
        prev = vols[k]
        fn = (lambda c, p: conds[c].dsd_lowpass(p, gs, ge)) if lowpass else (lambda c, p: conds[c].dsd_boxcar(p, gs, ge))
        refs.append(np.stack([fn(c, x[c, s0:s1]) for c in range(nch)]))

    def feed(k):
        S.set_gain(vols[k])
        s0 = starts[k]
        part = torch.from_numpy(np.ascontiguousarray(x[:, s0:min(n_total, s0 + cf)])).cuda()
        return S.process_dsd(cond.to_bits(part), part.shape[1])

    das = ol.DasStream(oracle, fi, DEPTH, K, L, nch)
    check_stream(S, das, refs, feed, nch)


def test_stream_dsd_rate_conversion(gpu, oracle):
    """DSD64 -> DSD128 as engine_process_block runs it: boxcar(8) of the
    +-1 stream, FIR, ramp, soft clip, DAS."""
    from foo_dsd_trellis_amd import cond, stream, trellis
    fs_in, fs_out, nch, cf = 2822400, 5644800, 2, 16384
    t = np.arange(3 * cf + 1000) / fs_in
    x = np.stack([sdm1(0.6 * np.sin(2 * np.pi * (1500 + 500 * c) * t)) for c in range(nch)])
    fg = float(np.float32(0.708))
    vol = 1.0
    fi = trellis.ntf_auto(fs_out)
    S = stream.Stream(fs_in, fs_out, nch, cf, ntf_index=fi, K=K, L=L, depth=DEPTH, seg_len=SEG, gain=fg,
                      volume=vol, input=stream.IN_DSD, dsd_boxcar=8, soft_clip=True)
    R = S.R
    assert R == 2
    y = np.stack([ol.fir_chain_d(oracle, fs_in, fs_out, ol.Cond(oracle, 8).dsd_boxcar(x[c], 1.0, 1.0))
                  for c in range(nch)])
    n_total = x.shape[1]
    starts = list(range(0, n_total, cf))
    refs = []
    for k, s0 in enumerate(starts):
        s1 = min(n_total, s0 + cf)
        gs, ge = ramp_pts(fg, 0.0 if k == 0 else vol, vol)
        refs.append(np.stack([ol.soft_clip(oracle, ol.gain_ramp(oracle, y[c, s0 * R:s1 * R], gs, ge))
                              for c in range(nch)]))

    def feed(k):
        s0 = starts[k]
        part = torch.from_numpy(np.ascontiguousarray(x[:, s0:min(n_total, s0 + cf)])).cuda()
        return S.process_dsd(cond.to_bits(part), part.shape[1])

    das = ol.DasStream(oracle, fi, DEPTH, K, L, nch)
    check_stream(S, das, refs, feed, nch)
