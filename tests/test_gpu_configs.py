"""The batched uneven workload and the streaming workload at their
configured shapes, at reduced lengths: bits against the oracle chunk by
chunk.

- Batched: 48 kHz -> DSD512/48 (24.576 MHz: the 127-tap first stage plus
  eight 63-tap stages), K=32, L=64, B=8, 6 channels, three tracks of uneven
  length, every segment of every channel and track in one trellis launch
  per chunk (at DSD512 and up the reference seeds every segment from the
  persistent state, src/dsp_plugin.c:2185).
- Streaming: 44.1 kHz -> DSD256 with UPOLS room correction at P = 8192
  (complex FP32 spectra), DAS K=32, L=64 seeded seg0-first as the reference
  seeds at DSD256; the SDM input is checked against FP64 linear convolution
  of the oracle's FIR output, the bits against the oracle's chunked DAS on
  the GPU's SDM input.
"""
import numpy as np
import pytest
import torch

import oracle_lib as ol

pytestmark = pytest.mark.gpu

K, L, B = 32, 64, 8


def _gpu_bits(words, nbits):
    w = words.cpu().numpy().view(np.uint64)
    bits = np.unpackbits(w.view(np.uint8), axis=1, bitorder="little")[:, :nbits]
    return np.where(bits > 0, 1.0, -1.0).astype(np.float32)


def test_batched_dsd512_48_k32_six_channels(gpu, oracle):
    from foo_dsd_trellis_amd import batch, fir, signals, trellis
    fs_in, fs_out, nch, cf, seg = 48000, 24576000, 6, 512, 32768
    R = fs_out // fs_in
    lens = [600, 350, 900]                       # frames: 2, 1 and 2 chunks (ragged ends)
    print("case: batched DSD512/48 K32 L64 6 ch, tracks %s frames, chunk %d, seg %d" % (lens, cf, seg),
          flush=True)
    pcm = [np.ascontiguousarray(signals.bandlimited_gauss(n, fs=fs_in, nch=nch, seed=30 + t))
           for t, n in enumerate(lens)]
    pcm[0][:, 100:300] *= 1e-3                   # a quiet stretch
    fi = trellis.ntf_auto(fs_out)
    assert fir.plan(fs_in, fs_out) == [127] + [63] * 8
    plan = batch.BatchPlan(fs_in, fs_out, fi, K=K, L=L, depth=B, nch=nch, chunk_frames=cf, seg_len=seg,
                           use_graph=False)
    plan.enqueue([torch.from_numpy(x).cuda() for x in pcm])
    bits = plan.run()
    for t, x in enumerate(pcm):
        y = np.stack([ol.fir_chain(oracle, fs_in, fs_out, x[c], fir.PCM_FIR_GAIN) for c in range(nch)])
        st = ol.DasStream(oracle, fi, B, K, L, nch)
        parts = [st.chunk(y[:, f0 * R:min(len(x[0]), f0 + cf) * R], seg, 0, 1) for f0 in range(0, x.shape[1], cf)]
        ref = np.concatenate(parts, axis=1)
        assert bits[t] == ref.shape[1], (t, bits[t], ref.shape)
        assert np.array_equal(_gpu_bits(plan.outputs[t], int(bits[t])), ref), "track %d" % t


def test_streaming_dsd256_room_p8192_k32(gpu, oracle):
    import scipy.signal
    from foo_dsd_trellis_amd import fir, signals, stream, trellis
    fs_in, fs_out, nch, cf, P, seg = 44100, 11289600, 2, 1024, 8192, 8192
    frames = 3 * cf
    print("case: streaming DSD256 K32 L64 P8192, %d ch x %d frames, chunk %d" % (nch, frames, cf), flush=True)
    pcm = np.ascontiguousarray(signals.pink_multitone(frames, fs=fs_in, nch=nch, seed=11))
    ir = 0.5 * signals.room_ir(1 << 17, fs=fs_out, nch=nch, rt60=0.05, seed=6)
    fi = trellis.ntf_auto(fs_out)
    S = stream.Stream(fs_in, fs_out, nch, cf, ir=ir, P=P, ntf_index=fi, K=K, L=L, depth=B, seg_len=seg)
    assert S.seg0_first
    R = S.R
    y = np.stack([ol.fir_chain(oracle, fs_in, fs_out, pcm[c], fir.PCM_FIR_GAIN) for c in range(nch)])
    yc = np.stack([scipy.signal.fftconvolve(y[c], ir[c])[:y.shape[1]] for c in range(nch)])
    das = ol.DasStream(oracle, fi, B, K, L, nch)
    d_pcm = torch.from_numpy(pcm).cuda()
    for k, f0 in enumerate(range(0, frames, cf)):
        n = S.process(d_pcm[:, f0:f0 + cf])
        sig = S.signal()
        err = np.max(np.abs(sig - yc[:, f0 * R:(f0 + cf) * R])) / np.max(np.abs(yc))
        assert err < 2e-5, (k, err)                # FP32 spectra against FP64 convolution
        want = das.chunk(sig, seg, 1, 1)
        assert n == want.shape[1] == (cf * R - L if k == 0 else cf * R), (k, n)
        assert np.array_equal(S.bits_numpy(n), want), k
