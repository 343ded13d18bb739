"""Streaming chain (FIR -> UPOLS -> DAS per chunk) against the oracle: the
SDM input of every chunk against the oracle's FP64 FIR chain followed by an
FP64 linear convolution of the whole stream (the convolution runs on complex
FP32 spectra: relative error bound stated below), and the DSD bits against
the oracle's per-chunk DAS with persistent contexts run on the same SDM
input (bit-exact)."""
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
        want = das.chunk(sig, SEG, 0, 1)
        assert n == want.shape[1], (f0, n, want.shape)
        assert np.array_equal(S.bits_numpy(n), want), "chunk at frame %d" % f0
        total += n
    assert total > (frames * R) * 0.9
