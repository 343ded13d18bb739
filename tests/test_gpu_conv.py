"""UPOLS convolution (complex FP32 spectra) against FP64 linear convolution
of the whole stream, over several chunks (the delay line crosses chunk
boundaries)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def run_stream(ir, x, P, chunk):
    from foo_dsd_trellis_amd import conv
    U = conv.Upols(ir, P)
    nch, n = x.shape
    d_x = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    d_y = torch.empty_like(d_x)
    for c0 in range(0, n, chunk):
        U.process_dev(d_x[:, c0:].data_ptr(), n, chunk, d_y[:, c0:].data_ptr(), n)
    torch.cuda.synchronize()
    return d_y.cpu().numpy(), U.partitions


@pytest.mark.parametrize("P,taps,chunk,nchunks", [(64, 700, 1024, 5), (256, 4096, 2048, 4),
                                                   (64, 64, 512, 3), (128, 3000, 8192, 2)])
def test_upols_matches_linear_convolution(gpu, P, taps, chunk, nchunks):
    rng = np.random.default_rng(P + taps)
    ir = rng.standard_normal((2, taps)) * np.exp(-np.arange(taps) / (taps / 3))
    x = 0.2 * rng.standard_normal((2, chunk * nchunks))
    y, npart = run_stream(ir, x, P, chunk)
    assert npart % 8 == 0 and npart * P >= taps
    for c in range(2):
        ref = np.convolve(x[c], ir[c])[:x.shape[1]]
        err = np.max(np.abs(y[c] - ref)) / np.max(np.abs(ref))
        assert err < 2e-5, err


def test_upols_room_ir_dsd256(gpu):
    """The streaming workload's shape: 6 channels, 2^20-tap room IRs,
    P = 8192 (128 partitions), chunks of 2,097,152 samples."""
    from scipy.signal import fftconvolve
    from foo_dsd_trellis_amd import signals
    ir = signals.room_ir(1 << 20, nch=6, seed=5)
    rng = np.random.default_rng(1)
    chunk, nchunks = 2097152, 2
    x = 0.1 * rng.standard_normal((6, chunk * nchunks))
    y, npart = run_stream(ir, x, 8192, chunk)
    assert npart == 128
    errs = []
    for c in range(6):
        ref = fftconvolve(x[c], ir[c])[:x.shape[1]]
        errs.append(np.max(np.abs(y[c] - ref)) / np.max(np.abs(ref)))
    print("UPOLS max error / max |y| per channel:", errs)
    assert max(errs) < 1e-4
