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


@pytest.mark.parametrize("P,taps,chunk,nchunks", [(64, 700, 320, 5), (256, 4096, 256 * 11, 3)])
def test_upols_any_block_count(gpu, P, taps, chunk, nchunks):
    """Chunks of any whole number of blocks (not a multiple of the 8-block
    tile of the spectral multiply-accumulate)."""
    rng = np.random.default_rng(P + taps + 1)
    ir = rng.standard_normal((2, taps)) * np.exp(-np.arange(taps) / (taps / 3))
    x = 0.2 * rng.standard_normal((2, chunk * nchunks))
    y, _ = run_stream(ir, x, P, chunk)
    for c in range(2):
        ref = np.convolve(x[c], ir[c])[:x.shape[1]]
        err = np.max(np.abs(y[c] - ref)) / np.max(np.abs(ref))
        assert err < 2e-5, err


def fifo_model(sizes, P):
    """Per chunk: (number of processed samples written, pending input,
    waiting output) by the counting of conv_process_direct
    (src/convolution.c:893-958): whole blocks of the pending input join the
    waiting output; exactly n are handed out, or all waiting ones when fewer
    (the rest of the chunk keeps its input) and nothing is carried over."""
    buf_pos, fifo, out = 0, 0, []
    for n in sizes:
        nb = (buf_pos + n) // P
        buf_pos = buf_pos + n - nb * P
        avail = fifo + nb * P
        if avail >= n:
            out.append(n)
            fifo = avail - n
        else:
            out.append(avail)
            fifo = 0
        yield out[-1], buf_pos, fifo


def test_upols_fifo_any_chunk_length(gpu):
    """The FIFO form over chunks of arbitrary length: every output position
    holds either the next processed sample in stream order (against FP64
    linear convolution, each processed sample exactly once) or, where the
    waiting output ran short, exactly the chunk's own input sample; the
    counts follow the reference's FIFO; in-place (d_out == d_in) too."""
    from foo_dsd_trellis_amd import conv
    P, taps = 256, 2000
    rng = np.random.default_rng(9)
    ir = rng.standard_normal((2, taps)) * np.exp(-np.arange(taps) / 600.0)
    sizes = [1000, 37, 500, 2048, 65, 3000, 1, 256, 777, 4096, 3, 1500]
    n = sum(sizes)
    x = 0.2 * rng.standard_normal((2, n))
    ref = np.stack([np.convolve(x[c], ir[c])[:n] for c in range(2)])
    for inplace in (False, True):
        U = conv.Upols(ir, P)
        d_x = torch.from_numpy(np.ascontiguousarray(x)).cuda()
        d_y = d_x if inplace else torch.zeros_like(d_x)
        model = fifo_model(sizes, P)
        c0, done = 0, 0
        for k, m in enumerate(sizes):
            U.process_fifo_dev(d_x[:, c0:].data_ptr(), n, m, d_y[:, c0:].data_ptr(), n)
            torch.cuda.synchronize()
            got, pend, wait = next(model)
            assert U.fifo_state() == (pend, wait), (k, U.fifo_state(), pend, wait)
            y = d_y[:, c0:c0 + m].cpu().numpy()
            proc = y[:, :got]
            err = np.max(np.abs(proc - ref[:, done:done + got])) / np.max(np.abs(ref)) if got else 0.0
            assert err < 2e-5, (k, err)
            assert np.array_equal(y[:, got:], x[:, c0 + got:c0 + m]), k
            done += got
            c0 += m
        assert done <= n


def test_upols_fifo_equals_block_form_on_whole_blocks(gpu):
    """Chunks that are whole blocks: the FIFO form writes exactly what the
    block form writes, bit for bit, and holds nothing between chunks."""
    from foo_dsd_trellis_amd import conv
    P, taps = 128, 1500
    rng = np.random.default_rng(10)
    ir = rng.standard_normal((2, taps)) * np.exp(-np.arange(taps) / 400.0)
    sizes = [P * 3, P * 8, P, P * 5]
    n = sum(sizes)
    x = torch.from_numpy(0.2 * rng.standard_normal((2, n))).cuda()
    outs = []
    for fifo in (False, True):
        U = conv.Upols(ir.copy(), P)
        y = torch.zeros_like(x)
        c0 = 0
        for m in sizes:
            f = U.process_fifo_dev if fifo else U.process_dev
            f(x[:, c0:].data_ptr(), n, m, y[:, c0:].data_ptr(), n)
            c0 += m
            if fifo:
                assert U.fifo_state() == (0, 0)
        torch.cuda.synchronize()
        outs.append(y.cpu().numpy())
    assert np.array_equal(outs[0], outs[1])
