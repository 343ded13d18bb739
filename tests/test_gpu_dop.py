"""DoP and raw-bit formats on the GPU against the oracle (or_dop.c, pinned
to the reference's dop.c): multi-channel interleaved frames, frame counts
that do not fill a 64-bit word, marker phases, partial bytes."""
import numpy as np
import pytest
import torch

import oracle_lib as ol

pytestmark = pytest.mark.gpu


def pm1(bits, n):
    w = bits.cpu().numpy().view(np.uint64)
    b = np.unpackbits(w.view(np.uint8), axis=1, bitorder="little")[:, :n]
    return np.where(b > 0, 1.0, -1.0).astype(np.float32), b


def dop_frames(rng, frames, nch):
    words = rng.integers(0, 1 << 16, (frames, nch))
    v = np.where(np.arange(frames)[:, None] & 1, 0xFA, 0x05) << 16 | words
    v = np.where(v & 0x800000, v - (1 << 24), v)
    return (v / 8388608.0).astype(np.float32)


@pytest.mark.parametrize("frames,nch", [(1, 1), (3, 2), (4, 6), (1001, 2), (8192, 6)])
def test_dop_unpack_pack(gpu, oracle, frames, nch):
    from foo_dsd_trellis_amd import dop
    od = ol.Dop(oracle, "ora")
    rng = np.random.default_rng(frames + nch)
    pcm = dop_frames(rng, frames, nch)
    assert dop.detect(pcm.ravel(), nch) == od.detect(pcm.ravel(), nch)
    bits = dop.unpack(torch.from_numpy(pcm).cuda())
    f, raw = pm1(bits, frames * 16)
    for c in range(nch):
        assert np.array_equal(f[c], od.unpack(pcm[:, c])), c
    # bits past the stream are zero
    tail = np.unpackbits(bits.cpu().numpy().view(np.uint64).view(np.uint8), axis=1, bitorder="little")
    assert not tail[:, frames * 16:].any()
    nbits = frames * 16 - (5 if frames > 1 else 0)
    out = dop.pack(bits, nbits).cpu().numpy()
    for c in range(nch):
        assert np.array_equal(out[:, c], od.pack(f[c, :nbits])), c
    for ph in (0, 1):
        o24, nph = dop.pack_i24(bits, nbits, ph)
        o24 = o24.cpu().numpy()
        for c in range(nch):
            ref, rph = od.pack_i24(f[c, :nbits], ph)
            assert rph == nph and np.array_equal(o24[:, c, :].ravel(), ref), (c, ph)


@pytest.mark.parametrize("n_bits,nch", [(5, 1), (64, 2), (1000, 3), (65536 + 13, 2)])
def test_bytes_and_f32(gpu, oracle, n_bits, nch):
    from foo_dsd_trellis_amd import dop
    od = ol.Dop(oracle, "ora")
    rng = np.random.default_rng(n_bits)
    src = rng.integers(0, 256, (nch, (n_bits + 7) // 8)).astype(np.uint8)
    bits = dop.bytes_unpack(torch.from_numpy(src).cuda(), n_bits)
    f, _ = pm1(bits, n_bits)
    for c in range(nch):
        assert np.array_equal(f[c], od.bits_unpack(src[c], n_bits)), c
    back = dop.bytes_pack(bits, n_bits).cpu().numpy()
    for c in range(nch):
        assert np.array_equal(back[c], od.bits_pack(f[c])), c
    assert np.array_equal(dop.to_f32(bits, n_bits).cpu().numpy(), f)
