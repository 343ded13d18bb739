"""DoP / raw-bit formats: the oracle (or_dop.c) against the known answers of
test/test_dop.c restated, and against the reference's own src/dop.c
(compiled into oracle/_ref) on random frames."""
import numpy as np
import pytest

import oracle_lib as ol

A, B = 0x05, 0xFA


def dop_sample(marker, bits16):
    """test/test_dop.c:11-17: a float carrying a 24-bit DoP word."""
    v = (marker << 16) | bits16
    if v & 0x800000:
        v -= 1 << 24
    return np.float32(v / 8388608.0)


@pytest.fixture(scope="module")
def od(oracle):
    return ol.Dop(oracle, "ora")


def test_detect_known(od):
    assert not od.detect(np.array([dop_sample(A, 0)], np.float32))
    assert od.detect(np.array([dop_sample(A, 0xAAAA), dop_sample(B, 0x5555)], np.float32))
    pcm8 = np.array([dop_sample(B if i & 1 else A, (i * 0x1111) & 0xFFFF) for i in range(8)], np.float32)
    assert od.detect(pcm8)
    big = np.concatenate([pcm8, np.full(56, 0.25, np.float32)])
    assert od.detect(big)                                          # only the first 8 are checked
    assert not od.detect(np.array([dop_sample(A, 1), dop_sample(A, 2)], np.float32))
    assert not od.detect(np.array([dop_sample(B, 1), dop_sample(A, 2)], np.float32))
    rng = np.random.default_rng(1)
    assert not od.detect(rng.uniform(-1, 1, 8).astype(np.float32))
    inter = np.stack([pcm8, np.zeros(8, np.float32)], axis=1).ravel()
    assert od.detect(inter, 2)


def test_unpack_known(od):
    assert np.all(od.unpack(np.array([dop_sample(A, 0xFFFF)], np.float32)) == 1.0)
    assert np.all(od.unpack(np.array([dop_sample(A, 0x0000)], np.float32)) == -1.0)
    alt = od.unpack(np.array([dop_sample(A, 0xAAAA)], np.float32))
    assert np.array_equal(alt, np.tile([1.0, -1.0], 8).astype(np.float32))


@pytest.mark.parametrize("pat", [[0xDEAD, 0xBEEF, 0x1234, 0x5678], [0x6969] * 8])
def test_roundtrip_known(od, pat):
    pcm = np.array([dop_sample(B if i & 1 else A, p) for i, p in enumerate(pat)], np.float32)
    out = od.pack(od.unpack(pcm))
    assert np.array_equal(np.round(pcm.astype(np.float64) * 8388608), np.round(out.astype(np.float64) * 8388608))


def test_bits_known(od):
    assert np.all(od.bits_unpack(np.array([0xFF], np.uint8), 8) == 1.0)
    assert np.all(od.bits_unpack(np.array([0x00], np.uint8), 8) == -1.0)
    assert np.array_equal(od.bits_unpack(np.array([0xA5], np.uint8), 8),
                          np.array([1, -1, 1, -1, -1, 1, -1, 1], np.float32))
    src = np.array([0xDE, 0xAD, 0xBE, 0xEF], np.uint8)
    assert np.array_equal(od.bits_pack(od.bits_unpack(src, 32)), src)
    part = od.bits_pack(od.bits_unpack(np.array([0xF8], np.uint8), 5))
    assert part[0] & 0xF8 == 0xF8


def test_pack_i24_phase(od):
    rng = np.random.default_rng(2)
    bits = np.where(rng.random(16 * 5) < 0.5, 1.0, -1.0).astype(np.float32)
    out, ph = od.pack_i24(bits, 1)
    assert ph == 0 and list(out[2::3]) == [B, A, B, A, B]
    out0, ph0 = od.pack_i24(bits, 0)
    assert ph0 == 1 and np.array_equal(out0[0::3], out[0::3]) and np.array_equal(out0[1::3], out[1::3])


def test_against_reference_dop(oracle, reference):
    od, rd = ol.Dop(oracle, "ora"), ol.Dop(reference, "ref")
    rng = np.random.default_rng(3)
    for trial in range(20):
        n = int(rng.integers(1, 300))
        words = rng.integers(0, 1 << 16, n)
        pcm = np.array([dop_sample(B if i & 1 else A, int(w)) for i, w in enumerate(words)], np.float32)
        if trial % 4 == 3:
            pcm = rng.uniform(-1, 1, n).astype(np.float32)        # not DoP: unpack still defined
        assert od.detect(pcm) == rd.detect(pcm)
        assert np.array_equal(od.unpack(pcm), rd.unpack(pcm))
        bits = np.where(rng.random(n * 16 + int(rng.integers(0, 16))) < 0.5, 1.0, -1.0).astype(np.float32)
        assert np.array_equal(od.pack(bits), rd.pack(bits))
        ph = int(rng.integers(0, 2))
        a, pa = od.pack_i24(bits, ph)
        b, pb = rd.pack_i24(bits, ph)
        assert pa == pb and np.array_equal(a, b)
        nb = int(rng.integers(1, 200))
        src = rng.integers(0, 256, (nb + 7) // 8).astype(np.uint8)
        assert np.array_equal(od.bits_unpack(src, nb), rd.bits_unpack(src, nb))
        f = np.where(rng.random(nb) < 0.5, 1.0, -1.0).astype(np.float32)
        assert np.array_equal(od.bits_pack(f), rd.bits_pack(f))
