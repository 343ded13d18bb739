"""The plain CPU-runnable configuration end to end on the GPU: 1 s of
stereo 44.1 kHz float32 PCM (997 Hz sine at -6 dBFS plus seeded Gaussian
noise at -100 dBFS) -> six 2x 63-tap half-band stages to DSD64 -> trellis
with the default NTF for DSD64, K = 8, L = 16, B = 8, one segment per
channel.  The FIR output equals the oracle's doubles; the DSD bits equal
those of the reference's own trellis.c (compiled into oracle/_ref) run on
that FIR output, and the oracle's."""
import numpy as np
import pytest

import oracle_lib as ol
from foo_dsd_trellis_amd import fir as FIR
from foo_dsd_trellis_amd import trellis as T

pytestmark = pytest.mark.gpu


def test_c1_pcm_to_dsd64_bit_exact(gpu, oracle, reference):
    fs, fs_out, n, nch = 44100, 2822400, 44100, 2
    rng = np.random.default_rng(1)
    t = np.arange(n) / fs
    x = np.stack([0.5012 * np.sin(2 * np.pi * 997 * t) + 1e-5 * rng.standard_normal(n)
                  for _ in range(nch)]).astype(np.float32)
    f = FIR.FirChain(fs, fs_out, FIR.PCM_FIR_GAIN, nch=nch)
    y = f.process(x)
    assert y.shape == (nch, n * (fs_out // fs))
    fi = oracle.ntf_auto(fs_out)
    assert oracle.ntf_get(fi)["name"] == "clans-5"
    pool = T.SdmPool(fi, 8, 8, 16, 0.0, nch)
    for c in range(nch):
        assert np.array_equal(y[c], ol.fir_chain(oracle, fs, fs_out, x[c], FIR.PCM_FIR_GAIN))
        bits = pool.process(c, y[c])
        ref = reference.ctx(fi, 8, 8, 16).process(y[c])
        assert len(bits) == len(ref) == y.shape[1] - 16
        bad = np.nonzero(bits != ref)[0]
        assert bad.size == 0, "ch%d first mismatch at %d" % (c, bad[0])
        assert np.array_equal(ref, oracle.ctx(fi, 8, 8, 16).process(y[c]))
