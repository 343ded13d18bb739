"""The engine's pre-SDM conditioning on the GPU against the oracle
(or_cond.c), bit for bit: DSD boxcar at the engine's tap counts and the
lowpass over chunks with gain ramps, the sequential PCM boxcar, ramp and
soft clip."""
import numpy as np
import pytest
import torch

import oracle_lib as ol

pytestmark = pytest.mark.gpu

SIZES = [1000, 1, 63, 4096, 777]
GAINS = [(0.708, 0.708), (0.354, 0.9), (0.9, 0.2), (1.0, 1.0), (0.5, 0.5)]


def dsd_input(seed, nch, n):
    rng = np.random.default_rng(seed)
    return np.where(rng.random((nch, n)) < 0.55, 1.0, -1.0).astype(np.float32)


@pytest.mark.parametrize("taps", [8, 16, 32, 64])
def test_dsd_boxcar(gpu, oracle, taps):
    from foo_dsd_trellis_amd import cond
    nch = 3
    x = dsd_input(taps, nch, sum(SIZES))
    g = cond.Conditioner(nch, taps)
    refs = [ol.Cond(oracle, taps) for _ in range(nch)]
    pos = 0
    for k, n in enumerate(SIZES):
        gs, ge = GAINS[k]
        part = torch.from_numpy(np.ascontiguousarray(x[:, pos:pos + n])).cuda()
        y = g.dsd_boxcar(cond.to_bits(part), n, gs, ge).cpu().numpy()
        for c in range(nch):
            assert np.array_equal(y[c], refs[c].dsd_boxcar(x[c, pos:pos + n], gs, ge)), (k, c)
        pos += n


def test_dsd_lowpass(gpu, oracle):
    from foo_dsd_trellis_amd import cond
    nch, rate = 2, 5644800
    assert np.array_equal(cond.lowpass_taps(rate), ol.lowpass_taps(oracle, rate))
    x = dsd_input(9, nch, sum(SIZES))
    g = cond.Conditioner(nch, 8, rate)
    refs = [ol.Cond(oracle, 8, rate) for _ in range(nch)]
    pos = 0
    for k, n in enumerate(SIZES):
        gs, ge = GAINS[k]
        part = torch.from_numpy(np.ascontiguousarray(x[:, pos:pos + n])).cuda()
        y = g.dsd_lowpass(cond.to_bits(part), n, gs, ge).cpu().numpy()
        for c in range(nch):
            assert np.array_equal(y[c], refs[c].dsd_lowpass(x[c, pos:pos + n], gs, ge)), (k, c)
        pos += n


def test_pcm_boxcar(gpu, oracle):
    from foo_dsd_trellis_amd import cond
    nch, taps = 2, 8
    rng = np.random.default_rng(4)
    x = (0.3 * rng.standard_normal((nch, sum(SIZES) + 50000))).astype(np.float32)
    x[:, 100:300] *= 1e-7
    g = cond.Conditioner(nch, taps)
    refs = [ol.Cond(oracle, taps) for _ in range(nch)]
    pos = 0
    for n in SIZES + [50000]:
        part = torch.from_numpy(np.ascontiguousarray(x[:, pos:pos + n])).cuda()
        y = g.pcm_boxcar(part).cpu().numpy()
        for c in range(nch):
            assert np.array_equal(y[c], refs[c].pcm_boxcar(x[c, pos:pos + n])), (n, c)
        pos += n


def test_ramp_and_clip(gpu, oracle):
    from foo_dsd_trellis_amd import cond
    rng = np.random.default_rng(5)
    x = rng.uniform(-1.6, 1.6, (2, 10000))
    y = cond.gain_ramp(torch.from_numpy(x.copy()).cuda(), 0.3, 0.9).cpu().numpy()
    for c in range(2):
        assert np.array_equal(y[c], ol.gain_ramp(oracle, x[c], 0.3, 0.9))
    z = cond.soft_clip(torch.from_numpy(x.copy()).cuda()).cpu().numpy()
    for c in range(2):
        assert np.array_equal(z[c], ol.soft_clip(oracle, x[c]))
