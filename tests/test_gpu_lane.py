"""K <= 2 trellis (one lane per context, 64 contexts per wave) against the
oracle: bits and final loop-filter states bit-exact over orders 4-8,
histories 1-128, path depths 1-16, the state limiter, block-wise
continuation, copy + seed, DAS with many segments (several workgroups and a
partial one), the engine's same-rate DSD re-encode settings streamed chunk
by chunk, and the one-wave-per-job kernel on the same jobs."""
import numpy as np
import pytest
import torch

import oracle_lib as ol
from foo_dsd_trellis_amd import das as D
from foo_dsd_trellis_amd import signals as SG
from foo_dsd_trellis_amd import trellis as T

pytestmark = pytest.mark.gpu


def _x(kind, n, rate, seed):
    rng = np.random.default_rng(seed)
    t = np.arange(n) / rate
    if kind == "sine":
        return 0.5012 * np.sin(2 * np.pi * 997 * 16 * t) + 1e-5 * rng.standard_normal(n)
    if kind == "pink":
        return np.clip(0.18 * SG.pink(n, rng), -0.5, 0.5)
    if kind == "loud":
        return 0.45 * np.sin(2 * np.pi * 997 * 16 * t) + 0.05 * rng.standard_normal(n)
    raise ValueError(kind)


def pick(oracle, order, rate):
    """First table filter of this order whose design rate is <= rate."""
    for i in range(oracle.ntf_count()):
        f = oracle.ntf_get(i)
        if f["order"] == order and f["freq"] <= rate:
            return i
    raise LookupError((order, rate))


LANE_CASES = [
    # K, L, B, rate, n, kind, order, state_limit
    (2, 16, 4, 11289600, 40000, "pink", 6, 0.0),
    (2, 16, 16, 2822400, 30000, "sine", 6, 0.0),
    (2, 128, 16, 11289600, 20000, "loud", 4, 0.0),
    (2, 65, 8, 5644800, 20000, "pink", 8, 0.0),
    (2, 64, 1, 5644800, 10000, "sine", 5, 0.0),
    (2, 32, 4, 22579200, 20000, "loud", 7, 1.5),
    (2, 1, 2, 2822400, 8000, "loud", 6, 0.0),
    (1, 1, 1, 2822400, 8000, "sine", 5, 0.0),
    (1, 16, 4, 11289600, 20000, "pink", 8, 0.0),
]


@pytest.mark.parametrize("K,L,B,rate,n,kind,order,lim", LANE_CASES)
def test_lane_trellis_bit_exact(gpu, oracle, K, L, B, rate, n, kind, order, lim):
    x = _x(kind, n, rate, K * 31 + L + order)
    fi = pick(oracle, order, rate)
    o = oracle.ctx(fi, B, K, L, lim)
    ref = o.process(x)
    p = T.SdmPool(fi, B, K, L, lim, 1)
    y = p.process(0, x)
    assert len(y) == len(ref)
    bad = np.nonzero(y != ref)[0]
    assert bad.size == 0, "first mismatch at %d of %d" % (bad[0], len(ref))
    assert np.array_equal(p.best_state(0)[:order], o.best_state()[:order])
    w = T.SdmPool(fi, B, K, L, lim, 1)
    w.wave_only(True)
    assert np.array_equal(w.process(0, x), ref)
    assert np.array_equal(w.best_state(0), p.best_state(0))


def test_lane_continuation_copy_seed(gpu, oracle):
    """Block-wise processing continues exactly through the stored context;
    a copied and seeded context continues like the oracle's."""
    rate = 11289600
    fi = pick(oracle, 6, rate)
    x = _x("pink", 30000, rate, 17)
    o = oracle.ctx(fi, 4, 2, 16)
    ref = o.process(x[:20000])
    p = T.SdmPool(fi, 4, 2, 16, 0.0, 2)
    ys = [p.process(0, x[a:b]) for a, b in [(0, 7), (7, 5000), (5000, 5001), (5001, 20000)]]
    assert np.array_equal(np.concatenate(ys), ref)
    o2 = oracle.ctx(fi, 4, 2, 16)
    o2.copy_from(o)
    st = o.best_state() * 0.5
    o2.seed(st)
    ref2 = o2.process(x[20000:])
    p.copy(1, 0)
    p.seed(1, st)
    assert np.array_equal(p.process(1, x[20000:]), ref2)


DAS_CASES = [
    # K, L, B, rate, F, seg_len, seg0_first, kind, order
    (2, 16, 4, 11289600, 300000, 2048, 0, "pink", 6),   # 146 segments per channel
    (2, 16, 4, 11289600, 300000, 2048, 1, "pink", 6),
    (2, 128, 16, 22579200, 200000, 16384, 0, "loud", 4),
    (1, 32, 8, 5644800, 150000, 4096, 0, "sine", 8),
]


@pytest.mark.parametrize("K,L,B,rate,F,seg_len,seg0,kind,order", DAS_CASES)
def test_lane_das_bit_exact(gpu, oracle, K, L, B, rate, F, seg_len, seg0, kind, order):
    nch = 2
    x = np.stack([_x(kind, F, rate, 40 + c) for c in range(nch)])
    fi = pick(oracle, order, rate)
    ref, n, sp, plan = ol.das(oracle, fi, B, K, L, x, seg_len=seg_len, nseg=0, seg0_first=seg0,
                              reposition=1)
    outs = []
    for wave in (False, True):
        d = D.Das(fi, B, K, L, nch, seg_len=seg_len, nseg=0, seg0_first=bool(seg0), reposition=True)
        if wave:
            d.wave_only(True)
        y, m = d.run(x)
        assert m == n
        S, slow, pos = d.last_info(len(sp))
        assert S == len(sp) and list(pos[1:]) == list(sp[1:])
        outs.append(y)
        for c in range(nch):
            bad = np.nonzero(y[c, :n] != ref[c, :n])[0]
            assert bad.size == 0, "wave_only=%s ch%d first mismatch %d" % (wave, c, bad[0])
    assert np.array_equal(outs[0][:, :n], outs[1][:, :n])


def test_lane_same_rate_reencode_stream(gpu, oracle):
    """The engine's DSD256 same-rate settings (boxcar 64, CLANS-6, K=2,
    L=16, B=4) over successive chunks with persistent contexts."""
    from foo_dsd_trellis_amd import cond, stream, trellis
    rate, nch, cf, seg = 11289600, 2, 65536, 4096
    t = np.arange(3 * cf + 1000) / rate
    rng = np.random.default_rng(5)
    src = np.stack([np.where(0.5 * np.sin(2 * np.pi * (997 + 300 * c) * t) + 0.3 * rng.standard_normal(t.size)
                             >= 0, 1.0, -1.0) for c in range(nch)]).astype(np.float32)
    fi = trellis.ntf_lookup(4, rate)
    assert oracle.ntf_get(fi)["name"] == "clans-6"
    fg = float(np.float32(0.708))
    S = stream.Stream(rate, rate, nch, cf, ntf_index=fi, K=2, L=16, depth=4, seg_len=seg, gain=fg,
                      input=stream.IN_DSD)
    conds = [ol.Cond(oracle, 64, 0) for _ in range(nch)]
    das = ol.DasStream(oracle, fi, 4, 2, 16, nch)
    n_total = src.shape[1]
    for s0 in range(0, n_total, cf):
        s1 = min(n_total, s0 + cf)
        part = torch.from_numpy(np.ascontiguousarray(src[:, s0:s1])).cuda()
        n = S.process_dsd(cond.to_bits(part), s1 - s0)
        sdm_in = np.stack([conds[c].dsd_boxcar(src[c, s0:s1], fg, fg) for c in range(nch)])
        sig = S.signal()
        assert np.array_equal(sig, sdm_in), s0
        ref = das.chunk(sig, seg, int(S.seg0_first), 1)
        assert n == ref.shape[1], s0
        assert np.array_equal(S.bits_numpy(n), ref), s0
