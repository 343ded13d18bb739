"""PreCorr on the GPU against the oracle (pinned to the reference's
precorr.c): table, bits and final states for orders 4..8, state limit,
several channels, chunked; and the seam entry point."""
import numpy as np
import pytest
import torch

import oracle_lib as ol

pytestmark = pytest.mark.gpu
RATES = {4: 2822400, 5: 2822400, 6: 5644800, 7: 11289600, 8: 11289600}


def pm1(bits, n):
    w = bits.cpu().numpy().view(np.uint64)
    b = np.unpackbits(w.view(np.uint8), axis=1, bitorder="little")[:, :n]
    return np.where(b > 0, 1.0, -1.0).astype(np.float32)


@pytest.mark.parametrize("order,lim", [(4, 0.0), (5, 12.0), (6, 0.0), (7, 0.0), (8, 0.0), (8, 12.0)])
def test_precorr_matches_oracle(gpu, oracle, order, lim):
    from foo_dsd_trellis_amd import precorr
    nch = 3
    fi = oracle.ntf_lookup(2 * (order - 4), RATES[order])
    rng = np.random.default_rng(order)
    n = 30000
    t = np.arange(n)
    x = np.stack([(0.5 + 0.1 * c) * np.sin(2 * np.pi * t * (2e-4 + 1e-4 * c)) + 0.05 * rng.standard_normal(n)
                  for c in range(nch)])
    G = precorr.PreCorr(fi, nch, lim)
    refs = [ol.PreCorr(oracle, "ora", fi, lim) for _ in range(nch)]
    assert np.array_equal(G.table(), refs[0].table())
    d_x = torch.from_numpy(x).cuda()
    for s0, s1 in [(0, 1), (1, 64), (64, 10000), (10000, 10077), (10077, n)]:
        bits = G.process(d_x[:, s0:s1])
        y = pm1(bits, s1 - s0)
        for c in range(nch):
            assert np.array_equal(y[c], refs[c].process(x[c, s0:s1])), (s0, c)
    for c in range(nch):
        st, h, ph, py = G.state(c)
        rs, rh, rph, rpy = refs[c].state()
        assert np.array_equal(st[:order], rs[:order]) and (h, ph, py) == (rh, rph, rpy)


def test_precorr_seam(gpu, oracle):
    from foo_dsd_trellis_amd import precorr
    fi = oracle.ntf_auto(11289600)
    ref = ol.PreCorr(oracle, "ora", fi)
    r = oracle.ntf_get(fi)
    n = 20000
    x = (0.4 * np.sin(2 * np.pi * np.arange(n) * 5e-4)).astype(np.float32)
    y, fin = precorr.seam_process(x, r["a"][:r["order"]], r["g"][:r["order"]], ref.table())
    assert np.array_equal(y, ref.process(x.astype(np.float64)))
    st, h, ph, py = ref.state()
    assert (fin.history, fin.phase, fin.prev_y) == (h, ph, py)
