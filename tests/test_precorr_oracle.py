"""PreCorr in the oracle (or_precorr.c) against the reference's own
src/precorr.c compiled into oracle/_ref: training table, outputs and final
state, orders 4..8, with and without the state limit, chunked."""
import numpy as np
import pytest

import oracle_lib as ol

RATES = {4: 2822400, 5: 2822400, 6: 5644800, 7: 11289600, 8: 11289600}


def signal(n, seed, amp):
    rng = np.random.default_rng(seed)
    t = np.arange(n)
    return amp * np.sin(2 * np.pi * t * 3e-4) + 0.05 * rng.standard_normal(n)


@pytest.mark.parametrize("order", [4, 5, 6, 7, 8])
@pytest.mark.parametrize("lim", [0.0, 12.0])
def test_precorr_matches_reference(oracle, reference, order, lim):
    fi = reference.ntf_lookup(2 * (order - 4), RATES[order])
    a = ol.PreCorr(oracle, "ora", fi, lim)
    b = ol.PreCorr(reference, "ref", fi, lim)
    assert np.array_equal(a.table(), b.table())
    x = signal(20000, order, 0.6 if lim else 0.4)
    for part in np.split(x, [1, 4000, 4063, 15000]):
        assert np.array_equal(a.process(part), b.process(part))
    sa, sb = a.state(), b.state()
    assert np.array_equal(sa[0][:order], sb[0][:order]) and sa[1:] == sb[1:]


def test_precorr_properties(oracle):
    fi = oracle.ntf_auto(11289600)
    c = ol.PreCorr(oracle, "ora", fi)
    t = c.table()
    assert t.shape == (2048,) and np.all(np.isfinite(t)) and np.any(t != 0.0)
    y = c.process(signal(5000, 1, 0.4))
    assert y.size == 5000 and set(np.unique(y)) <= {-1.0, 1.0}          # no latency, +-1
    # density follows the input's sign over long stretches
    yp = c.process(np.full(20000, 0.3))
    assert 0.55 < np.mean(yp > 0) < 0.75
