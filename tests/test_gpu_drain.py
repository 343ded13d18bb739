"""sdm_drain on the device (dsdhip_trellis_drain) against the oracle's
restatement of src/trellis.c:1042-1062: a filled context, an unfilled one
(the latency flush on zeros), partial drains, a drain without processing,
and processing again after a drain, on the one-wave (K = 8) and the
one-lane (K = 2) kernels."""
import numpy as np
import pytest

from foo_dsd_trellis_amd import trellis as T

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("K,L,B", [(8, 16, 8), (2, 16, 4), (32, 64, 8)])
@pytest.mark.parametrize("case", ["filled", "unfilled", "partial", "empty"])
def test_drain_matches_oracle(gpu, oracle, K, L, B, case):
    rate = 5644800
    fi = oracle.ntf_auto(rate)
    rng = np.random.default_rng(K * 7 + L)
    x = 0.3 * np.sin(2 * np.pi * 997 * 8 * np.arange(6000) / rate) + 0.01 * rng.standard_normal(6000)
    o = oracle.ctx(fi, B, K, L)
    p = T.SdmPool(fi, B, K, L, 0.0, 1)
    n_in = {"filled": 5000, "unfilled": 5, "partial": 3000, "empty": 0}[case]
    if n_in:
        assert np.array_equal(p.process(0, x[:n_in]), o.process(x[:n_in]))
    if case == "partial":
        for m in (7, 100):
            assert np.array_equal(p.drain(0, m), o.drain(m))
    else:
        assert np.array_equal(p.drain(0, 100), o.drain(100))
    # processing continues from the drained state
    assert np.array_equal(p.process(0, x[n_in:n_in + 500]), o.process(x[n_in:n_in + 500]))
    assert np.array_equal(p.best_state(0), o.best_state())
