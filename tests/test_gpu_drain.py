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


@pytest.mark.parametrize("levels,lim", [(16, 0.0), (2, 0.0), (16, 1.5), (5, 0.0)])
def test_estimate_state_matches_oracle(gpu, oracle, levels, lim):
    """sdm_estimate_state(_multibit) on the device (dsdhip_estimate_state)
    against the oracle, which the compiled reference pins."""
    import ctypes as C
    from foo_dsd_trellis_amd._lib import lib
    rate = 11289600
    fi = oracle.ntf_auto(rate)
    f = oracle.ntf_get(fi)
    P = f["order"]
    rng = np.random.default_rng(levels)
    x = np.clip(0.2 * rng.standard_normal(50000), -0.5, 0.5)
    init = 0.01 * rng.standard_normal(10)
    init[P:] = 0.0
    want = oracle.estimate(fi, init, x, lim, levels)
    out = np.zeros(10)
    dp = C.POINTER(C.c_double)
    a = np.ascontiguousarray(f["a"][:10])
    g = np.ascontiguousarray(f["g"][:10])
    rc = lib().dsdhip_estimate_state(a.ctypes.data_as(dp), g.ctypes.data_as(dp), P, init.ctypes.data_as(dp),
                                     x.ctypes.data_as(dp), x.size, lim, levels, out.ctypes.data_as(dp))
    assert rc == 0
    assert np.array_equal(out[:P], want[:P])


def test_state_distance(gpu, oracle):
    import ctypes as C
    from foo_dsd_trellis_amd._lib import lib
    rate = 5644800
    fi = oracle.ntf_auto(rate)
    P = oracle.ntf_get(fi)["order"]
    rng = np.random.default_rng(2)
    x = 0.3 * rng.standard_normal(8000)
    p = T.SdmPool(fi, 8, 8, 16, 0.0, 2)
    p.process(0, x[:5000])
    p.process(1, x[1000:6000])
    d = C.c_double(0)
    assert lib().dsdhip_trellis_state_distance(p.h, 0, 1, C.byref(d)) == 0
    sa, sb = p.best_state(0)[:P], p.best_state(1)[:P]
    want = 0.0
    for k in range(P):
        want = want + (sa[k] - sb[k]) * (sa[k] - sb[k])
    assert d.value == want
