"""Live comparison of the oracle with the compiled reference trellis.c on a
seeded random matrix (runs where oracle/_ref was built from the reference
sources; elsewhere the committed fixtures carry the same check)."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as ol

RATES = {4: 2822400, 5: 2822400, 6: 5644800, 7: 11289600, 8: 11289600}


@pytest.mark.parametrize("seed", range(10))
def test_random_matrix(oracle, reference, seed):
    rng = np.random.default_rng(seed)
    K = int(rng.choice([1, 2, 3, 8, 16, 31, 32]))
    L = int(rng.choice([1, 2, 16, 33, 64, 128]))
    B = int(rng.choice([1, 4, 8, 9, 16, 32]))
    order = int(rng.integers(4, 9))
    rate = RATES[order]
    nid = 2 * (order - 4)                      # clans-<order>
    fi = reference.ntf_lookup(nid, rate)
    lim = float(rng.choice([0.0, 0.0, 2.0, 8.0]))
    n = 3000
    t = np.arange(n)
    amp = float(rng.uniform(0.05, 0.5))
    x = amp * np.sin(2 * np.pi * t * rng.uniform(1e-4, 1e-2)) + float(rng.uniform(0, 0.1)) * rng.standard_normal(n)
    a = oracle.ctx(fi, B, K, L, lim)
    b = reference.ctx(fi, B, K, L, lim)
    # feed in irregular blocks to cover the fill / continuation logic
    cuts = np.sort(rng.integers(0, n, size=4))
    parts = np.split(x, cuts)
    ya = np.concatenate([a.process(p) for p in parts])
    yb = np.concatenate([b.process(p) for p in parts])
    assert np.array_equal(ya, yb)
    assert a.counters() == b.counters()
    assert np.array_equal(a.drain(4096), b.drain(4096))


def test_bucket0_state_per_step(oracle, reference):
    """At K=32 the reference's full-list insertion writes one element past
    act[31], into the first path-hash bucket; the oracle's hash state must
    match step by step."""
    R, O = reference, oracle
    R.lib.ref_hash0_offset.argtypes = [C.c_void_p]
    R.lib.ref_hash0_offset.restype = C.c_long
    O.lib.ora_hash0_slot.argtypes = [C.c_void_p]
    O.lib.ora_hash0_slot.restype = C.c_long

    def dec(off):
        if off < 0:
            return -1
        b, r = divmod(off, 14592)
        return b * 256 + r // 112

    rng = np.random.default_rng(5)
    n = 1500
    x = 0.45 * np.sin(2 * np.pi * np.arange(n) * 997 / 11289600 * 16) + 0.05 * rng.standard_normal(n)
    fi = R.ntf_auto(11289600)
    cr, co = R.ctx(fi, 8, 32, 64), O.ctx(fi, 8, 32, 64)
    for i in range(n):
        cr.process(x[i:i + 1])
        co.process(x[i:i + 1])
        assert dec(R.lib.ref_hash0_offset(cr.h)) == O.lib.ora_hash0_slot(co.h), i


def test_estimator_live(oracle, reference):
    rng = np.random.default_rng(3)
    x = 0.2 * rng.standard_normal(5000)
    for fi in range(0, oracle.ntf_count(), 5):
        for lv in (2, 16):
            init = rng.standard_normal(10) * 0.01
            assert np.array_equal(oracle.estimate(fi, init, x, 0.0, lv),
                                  reference.estimate(fi, init, x, 0.0, lv))
