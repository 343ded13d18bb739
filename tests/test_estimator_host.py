"""The DAS seeding estimate on host threads (dsdhip_estimate_boundaries in
libdsdhip.so, pure host code): bit-exact with the reference's own
sdm_estimate_state_multibit (src/trellis.c:1118-1187, compiled into
oracle/_ref where the sources are present) and with the oracle, chained over
boundaries as the estimation block of src/threadpool.c:345-366 does.  Runs
without a GPU."""
import math

import numpy as np
import pytest

import oracle_lib as ol


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def ref_idx(v, a0, levels):
    """The reference's quantizer index (src/trellis.c:1162-1168) on Python
    floats (IEEE doubles), with x86 cvttsd2si's INT_MIN for NaN / out of range."""
    step = 2.0 / (levels - 1)
    w = (v / a0 + 1.0) / step + 0.5
    if not (w < 2147483648.0) or not (w > -2147483649.0):
        i = -2 ** 31
    else:
        i = int(w)
    return min(max(i, 0), levels - 1)


def _key(v):
    u = int(np.array(v, np.float64).view(np.uint64))
    return (~u & 0xFFFFFFFFFFFFFFFF) if u >> 63 else u | (1 << 63)


def _unkey(k):
    u = (k & ~(1 << 63)) if k >> 63 else (~k & 0xFFFFFFFFFFFFFFFF)
    return float(np.array(u, np.uint64).view(np.float64))


def first_at_least(k, a0, levels, lo, hi):
    """Smallest double v in (lo, hi] with ref_idx(v) >= k, by bisection over
    the ordered doubles (ref_idx(lo) < k <= ref_idx(hi))."""
    a, b = _key(lo), _key(hi)
    while b - a > 1:
        m = (a + b) // 2
        if ref_idx(_unkey(m), a0, levels) >= k:
            b = m
        else:
            a = m
    return _unkey(b)


def chained(lib_, fi, init, x, bnd, lim, levels):
    """The threadpool estimation loop over a single-shot estimator."""
    st = np.array(init, np.float64)
    pos, out = 0, []
    for b in bnd:
        if pos < b <= x.size:
            st = lib_.estimate(fi, st, x[pos:b], lim, levels)
            pos = b
        out.append(st.copy())
    return np.array(out)


def _inputs(rng, n, kind):
    t = np.arange(n)
    if kind == "music":
        return 0.3 * np.sin(2 * np.pi * t * 1e-4) + 0.1 * np.sin(t * 0.0031) + 0.05 * rng.standard_normal(n)
    if kind == "loud":
        return np.clip(0.9 * rng.standard_normal(n), -2, 2)
    if kind == "dc":
        return np.full(n, 0.25)
    if kind == "zero":
        return np.zeros(n)
    raise ValueError(kind)


@pytest.mark.parametrize("levels", [16, 2, 5, 17])
@pytest.mark.parametrize("kind", ["music", "loud", "dc", "zero"])
def test_boundaries_match_reference_and_oracle(oracle, kind, levels):
    from foo_dsd_trellis_amd import das
    ref = ol.reference()
    rng = np.random.default_rng(hash((kind, levels)) & 0xffff)
    for fi in (oracle.ntf_auto(r) for r in (2822400, 5644800, 11289600, 24576000)):
        for lim in (0.0, 1.5):
            n = 20000
            x = _inputs(rng, n, kind)
            init = rng.standard_normal(10) * 0.01
            bnd = np.array([0, 1, 1, 999, 1024, 1025, 4097, 12000, 12000, 19999, n, n + 5], np.int64)
            P = oracle.ntf_get(fi)["order"]      # states beyond the order carry nothing
            got, _ = das.estimate_boundaries(fi, init, x, bnd, lim, levels)
            want = chained(oracle, fi, init, x, bnd, lim, levels)
            assert np.array_equal(_bits(got[:, :P]), _bits(want[:, :P])), (fi, lim)
            if ref is not None:
                want = chained(ref, fi, init, x, bnd, lim, levels)
                assert np.array_equal(_bits(got[:, :P]), _bits(want[:, :P])), (fi, lim)


def test_index_guess_does_not_change_results(oracle):
    """A biased index guess sends blocks to the reference form; the states
    stay the reference's."""
    from foo_dsd_trellis_amd import das
    rng = np.random.default_rng(9)
    fi = oracle.ntf_auto(11289600)
    x = _inputs(rng, 30000, "music")
    bnd = np.arange(0, 30001, 3000)
    base, redo0 = das.estimate_boundaries(fi, None, x, bnd)
    assert redo0 == 0
    for bias in (0.5, -0.7, 3.0):
        got, redo = das.estimate_boundaries(fi, None, x, bnd, index_bias=bias)
        assert redo > 0
        assert np.array_equal(_bits(got), _bits(base))
    assert np.array_equal(_bits(base), _bits(chained(oracle, fi, np.zeros(10), x, bnd, 0.0, 16)))


def test_extreme_inputs(oracle):
    """Overflowing quantizer input (the int conversion's INT_MIN, idx 0),
    infinities and NaN follow the reference's semantics."""
    from foo_dsd_trellis_amd import das
    fi = oracle.ntf_auto(11289600)
    for x in (np.full(50, 1e12), np.concatenate([np.zeros(10), [np.inf], np.zeros(10)]),
              np.concatenate([np.full(5, 0.1), [np.nan], np.full(5, 0.1)]), np.full(40, -3e9)):
        bnd = np.array([x.size // 2, x.size])
        got, _ = das.estimate_boundaries(fi, None, x, bnd)
        assert np.array_equal(_bits(got), _bits(chained(oracle, fi, np.zeros(10), x, bnd, 0.0, 16)))


@pytest.mark.parametrize("levels", [3, 4, 16, 17, 32])
def test_interval_table_is_the_reference_quantizer(oracle, levels):
    """The exact interval table equals the reference's index formula at
    every transition point and at the doubles either side of it, at the
    overflow point, and on random values."""
    from foo_dsd_trellis_amd import das
    rng = np.random.default_rng(levels)
    for r in (2822400, 11289600, 24576000):
        a0 = float(oracle.ntf_get(oracle.ntf_auto(r))["a"][0])
        step = 2.0 / (levels - 1)
        vals = list(rng.uniform(-3 * a0, 3 * a0, 2000)) + [0.0, -0.0, math.inf, -math.inf, math.nan, 1e300,
                                                           -1e300, 5e-324]
        for k in range(1, levels):
            v = first_at_least(k, a0, levels, -2.0 * a0, 2.0 * a0)   # the exact transition point
            vals += [float(np.nextafter(v, -math.inf)), v, float(np.nextafter(v, math.inf))]
        big = a0 * (2 ** 31 * step)          # where the int conversion overflows
        v = big
        for _ in range(5):
            vals += [float(v), float(np.nextafter(v, math.inf)), float(np.nextafter(v, -math.inf))]
            v = v * (1 + 1e-12)
        for v in vals:
            assert das.table_index(a0, levels, v) == ref_idx(v, a0, levels), (a0, levels, v)


@pytest.mark.parametrize("bias", [0.0, 3e-5, -3e-5, 0.5])
def test_two_thread_chain_matches_reference(oracle, bias):
    """Long chains run the index guess on a second thread and check it on
    the calling one; rare wrong guesses (a small bias) restart the guesser
    from the checked state, frequent ones (a large bias) end in the
    single-thread form.  Every boundary state stays the reference's."""
    from foo_dsd_trellis_amd import das
    rng = np.random.default_rng(21)
    fi = oracle.ntf_auto(11289600)
    n = (1 << 20) + 12345
    x = _inputs(rng, n, "music")
    x[700001] = np.nan                               # one non-finite input mid-chain
    x[700002:700100] = 1e12
    bnd = np.array([0, 5, 65536, 65537, 300000, 524287, 524288, 700000, 700050, 900001, n, n + 1], np.int64)
    init = rng.standard_normal(10) * 0.01
    got, redo = das.estimate_boundaries(fi, init, x, bnd, index_bias=bias)
    want = chained(oracle, fi, init, x, bnd, 0.0, 16)
    assert np.array_equal(_bits(got[:, :8]), _bits(want[:, :8]))
    if bias != 0.0:
        assert redo > 0
