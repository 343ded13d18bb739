"""sdm_process_block / sdm_drain of include/dsd_sdm.h on the GPU, called
through the C ABI on contexts in the reference's layout, against the
compiled reference (oracle/_ref) on a twin context: every output bit and
everything the context carries into its next steps (survivors, every
node's path, metric and chain link) equal after every call, with calls to
the two implementations interleaved on the same context."""
import numpy as np
import pytest

import sdm_helpers as sh

pytestmark = pytest.mark.gpu


def _input(seed, n):
    rng = np.random.default_rng(seed)
    t = np.arange(n)
    return 0.3 * np.sin(t * 0.013) + rng.normal(0, 0.3, n)


@pytest.mark.parametrize("K,lat,lim", [(32, 64, 0.0), (32, 300, 0.0), (8, 64, 2.0), (2, 64, 0.0),
                                       (17, 128, 0.0), (1, 16, 0.0)])
def test_interleaved_calls_match_reference(gpu, reference, K, lat, lim):
    fi = reference.ntf_auto(11289600)
    print("case: K=%d L=%d limit=%g" % (K, lat, lim), flush=True)
    a = reference.ctx(fi, 8, K, lat, lim)
    b = reference.ctx(fi, 8, K, lat, lim)
    x = _input(K * 7 + lat, 12000)
    # (who runs context b, samples): starts on the GPU from a fresh context,
    # with a call shorter than the latency fill
    plan = [("gpu", lat // 2), ("gpu", 1500), ("ref", 700), ("gpu", 2500), ("ref", 300), ("gpu", 4000)]
    pos = 0
    for k, (who, n) in enumerate(plan):
        seg = x[pos:pos + n]
        pos += n
        ya = a.process(seg)
        yb = sh.process(b.h, seg) if who == "gpu" else b.process(seg)
        assert ya.size == yb.size and np.array_equal(ya, yb), (k, who)
        sh.assert_same_carried(sh.export(reference, a.h), sh.export(reference, b.h), 8, "call %d" % k)
    da = a.drain(lat // 3)
    db = sh.drain(b.h, lat // 3)
    assert np.array_equal(da, db)
    sh.assert_same_carried(sh.export(reference, a.h), sh.export(reference, b.h), 8, "drain 1")
    da = a.drain(4 * lat)
    db = sh.drain(b.h, 4 * lat)
    assert np.array_equal(da, db) and da.size > 0
    sh.assert_same_carried(sh.export(reference, a.h), sh.export(reference, b.h), 8, "drain 2")


def test_unfilled_drain_matches_reference(gpu, reference):
    fi = reference.ntf_auto(5644800)
    a = reference.ctx(fi, 6, 16, 200)
    b = reference.ctx(fi, 6, 16, 200)
    x = _input(3, 120)                               # fewer inputs than the latency
    assert a.process(x).size == sh.process(b.h, x).size == 0
    assert np.array_equal(a.drain(500), sh.drain(b.h, 500))
    sh.assert_same_carried(sh.export(reference, a.h), sh.export(reference, b.h), 6)


def test_seeded_copy_matches_reference(gpu, reference):
    """DAS seeding on the caller's contexts (src/dsp_plugin.c:2325-2338): a
    copy of the persistent context, every survivor seeded, then a segment;
    the copy keeps hash-chain links into the persistent context, which the
    GPU job follows as the reference does."""
    fi = reference.ntf_auto(11289600)
    rng = np.random.default_rng(11)
    persist = reference.ctx(fi, 8, 32, 64)
    persist.process(_input(5, 6000))
    a = reference.ctx(0, 2, 1, 1)
    b = reference.ctx(0, 2, 1, 1)
    reference._f("copy_state")(a.h, persist.h)
    sh.ours().sdm_context_copy_state(b.h, persist.h)
    st = rng.normal(0, 0.2, 8)
    a.seed(st)
    b.seed(st)
    assert (sh.export(reference, b.h, persist.h)["nlink"] >= 1000).any()
    x = _input(9, 20000)
    ya, yb = a.process(x), sh.process(b.h, x)
    assert ya.size == yb.size == x.size
    assert np.array_equal(ya, yb)
    sh.assert_same_carried(sh.export(reference, a.h, persist.h), sh.export(reference, b.h, persist.h), 8)


def test_unsupported_context_is_left_unchanged(gpu, reference):
    import ctypes as C
    fi = reference.ntf_auto(11289600)
    b = reference.ctx(fi, 8, 8, 64)
    b.process(_input(1, 500))
    before = sh.normalized(reference, b.h)
    C.c_uint.from_address(b.h + sh.MB_LEVELS_OFFSET).value = 4      # SDM_MB_LEVELS
    assert sh.process(b.h, _input(2, 1000)).size == 0
    C.c_uint.from_address(b.h + sh.MB_LEVELS_OFFSET).value = 2
    assert sh.normalized(reference, b.h) == before
