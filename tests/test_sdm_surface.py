"""The sdm_* / ntf_* surface (include/dsd_sdm.h) on the CPU: the context
layout against the compiled reference's, the host-side entry points
(init, reset, free, copy, distance, estimator, NTF table) byte for byte
against the reference's own, and the conversion to the device form and
back (what sdm_process_block does around its GPU job) shown to keep
everything the trellis reads again: a context that went through it
continues exactly as the untouched one under the reference's own
sdm_process_block, including stale hash chains at K = 32 and chains into
a copy source."""
import ctypes as C
import re

import numpy as np
import pytest

import oracle_lib as ol
import sdm_helpers as sh

_dp = C.POINTER(C.c_double)


def _declared():
    txt = open(sh.ROOT + "/include/dsd_sdm.h").read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b((?:sdm|ntf|dsdhip)_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_sdm_header():
    L = C.CDLL(sh.SO)
    names = _declared()
    assert len(names) >= 21
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


class _State(C.Structure):
    pass


_State._fields_ = [("state", C.c_double * 10), ("cost", C.c_double), ("path", C.c_uint32),
                   ("next", C.c_uint8), ("hist", C.c_uint8), ("hist_used", C.c_uint8),
                   ("parent", C.c_void_p), ("path_list", C.c_void_p)]


class _Trellis(C.Structure):
    _fields_ = [("sdm", _State * 128), ("act", C.c_void_p * 32)]


class _Ctx(C.Structure):
    """include/dsd_sdm.h's sdm_context_t, restated for the offsets."""
    _fields_ = [("trellis", _Trellis * 2), ("path_hash", C.c_void_p * 128), ("hist_free", C.c_uint8 * 64),
                ("hist_fnum", C.c_uint), ("trellis_mask", C.c_uint32), ("trellis_num", C.c_uint32),
                ("trellis_lat", C.c_uint32), ("num_cands", C.c_uint), ("mb_levels", C.c_uint),
                ("pos", C.c_uint), ("pending", C.c_uint), ("draining", C.c_uint), ("idx", C.c_uint),
                ("filter", C.c_void_p), ("prev_y", C.c_double), ("conv_fail", C.c_uint64),
                ("cands_collapse", C.c_uint64), ("next_filter_drops", C.c_uint64),
                ("total_children", C.c_uint64), ("state_limit", C.c_double), ("hist", (C.c_uint8 * 256) * 64)]


def test_layout_matches_reference(reference):
    lay = (C.c_long * 22)()
    reference.lib.ref_layout(lay)
    want = [C.sizeof(_State), C.sizeof(_Trellis), C.sizeof(_Ctx), _Trellis.act.offset,
            _Ctx.path_hash.offset, _Ctx.hist_free.offset, _Ctx.hist_fnum.offset, _Ctx.trellis_mask.offset,
            _Ctx.num_cands.offset, _Ctx.idx.offset, _Ctx.filter.offset, _Ctx.prev_y.offset,
            _Ctx.conv_fail.offset, _Ctx.total_children.offset, _Ctx.state_limit.offset, _Ctx.hist.offset,
            _State.cost.offset, _State.path.offset, _State.next.offset, _State.hist.offset,
            _State.parent.offset, _State.path_list.offset]
    assert list(lay) == want
    assert want[2] == 46752


def _filter_ptr(reference, fi):
    """the reference's own filter record (its context init stores it)"""
    h = reference._f("ctx_new")(fi, 2, 1, 1, 0.0)
    p = C.c_void_p.from_address(h + _Ctx.filter.offset).value
    reference._f("ctx_free")(h)
    return p


def _blank(reference):
    return reference.ctx(0, 2, 1, 1)


def test_init_reset_free_match_reference_bytes(reference):
    L = sh.ours()
    fi = reference.ntf_auto(11289600)
    f = _filter_ptr(reference, fi)
    for depth, K, lat in [(8, 32, 64), (4, 3, 100), (32, 1, 1), (16, 32, 2048)]:
        a = reference.ctx(fi, depth, K, lat)
        b = _blank(reference)
        assert L.sdm_context_init(b.h, f, depth, K, lat) == 0
        assert sh.normalized(reference, a.h) == sh.normalized(reference, b.h), (depth, K, lat)
        x = np.random.default_rng(depth).normal(0, 0.3, 700)
        a.process(x)
        b.process(x)
        a.reset()
        L.sdm_context_reset(b.h)
        assert sh.normalized(reference, a.h) == sh.normalized(reference, b.h), ("reset", depth, K, lat)
    b = _blank(reference)
    for bad in [(0, 8, 64), (33, 8, 64), (8, 0, 64), (8, 33, 64), (8, 8, 0), (8, 8, 2049)]:
        assert L.sdm_context_init(b.h, f, *bad) == -1, bad
    assert L.sdm_context_init(b.h, None, 8, 8, 64) == -1
    L.sdm_context_free(b.h)
    assert sh.raw(reference, b.h) == bytes(46752)


def test_copy_state_and_distance_match_reference(reference):
    L = sh.ours()
    fi = reference.ntf_auto(11289600)
    src = reference.ctx(fi, 8, 32, 64)
    src.process(np.random.default_rng(1).normal(0, 0.35, 3000))
    a, b = _blank(reference), _blank(reference)
    reference._f("copy_state")(a.h, src.h)
    L.sdm_context_copy_state(b.h, src.h)
    assert sh.normalized(reference, a.h) == sh.normalized(reference, b.h)
    other = reference.ctx(fi, 8, 32, 64)
    other.process(np.random.default_rng(2).normal(0, 0.35, 2000))
    d = L.sdm_state_distance(src.h, other.h)
    sa, so = src.best_state()[:8], other.best_state()[:8]
    want = 0.0
    for k in range(8):                            # left-to-right sum of squares
        want = want + (sa[k] - so[k]) * (sa[k] - so[k])
    assert d == want


def test_estimator_and_ntf_table_match_reference(reference):
    L = sh.ours()
    rng = np.random.default_rng(5)
    for i in range(reference.ntf_count()):
        p = L.ntf_get_by_index(i)
        a = np.frombuffer(C.string_at(p, 80), dtype=np.float64)
        g = np.frombuffer(C.string_at(p + 80, 80), dtype=np.float64)
        r = reference.ntf_get(i)
        assert np.array_equal(a, r["a"]) and np.array_equal(g, r["g"])
        assert C.c_int.from_address(p + 160).value == r["order"]
        assert C.string_at(C.c_void_p.from_address(p + 168).value).decode() == r["name"]
    for rate in (2822400, 5644800, 11289600, 22579200):
        assert L.ntf_auto_select(rate) == L.ntf_get_by_index(reference.ntf_auto(rate))
        for ntf_id in range(10):
            k = reference.ntf_lookup(ntf_id, rate)
            assert L.ntf_get_filter(ntf_id, rate) == (L.ntf_get_by_index(k) if k >= 0 else None)
    assert L.ntf_auto_select_precorr(2822400) == L.ntf_get_filter(4, 2822400)      # clans-6
    assert L.ntf_auto_select_precorr(11289600) == L.ntf_get_filter(6, 11289600)    # clans-7
    fi = reference.ntf_auto(11289600)
    f = L.ntf_get_by_index(fi)
    x = rng.normal(0, 0.3, 20000)
    init = np.concatenate([rng.normal(0, 0.1, 8), np.zeros(2)])
    for levels, lim in [(16, 0.0), (2, 0.0), (16, 1.5), (9, 0.0)]:
        out = np.zeros(10)
        L.sdm_estimate_state_multibit(f, init.ctypes.data_as(_dp), x.ctypes.data_as(_dp), x.size, lim, levels,
                                      out.ctypes.data_as(_dp))
        want = reference.estimate(fi, init, x, lim, levels)
        assert np.array_equal(out[:8], want[:8]), (levels, lim)


def _twin(reference, fi, depth, K, lat, lim, x0):
    a = reference.ctx(fi, depth, K, lat, lim)
    b = reference.ctx(fi, depth, K, lat, lim)
    a.process(x0)
    b.process(x0)
    return a, b


@pytest.mark.parametrize("K,lat,lim", [(32, 64, 0.0), (32, 300, 0.0), (8, 64, 2.0), (2, 64, 0.0), (17, 128, 0.0)])
def test_device_form_roundtrip_keeps_what_the_trellis_reads(reference, K, lat, lim):
    fi = reference.ntf_auto(11289600)
    rng = np.random.default_rng(K * 1000 + lat)
    x = rng.normal(0, 0.35, 6000) + 0.3 * np.sin(np.arange(6000) * 0.01)
    a, b = _twin(reference, fi, 8, K, lat, lim, x[:2500])
    assert sh.ours().dsdhip_sdm_roundtrip(b.h) == 0
    sh.assert_same_carried(sh.export(reference, a.h), sh.export(reference, b.h), 8, "after roundtrip")
    ya, yb = a.process(x[2500:]), b.process(x[2500:])
    assert np.array_equal(ya, yb)
    sh.assert_same_carried(sh.export(reference, a.h), sh.export(reference, b.h), 8, "continued")
    # roundtrips between every few hundred steps
    for k in range(6):
        seg = rng.normal(0, 0.35, 300 + 37 * k)
        assert sh.ours().dsdhip_sdm_roundtrip(b.h) == 0
        assert np.array_equal(a.process(seg), b.process(seg)), k
    assert np.array_equal(a.drain(), b.drain())


def test_roundtrip_follows_chains_into_the_copy_source(reference):
    """DAS seeding: a copy of the persistent context (its chain links keep
    pointing into the persistent one), every survivor seeded; the copy made
    by this library's sdm_context_copy_state, which registers the source."""
    fi = reference.ntf_auto(11289600)
    rng = np.random.default_rng(7)
    persist = reference.ctx(fi, 8, 32, 64)
    persist.process(rng.normal(0, 0.35, 4000))
    a, b = _blank(reference), _blank(reference)
    reference._f("copy_state")(a.h, persist.h)
    sh.ours().sdm_context_copy_state(b.h, persist.h)
    seed = rng.normal(0, 0.2, 8)
    a.seed(seed)
    b.seed(seed)
    ea = sh.export(reference, a.h, persist.h)
    assert (ea["nlink"] >= 1000).any()              # links into the source are present
    assert sh.ours().dsdhip_sdm_roundtrip(b.h) == 0
    sh.assert_same_carried(ea, sh.export(reference, b.h, persist.h), 8, "copy roundtrip")
    x = rng.normal(0, 0.35, 5000)
    assert np.array_equal(a.process(x), b.process(x))
    sh.assert_same_carried(sh.export(reference, a.h, persist.h), sh.export(reference, b.h, persist.h), 8)


def test_unregistered_outside_links_are_refused(reference):
    """A copy this library did not make (the reference's own copy) has links
    into memory it does not know: the conversion refuses rather than read
    through them."""
    fi = reference.ntf_auto(11289600)
    persist = reference.ctx(fi, 8, 32, 64)
    persist.process(np.random.default_rng(3).normal(0, 0.35, 3000))
    c = _blank(reference)
    reference._f("copy_state")(c.h, persist.h)
    sh.ours().sdm_context_free(c.h)                  # drops any registration
    reference._f("copy_state")(c.h, persist.h)
    e = sh.export(reference, c.h, persist.h)
    if (e["nlink"] >= 1000).any():
        assert sh.ours().dsdhip_sdm_roundtrip(c.h) == -1
