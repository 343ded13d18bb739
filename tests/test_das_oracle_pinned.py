"""Pins the oracle's DAS stitch, reposition, seeding and assembly
(src/dsp_plugin.c:1963-2034, 2185-2363, 2411-2724) by cases whose answers
follow from their construction, by independent numpy computations, and by
the reference's own trellis.c (compiled into oracle/_ref) for the segment
contexts.  The reference's DAS driver itself needs Win32 + IPP and does not
build here.  CPU only."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_lib as ol

_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lp = C.POINTER(C.c_long)


def detail(oracle, prev, cur, lat):
    L = oracle.lib
    L.ora_stitch_detail.argtypes = [_fp, _fp, C.c_int, C.c_int, _ip, _ip, _dp, _ip]
    prev = np.ascontiguousarray(prev, np.float32)
    cur = np.ascontiguousarray(cur, np.float32)
    ints = np.zeros(6, np.int32)
    ci = np.zeros(99, np.int32)
    cd = np.zeros(66)
    dm = np.zeros(prev.size, np.int32)
    pos = L.ora_stitch_detail(prev.ctypes.data_as(_fp), cur.ctypes.data_as(_fp), prev.size, lat,
                              ints.ctypes.data_as(_ip), ci.ctypes.data_as(_ip), cd.ctypes.data_as(_dp),
                              dm.ctypes.data_as(_ip))
    n = int(ints[4])
    cands = [dict(pos=int(ci[i]), run=int(ci[33 + i]), density=int(ci[66 + i]), quality=float(cd[i]),
                  max_decoded=float(cd[33 + i])) for i in range(n)]
    return dict(pos=pos, half_w=int(ints[0]), peak=int(ints[1]), peak_pos=int(ints[2]),
                scanned=int(ints[3]), best=int(ints[5]), cands=cands, dm=dm)


def streams(ovl, runs, seed=1):
    """prev random +-1; cur equal to prev exactly on the given [a, b) runs"""
    rng = np.random.default_rng(seed)
    prev = np.where(rng.random(ovl) > 0.5, 1.0, -1.0).astype(np.float32)
    cur = -prev
    for a, b in runs:
        cur[a:b] = prev[a:b]
    return prev, cur


# ---- density map: the windowed match count over [p - hw, p + hw - 1] ----

@pytest.mark.parametrize("ovl,lat", [(4096, 64), (4096, 16), (100, 64), (6, 64), (1, 16), (5000, 128)])
def test_density_map_is_windowed_match_count(oracle, ovl, lat):
    rng = np.random.default_rng(ovl + lat)
    prev = np.where(rng.random(ovl) > 0.5, 1.0, -1.0).astype(np.float32)
    cur = np.where(rng.random(ovl) > 0.3, prev, -prev).astype(np.float32)
    d = detail(oracle, prev, cur, lat)
    hw = max(min(lat, ovl // 2), 4)
    assert d["half_w"] == hw
    m = np.concatenate([[0], np.cumsum(prev == cur)])
    p = np.arange(ovl)
    lo = np.clip(p - hw, 0, ovl)
    hi = np.clip(p + hw, 0, ovl)          # window [p - hw, p + hw - 1]
    want = m[hi] - m[lo]
    assert np.array_equal(d["dm"], want)
    assert d["peak"] == want.max() and d["peak_pos"] == int(np.argmax(want))


# ---- candidate list on constructed overlaps (hw = 64, peak 128, floor 115) ----

def test_single_run_single_candidate(oracle):
    """One run [500, 700): the first position whose density reaches the
    floor 115 is 551 (density p - 436); the peak run (at the first full
    window, 564) lies within hw of it, so it is the only candidate and is
    taken without a quality comparison."""
    prev, cur = streams(4096, [(500, 700)])
    d = detail(oracle, prev, cur, 64)
    assert d["peak"] == 128 and d["peak_pos"] == 564
    assert d["scanned"] == 1 and len(d["cands"]) == 1
    assert d["cands"][0]["pos"] == 551 and d["cands"][0]["run"] == 149 and d["cands"][0]["density"] == 115
    assert d["cands"][0]["quality"] == 0.0 and d["pos"] == 551


def test_two_runs_two_candidates(oracle):
    prev, cur = streams(4096, [(500, 700), (2000, 2300)])
    d = detail(oracle, prev, cur, 64)
    got = [(c["pos"], c["run"], c["density"]) for c in d["cands"]]
    assert got == [(551, 149, 115), (2051, 249, 115)]
    assert d["scanned"] == 2


def test_min_spacing_and_peak_run_candidate(oracle):
    """Runs [1000, 1100) and [1105, 1400): the first qualifying position is
    1056 (100 + (p - 1041) matches, run 44); 1105 qualifies (density 123)
    but lies 49 < hw after it and is dropped; the peak (128, first at 1169)
    is 113 >= hw away, so its run of 8 is added last."""
    prev, cur = streams(4096, [(1000, 1100), (1105, 1400)])
    d = detail(oracle, prev, cur, 64)
    assert d["peak"] == 128 and d["peak_pos"] == 1169
    got = [(c["pos"], c["run"], c["density"]) for c in d["cands"]]
    assert got == [(1056, 44, 115), (1169, 8, 128)]
    assert d["scanned"] == 1


def test_peak_run_fallback_takes_last_single_match(oracle):
    """No run of 4 anywhere (pairs of matches every 3 samples): the scan
    finds nothing; the peak-run search tries runs of 8, 4, 2, 1 over
    r = 0..hw at pp - r then pp + r and, never reaching a run of 4, keeps
    the last position where it found one: the farthest single match, pp + r
    before pp - r."""
    ovl, hw = 4096, 64
    rng = np.random.default_rng(3)
    prev = np.where(rng.random(ovl) > 0.5, 1.0, -1.0).astype(np.float32)
    cur = -prev
    match = np.zeros(ovl, bool)
    for k in range(333):
        match[1000 + 3 * k] = match[1001 + 3 * k] = True
    cur[match] = prev[match]
    d = detail(oracle, prev, cur, hw)
    assert d["scanned"] == 0
    pp = d["peak_pos"]
    order = [t for r in range(hw + 1) for t in (pp - r, pp + r) if 0 <= t < ovl]
    last = [t for t in order if match[t]][-1]
    assert len(d["cands"]) == 1
    assert d["cands"][0]["pos"] == last and d["cands"][0]["run"] == 1
    assert d["pos"] == last


def test_no_match_at_all(oracle):
    prev, cur = streams(4096, [])
    d = detail(oracle, prev, cur, 64)
    assert d["peak"] == 0 and len(d["cands"]) == 1
    assert d["cands"][0]["pos"] == 0 and d["pos"] == 0


# ---- quality: largest |255-tap decode| of the masked difference over +-512 ----

def qfir_numpy():
    """255-tap Kaiser (beta 8) low-pass at 20 kHz / 5.6448 MHz, unit sum,
    with numpy's exact I0 (np.kaiser)."""
    fc = 20000.0 / (5644800.0 / 2.0)
    n = np.arange(255) - 127.0
    h = np.where(n == 0, 2 * fc, np.sin(2 * np.pi * fc * n) / (np.pi * np.where(n == 0, 1, n)))
    h = h * np.kaiser(255, 8.0)
    return h / h.sum()


def test_qfir_taps(oracle):
    L = oracle.lib
    L.or_das_qfir.restype = _dp
    q = np.ctypeslib.as_array(L.or_das_qfir(), shape=(255,)).copy()
    ref = qfir_numpy()
    assert np.max(np.abs(q - ref)) < 1e-12 * np.max(np.abs(ref))


def decoded_max(prev, cur, sp, q):
    """max over j in [-512, 512) of |sum_k q[k] d[sp + j - 127 + k]|, d the
    difference (cur - prev) from sp on (zero before sp and outside the
    overlap), by numpy correlation."""
    d = (cur.astype(np.float64) - prev.astype(np.float64))
    d[:sp] = 0.0
    full = np.convolve(d, q[::-1])         # full[m + 254] = sum_k q[k] d[m + k]
    vals = []
    for j in range(-512, 512):
        i = sp + j - 127 + 254
        vals.append(abs(full[i]) if 0 <= i < full.size else 0.0)
    return max(vals)


@pytest.mark.parametrize("runs", [[(500, 700), (2000, 2300)], [(1000, 1100), (1105, 1400)],
                                  [(100, 400), (1300, 1500), (3000, 3900)]])
def test_candidate_quality_is_decoded_difference(oracle, runs):
    prev, cur = streams(4096, runs, seed=7)
    d = detail(oracle, prev, cur, 64)
    assert len(d["cands"]) > 1
    q = qfir_numpy()
    for c in d["cands"]:
        want = decoded_max(prev, cur, c["pos"], q)
        assert abs(c["max_decoded"] - want) <= 1e-12 * max(want, 1.0), c
        ratio = max(c["density"] / max(d["peak"], 1), 0.5)
        assert c["quality"] == c["max_decoded"] / ratio


def test_decode_window_is_plus_minus_512(oracle):
    """Mismatch on [0, 1000), match after it except one sample at 1651:
    candidates 1051 (run 600, density 115) and 1652 (density 127); the
    peak run at 1064 lies within hw of 1051.  For 1051 the masked
    difference is one +-2 impulse 600 samples on, so over j in [-512, 512)
    the decode sees only taps 216..254: max_decoded = 2 max|q[216:255]|
    (a wider window would reach the centre tap).  For 1652 the difference is
    zero, so its quality 0 wins."""
    prev, cur = streams(4096, [(1000, 1651), (1652, 4096)])
    d = detail(oracle, prev, cur, 64)
    got = [(c["pos"], c["run"], c["density"]) for c in d["cands"]]
    assert got == [(1051, 600, 115), (1652, 2444, 127)]
    q = qfir_numpy()
    want = 2.0 * np.max(np.abs(q[216:255]))
    assert abs(d["cands"][0]["max_decoded"] - want) <= 1e-12 * want
    assert d["cands"][1]["max_decoded"] == 0.0 and d["cands"][1]["quality"] == 0.0
    assert d["best"] == 1 and d["pos"] == 1652


# ---- the 0.8 / 1.2 / density rule (src/dsp_plugin.c:2641-2652) ----

@pytest.mark.parametrize("quality,density,best", [
    ([1.0, 0.79], [10, 5], 1),             # clearly better pop: taken
    ([1.0, 0.81], [10, 5], 0),             # not below 0.8; not denser
    ([1.0, 0.8], [10, 5], 1),              # 0.8 / (1 + 1e-15) is just below 0.8
    ([1.0, 0.81], [10, 11], 1),            # within 1.2 and denser
    ([1.0, 1.19], [10, 11], 1),            # within 1.2 and denser
    ([1.0, 1.25], [10, 99], 0),            # worse pop: density does not matter
    ([1.0, 1.1], [10, 10], 0),             # equal density is not higher
    ([1.0, 0.9, 0.75], [10, 20, 5], 1),    # 0.75 / 0.9 = 0.83: not clearly better than the kept one
    ([1.0, 0.9, 0.7], [10, 20, 5], 2),     # 0.7 / 0.9 = 0.78: clearly better
    ([2.0, 1.0, 1.5], [10, 1, 30], 1),     # 1.5 is compared with the kept 1.0, not the first 2.0
    ([0.0, 0.0], [5, 1], 1),               # zero pops: 0 / 1e-15 = 0 < 0.8
])
def test_pick_rule(oracle, quality, density, best):
    L = oracle.lib
    L.ora_pick.argtypes = [C.c_int, _dp, _ip]
    q = np.array(quality, np.float64)
    d = np.array(density, np.int32)
    assert L.ora_pick(len(q), q.ctypes.data_as(_dp), d.ctypes.data_as(_ip)) == best


def test_stitch_returns_the_picked_candidate(oracle):
    L = oracle.lib
    L.ora_pick.argtypes = [C.c_int, _dp, _ip]
    for seed in range(5):
        prev, cur = streams(4096, [(200, 500), (1500, 1800), (2500, 2560), (3000, 3600)], seed=seed)
        d = detail(oracle, prev, cur, 64)
        q = np.array([c["quality"] for c in d["cands"]])
        den = np.array([c["density"] for c in d["cands"]], np.int32)
        b = L.ora_pick(len(q), q.ctypes.data_as(_dp), den.ctypes.data_as(_ip))
        assert d["best"] == b and d["pos"] == d["cands"][b]["pos"]
        assert ol.stitch(oracle, prev, cur, 64)[0] == d["pos"]


# ---- reposition (src/dsp_plugin.c:1963-2034): closed-form quiet stretches ----

def test_reposition_constructed():
    """Seven 65,536-sample segments (lat 16: O = 1024; max shift
    min(0.4 * 65536, 200000) = 26214).  The signal is +-A everywhere except
    silent stretches, so the RMS of the 2048-sample window at c is
    A * sqrt(n / 2048) with n its loud samples.  Boundary 2: silent over
    [b - 4000, b + 6000): ring 3 (3072) first reaches the left loud part
    (96 samples) -> b - 3072.  Boundary 3: silent over [b - 4000, b + 4000),
    louder right side: both sides reach 96 loud samples at ring 3, the right
    one is louder -> b + 3072.  Boundary 4 is loud: unchanged.  Boundary 5
    is silent beyond every ring: unchanged.  Boundary 6 (last): silent over
    [b - 30000, b + 2000): ring 1 reaches 48 loud samples on the right ->
    b + 1024."""
    oracle = ol.oracle()
    Ls = oracle.lib
    Ls.ora_reposition.argtypes = [C.c_size_t, C.c_size_t, C.c_int, _dp, _lp, _lp, _ip]
    seg, S = 65536, 7
    F = seg * S
    n = np.arange(F)
    x = np.where(n % 2 == 0, 0.1, -0.1)
    b = [seg * k for k in range(S)]
    x[b[2] - 4000:b[2] + 6000] = 0.0
    x[b[3] - 4000:b[3] + 4000] = 0.0
    x[b[3] + 4000:b[3] + 20000] *= 2.0          # louder right of boundary 3
    x[b[5] - 27300:b[5] + 27300] = 0.0
    x[b[6] - 30000:b[6] + 2000] = 0.0
    st = np.zeros(S, np.int64)
    sz = np.zeros(S, np.int64)
    ns = C.c_int(0)
    Ls.ora_reposition(F, seg, 16, x.ctypes.data_as(_dp), st.ctypes.data_as(_lp), sz.ctypes.data_as(_lp),
                      C.byref(ns))
    assert ns.value == S
    assert st.tolist() == [0, b[1], b[2] - 3072, b[3] + 3072, b[4], b[5], b[6] + 1024]
    assert sz.sum() == F and all(st[i] + sz[i] == st[i + 1] for i in range(S - 1))


def test_reposition_respects_neighbour_margin():
    """A candidate must stay more than 2 O from the previous boundary (and
    2 O before the next): with the previous boundary 3072 + 2048 away
    (reposition of boundary 2 above) the left ring-3 candidate of the next
    boundary is allowed, and one at 2 O is not."""
    oracle = ol.oracle()
    Ls = oracle.lib
    Ls.ora_reposition.argtypes = [C.c_size_t, C.c_size_t, C.c_int, _dp, _lp, _lp, _ip]
    seg, S, lat = 8192, 5, 16                 # O = 1024, max shift 3276
    F = seg * S
    n = np.arange(F)
    x = np.where(n % 2 == 0, 0.1, -0.1)
    b2 = 2 * seg
    # silent around b2 except loud again from b2 - 3000 leftwards: ring 2 (2048)
    # reaches the left loud part; ring 3 would be 3072 > max shift
    x[b2 - 3000:b2 + 5000] = 0.0
    st = np.zeros(S, np.int64)
    sz = np.zeros(S, np.int64)
    ns = C.c_int(0)
    Ls.ora_reposition(F, seg, lat, x.ctypes.data_as(_dp), st.ctypes.data_as(_lp), sz.ctypes.data_as(_lp),
                      C.byref(ns))
    # ring 2: left window [b2 - 3072, b2 - 1024) holds 72 loud samples -> b2 - 2048,
    # which is 6144 > 2 O after boundary 1
    assert st[2] == b2 - 2048


# ---- seeding and assembly against the reference's own trellis.c ----

def _x(nch, F, seed, rate=2822400):
    rng = np.random.default_rng(seed)
    t = np.arange(F) / rate
    return np.stack([0.3 * np.sin(2 * np.pi * (997 + 211 * c) * 16 * t) + 0.05 * rng.standard_normal(F)
                     for c in range(nch)])


def test_all_parallel_two_segments_from_reference(oracle, reference):
    """Two segments seeded from the persistent state, two chunks: seg0 is
    the persistent context over its nominal size plus O (kept to the
    nominal size), seg1 a copy of the persistent context with every
    survivor set to its best state at cost 0, run from nominal start - O
    with the first O outputs discarded; the output is seg0 up to the stitch
    point O before its end plus seg1 from the stitch position; the next
    chunk starts from a copy of seg1's context.  Built here from the
    reference's own sdm_* calls, against the oracle DAS."""
    K, L, B, fi_rate = 8, 16, 8, 2822400
    fi = reference.ntf_auto(fi_rate)
    O = 64 * L
    nch = 1
    das = ol.DasStream(oracle, fi, B, K, L, nch)
    persist = reference.ctx(fi, B, K, L)
    first = True
    for k, F in enumerate([40000, 52000]):
        x = _x(nch, F, 50 + k)[0]
        got = das.chunk(x[None, :], 0, 0, 0, nseg=2)[0]
        size0 = F // 2
        # seg1: copy + seed, from size0 - O, discard O outputs (copied, filled) or
        # L + (O - L) inputs (fresh)
        t = reference.ctx(fi, B, K, L)
        t.copy_from(persist)
        t.seed(persist.best_state())
        y1 = t.process(x[size0 - O:])
        kept1 = y1[O - L:] if first else y1[O:]
        # seg0: the persistent context over size0 + O, kept to size0
        y0 = persist.process(x[:size0 + O])
        y0 = y0[:size0]
        # stitch in the last O of seg0's output
        ovl = min(O, kept1.size)
        pos = ol.stitch(oracle, y0[y0.size - O:], kept1[:ovl], L)[0]
        want = np.concatenate([y0[:y0.size - O + pos], kept1[pos:]])
        assert got.size == want.size, (k, got.size, want.size)
        assert np.array_equal(got, want), k
        nxt = reference.ctx(fi, B, K, L)
        nxt.copy_from(t)
        persist = nxt
        first = False


def test_seg0_first_three_segments_from_reference(oracle, reference):
    """seg0-first with three segments: seg0 on the persistent context, seg1 a
    copy of it continuing over its segment plus O (all outputs kept), seg2 a
    copy seeded with the 16-level estimate run from the persistent best
    state across seg1 (the reference's sdm_estimate_state), from its start
    - O, first O outputs discarded, stitched into the last O of seg1's
    output.  The output count is F - L."""
    K, L, B = 8, 16, 8
    fi = reference.ntf_auto(2822400)
    O = 64 * L
    F = 3 * 20000
    x = _x(1, F, 61)[0]
    out, n, sp, plan = ol.das(oracle, fi, B, K, L, x[None, :], seg_len=0, nseg=3, seg0_first=1, reposition=0)
    st = plan[:, 0].tolist()
    assert st == [0, 20000, 40000]
    persist = reference.ctx(fi, B, K, L)
    y0 = persist.process(x[:st[1]])
    t1 = reference.ctx(fi, B, K, L)
    t1.copy_from(persist)
    y1 = t1.process(x[st[1]:st[2] + O])
    est = reference.estimate(fi, persist.best_state(), x[st[1]:st[2]])
    t2 = reference.ctx(fi, B, K, L)
    t2.copy_from(persist)
    t2.seed(est)
    y2 = t2.process(x[st[2] - O:])[O:]
    head = np.concatenate([y0, y1])
    pos = ol.stitch(oracle, head[head.size - O:], y2[:O], L)[0]
    assert pos == sp[2]
    want = np.concatenate([head[:head.size - O + pos], y2[pos:]])
    assert n == want.size == F - L
    assert np.array_equal(out[0, :n], want)
