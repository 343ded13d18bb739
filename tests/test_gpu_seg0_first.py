"""The reference's DSD256 DAS (seg0-first seeding, src/dsp_plugin.c:2185-
2266): seg0 on the persistent context, seg1 its exact continuation, every
later segment seeded by the 16-level estimate from seg1's start (host
threads, segments launched as their estimates arrive).  Bit-exact against
the oracle at the headline shape (DSD256, K=32, L=64, 65,536-sample segments),
with the output count the reference's plugin checks (F, or F - L on the
first chunk, :2733-2747) and the host-buffer output's tail."""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle_lib as ol
from foo_dsd_trellis_amd import das as D
from foo_dsd_trellis_amd import signals as SG

pytestmark = pytest.mark.gpu

RATE, K, L, B = 11289600, 32, 64, 8
SEG = 65536


def _pink(nch, F, seed):
    rng = np.random.default_rng(seed)
    return np.stack([np.clip(0.18 * SG.pink(F, rng), -0.5, 0.5) for _ in range(nch)])


@pytest.mark.parametrize("seg0", [1, 0])
def test_dsd256_k32_full_segments(gpu, oracle, seg0):
    """Four 65,536-sample segments plus a ragged tail, both seed modes."""
    nch, F = 2, 4 * SEG + 12345
    print("case: DSD256 K32 L64 seg 65536 F=%d seg0_first=%d" % (F, seg0), flush=True)
    x = _pink(nch, F, 5)
    fi = oracle.ntf_auto(RATE)
    ref, n, sp, plan = ol.das(oracle, fi, B, K, L, x, seg_len=SEG, nseg=0, seg0_first=seg0, reposition=1)
    d = D.Das(fi, B, K, L, nch, seg_len=SEG, seg0_first=bool(seg0), reposition=True)
    y, m = d.run(x)
    S, slow, pos = d.last_info(len(sp))
    assert S == len(sp) == 4
    assert m == n
    if seg0:
        assert n == F - L                 # first chunk: only the latency fill is missing
        ms, ns, redo, nb = d.last_estimate()
        size0 = F // 4 + (1 if F % 4 else 0)        # the reference's sizing, base + (s < rem)
        assert ns == nch * (F - size0) and nb >= 1
    assert list(pos[2 if seg0 else 1:]) == list(sp[2 if seg0 else 1:])
    for c in range(nch):
        bad = np.nonzero(y[c, :n] != ref[c, :n])[0]
        assert bad.size == 0, "ch%d first mismatch %d" % (c, bad[0])
    assert not np.any(y[:, n:])           # the tail past the produced count is 0.0


def test_seg0_first_chunks_keep_count_contract(gpu, oracle):
    """Chunks of changing size on persistent contexts: every chunk after the
    first produces exactly its input count, bits equal the oracle's."""
    nch, seg = 2, 16384
    fi = oracle.ntf_auto(RATE)
    d = D.Das(fi, B, 16, 32, nch, seg_len=seg, seg0_first=True, reposition=True)
    ref = ol.DasStream(oracle, fi, B, 16, 32, nch)
    for k, F in enumerate([90000, 40000, 131072, 16384, 70001]):
        x = _pink(nch, F, 40 + k)
        y, m = d.run(x)
        want = ref.chunk(x, seg, 1, 1)
        assert m == want.shape[1] == (F - 32 if k == 0 else F), (k, m, want.shape)
        assert np.array_equal(y[:, :m], want), k
        assert not np.any(y[:, m:]), k


def test_estimate_on_device_lane_gives_same_bits(gpu, oracle):
    """The seeding estimate on one device lane per channel (das_estimate)
    and on host threads give identical boundary states and bits."""
    nch, F, seg = 2, 200000, 16384
    x = _pink(nch, F, 77)
    fi = oracle.ntf_auto(RATE)
    outs = []
    for dev in (False, True):
        d = D.Das(fi, B, 16, 32, nch, seg_len=seg, seg0_first=True, reposition=True)
        d.estimator_device(dev)
        outs.append(d.run(x))
    assert outs[0][1] == outs[1][1]
    assert np.array_equal(outs[0][0], outs[1][0])
    # the estimator entries themselves
    init = np.random.default_rng(1).standard_normal((nch, 10)) * 0.01
    bnd = np.array([0, 1000, 50000, 50000, 150000, F], np.int64)
    dx = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    got = D.estimate_boundaries_dev(fi, init, dx.data_ptr(), F, nch, F, bnd)
    for c in range(nch):
        host, _ = D.estimate_boundaries(fi, init[c], x[c], bnd)
        P = oracle.ntf_get(fi)["order"]
        assert np.array_equal(got[c][:, :P].view(np.uint64), host[:, :P].view(np.uint64)), c


def _seam_setup(L, h, oracle, K_, Lat):
    fi = oracle.ntf_auto(RATE)
    f = oracle.ntf_get(fi)
    a = np.ascontiguousarray(f["a"][:10])
    g = np.ascontiguousarray(f["g"][:10])
    dp = lambda v: v.ctypes.data_as(C.POINTER(C.c_double))
    assert L.gpu_hip_trellis_setup(h, K_, f["order"], Lat, dp(a), dp(g), 0.0, B) == 0
    return fi


@pytest.mark.parametrize("rate_known", [False, True])
def test_seam_das_count_and_tail(gpu, oracle, rate_known):
    """gpu_hip_trellis_das(_ex) seeds by the reference's rule (seg0-first
    below DSD512, the default), reports the produced count (F - L, then F)
    and leaves 0.0 past it, also where an earlier, longer call wrote."""
    from foo_dsd_trellis_amd._lib import lib
    L = lib()
    h = L.gpu_hip_create()
    assert h
    try:
        Kc, Lat, seg, nch = 16, 32, 16384, 2
        fi = _seam_setup(L, h, oracle, Kc, Lat)
        if rate_known:
            assert L.gpu_hip_das_rate(h, RATE) == 0
        assert L.gpu_hip_das_config(h, seg, -1, 1) == 0
        das = ol.DasStream(oracle, fi, B, Kc, Lat, nch)
        out = np.full((nch, 100000), 7.0, np.float32)
        for k, F in enumerate([100000, 50000, 70000]):
            x = np.ascontiguousarray(_pink(nch, F, 90 + k))
            o = np.ascontiguousarray(out[:, :F]) if k else out
            cnt = C.c_size_t(0)
            rc = L.gpu_hip_trellis_das_ex(h, x.ctypes.data_as(C.POINTER(C.c_double)),
                                          o.ctypes.data_as(C.POINTER(C.c_float)), F, nch, C.byref(cnt))
            assert rc == 0
            want = das.chunk(x, seg, 1, 1)
            assert cnt.value == want.shape[1] == (F - Lat if k == 0 else F), (k, cnt.value)
            assert L.gpu_hip_trellis_das_count(h) == cnt.value
            assert np.array_equal(o[:, :cnt.value], want), k
            assert not np.any(o[:, cnt.value:F]), k
    finally:
        L.gpu_hip_destroy(h)


def test_stream_dsd256_seg0_first_count(gpu, oracle):
    """The streaming chain at DSD256 seeds seg0-first by default: no chunk
    loses time at the seam; counts are the chunk's samples (minus L on the
    first chunk), bits equal the oracle's chunked DAS on the same input."""
    from foo_dsd_trellis_amd import stream, trellis
    fs_in, nch, cf, seg = 44100, 2, 1024, 16384
    pcm = SG.bandlimited_gauss(3 * cf + 300, fs=fs_in, nch=nch, seed=12, fmax=18000.0)
    fi = trellis.ntf_auto(RATE)
    S = stream.Stream(fs_in, RATE, nch, cf, ntf_index=fi, K=16, L=32, depth=B, seg_len=seg)
    assert S.seg0_first
    das = ol.DasStream(oracle, fi, B, 16, 32, nch)
    d_pcm = torch.from_numpy(np.ascontiguousarray(pcm)).cuda()
    for k, f0 in enumerate(range(0, pcm.shape[1], cf)):
        f1 = min(pcm.shape[1], f0 + cf)
        n = S.process(d_pcm[:, f0:f1])
        sig = S.signal()
        want = das.chunk(sig, seg, 1, 1)
        F = (f1 - f0) * S.R
        assert n == want.shape[1] == (F - 32 if k == 0 else F), (k, n)
        assert np.array_equal(S.bits_numpy(n), want), k


def test_host_fed_launch_equals_batches(gpu, oracle):
    """The host-fed persistent launch (every segment starts once its estimate
    is published) and the batch form give the oracle's bits; six channels,
    so the estimate chains publish interleaved."""
    nch, F, seg = 6, 180000, 16384
    x = _pink(nch, F, 91)
    fi = oracle.ntf_auto(RATE)
    print("case: host-fed vs batches, 6 ch, F=%d seg=%d K=16 L=32" % (F, seg), flush=True)
    ref, n, _, _ = ol.das(oracle, fi, B, 16, 32, x, seg_len=seg, nseg=0, seg0_first=1, reposition=1)
    outs = []
    for batches in (False, True):
        d = D.Das(fi, B, 16, 32, nch, seg_len=seg, seg0_first=True, reposition=True)
        d.batch_launch(batches)
        y, m = d.run(x)
        ms, ns, redo, nb = d.last_estimate()
        assert m == n
        assert (nb == 1) if not batches else (nb >= 1)
        assert np.array_equal(y[:, :n], ref[:, :n]), "batches=%s" % batches
        outs.append(y)
    assert np.array_equal(outs[0], outs[1])


def test_host_fed_wait_limit_ends_the_launch(gpu, oracle):
    """A wave that waits longer than the limit for its segment leaves: with
    a zero limit the run fails with an error instead of waiting (the
    condition is forced from the host; every wave exits), and a new object
    with the default limit then runs bit-exact."""
    nch, F, seg = 2, 120000, 16384
    x = _pink(nch, F, 93)
    fi = oracle.ntf_auto(RATE)
    d = D.Das(fi, B, 16, 32, nch, seg_len=seg, seg0_first=True, reposition=True)
    d.feed_limit(0.0)
    with pytest.raises(Exception):
        d.run(x)
    ref, n, _, _ = ol.das(oracle, fi, B, 16, 32, x, seg_len=seg, nseg=0, seg0_first=1, reposition=1)
    d2 = D.Das(fi, B, 16, 32, nch, seg_len=seg, seg0_first=True, reposition=True)
    y, m = d2.run(x)
    assert m == n and np.array_equal(y[:, :n], ref[:, :n])


def test_host_fed_launch_with_streams_sharing_queues(gpu, oracle):
    """Regression: the host-fed launch waits for estimates whose input is
    copied on the DAS stream; it is queued behind those copies on that
    stream, so it cannot block them even when streams share hardware queues.
    With more streams in use than hardware queues, seg0-first runs on the
    default stream and on a side stream finish bit-exact within a short wait
    limit."""
    side = [torch.cuda.Stream() for _ in range(12)]
    for s in side:
        with torch.cuda.stream(s):
            torch.ones(16, device="cuda").sum()
    torch.cuda.synchronize()
    nch, F, seg = 2, 150000, 16384
    x = _pink(nch, F, 95)
    fi = oracle.ntf_auto(RATE)
    ref, n, _, _ = ol.das(oracle, fi, B, 16, 32, x, seg_len=seg, nseg=0, seg0_first=1, reposition=1)
    d_x = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    words = F // 64 + 2
    for use_side in (False, True):
        print("case: host-fed launch, device buffers, %s stream, 12 side streams"
              % ("side" if use_side else "default"), flush=True)
        d = D.Das(fi, B, 16, 32, nch, seg_len=seg, seg0_first=True, reposition=True)
        d.feed_limit(5.0)
        if use_side:
            d.set_stream(side[5].cuda_stream)
        d_o = torch.zeros((nch, words), dtype=torch.int64, device="cuda")
        torch.cuda.synchronize()
        m = d.run_dev(d_x.data_ptr(), F, F, d_o.data_ptr(), words)
        torch.cuda.synchronize()
        w = d_o.cpu().numpy().view(np.uint64)
        bits = np.unpackbits(w.view(np.uint8), axis=1, bitorder="little")[:, :m]
        assert m == n and np.array_equal(np.where(bits > 0, 1.0, -1.0).astype(np.float32), ref[:, :n])
