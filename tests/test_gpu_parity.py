"""GPU parity: HIP trellis / FIR / DAS against the CPU oracle on seeded
inputs (bit-exact for DSD bits, identical doubles for the FIR)."""
import numpy as np
import pytest

import oracle_lib as ol
from foo_dsd_trellis_amd import das as D
from foo_dsd_trellis_amd import fir as FIR
from foo_dsd_trellis_amd import signals as SG
from foo_dsd_trellis_amd import trellis as T

pytestmark = pytest.mark.gpu


def _x(kind, n, rate, seed):
    rng = np.random.default_rng(seed)
    t = np.arange(n) / rate
    if kind == "sine":
        return 0.5012 * np.sin(2 * np.pi * 997 * 16 * t) + 1e-5 * rng.standard_normal(n)
    if kind == "pink":
        return np.clip(0.18 * SG.pink(n, rng), -0.5, 0.5)
    if kind == "loud":   # drives bucket-0 lookups after the act[32] write at K=32
        return 0.45 * np.sin(2 * np.pi * 997 * 16 * t) + 0.05 * rng.standard_normal(n)
    if kind == "quiet":  # pink with near-silent stretches over segment boundaries
        x = np.clip(0.18 * SG.pink(n, rng), -0.5, 0.5)
        for a in (30000, 62000):
            x[a:a + 6000] *= 1e-3
        return x
    raise ValueError(kind)


TRELLIS_CASES = [
    # K, L, B, rate, n, kind
    (8, 16, 8, 2822400, 30000, "sine"),
    (32, 64, 8, 11289600, 30000, "pink"),
    (32, 64, 8, 11289600, 30000, "loud"),
    (16, 32, 8, 5644800, 20000, "pink"),
    (32, 100, 16, 11289600, 8000, "pink"),
    (4, 8, 4, 2822400, 8000, "sine"),
    (1, 1, 1, 2822400, 4000, "sine"),
    (32, 200, 8, 22579200, 6000, "loud"),
    (4, 2048, 8, 2822400, 7000, "pink"),      # longest history (32 words per survivor)
    (2, 129, 4, 11289600, 9000, "pink"),      # K = 2 past the lane kernel's history limit
    (32, 64, 32, 11289600, 6000, "loud"),     # 32-bit paths
    (32, 2048, 8, 11289600, 5000, "loud"),    # largest K and L together
]


@pytest.mark.parametrize("K,L,B,rate,n,kind", TRELLIS_CASES)
def test_trellis_bit_exact(gpu, oracle, K, L, B, rate, n, kind):
    x = _x(kind, n, rate, K * 131 + L)
    fi = oracle.ntf_auto(rate)
    ref = oracle.ctx(fi, B, K, L).process(x)
    p = T.SdmPool(fi, B, K, L, 0.0, 1)
    y = p.process(0, x)
    assert len(y) == len(ref)
    bad = np.nonzero(y != ref)[0]
    assert bad.size == 0, "first mismatch at %d of %d" % (bad[0], len(ref))


@pytest.mark.parametrize("K,L,B,rate,n,kind", TRELLIS_CASES[:4])
def test_trellis_literal_path_bit_exact(gpu, oracle, K, L, B, rate, n, kind):
    """Every step decided by the kernel's literal prune gives the same bits."""
    x = _x(kind, n // 2, rate, K * 7 + L)
    fi = oracle.ntf_auto(rate)
    ref = oracle.ctx(fi, B, K, L).process(x)
    p = T.SdmPool(fi, B, K, L, 0.0, 1)
    p.force_literal(True)
    y = p.process(0, x)
    assert np.array_equal(y, ref)


def test_trellis_state_limit_and_continuation(gpu, oracle):
    """Block-wise processing continues exactly (persistent device state)."""
    fi = oracle.ntf_auto(11289600)
    x = _x("loud", 12000, 11289600, 5)
    ref = oracle.ctx(fi, 8, 32, 64, 2.5).process(x)
    p = T.SdmPool(fi, 8, 32, 64, 2.5, 1)
    ys = [p.process(0, x[a:b]) for a, b in [(0, 33), (33, 5000), (5000, 5001), (5001, 12000)]]
    assert np.array_equal(np.concatenate(ys), ref)


def test_trellis_copy_and_seed(gpu, oracle):
    """copy_state + seeding as the DAS dispatch does, then continue."""
    fi = oracle.ntf_auto(11289600)
    x = _x("pink", 20000, 11289600, 9)
    o_src = oracle.ctx(fi, 8, 32, 64)
    o_src.process(x[:8000])
    o_dst = oracle.ctx(fi, 8, 32, 64)
    o_dst.copy_from(o_src)
    st = o_src.best_state() * 0.5
    o_dst.seed(st)
    ref = o_dst.process(x[8000:])
    p = T.SdmPool(fi, 8, 32, 64, 0.0, 2)
    p.process(0, x[:8000])
    p.copy(1, 0)
    p.seed(1, st)
    y = p.process(1, x[8000:])
    assert np.array_equal(y, ref)


@pytest.mark.parametrize("fs_in,fs_out,n,gain", [
    (44100, 2822400, 44100, FIR.PCM_FIR_GAIN),
    (44100, 11289600, 6000, FIR.PCM_FIR_GAIN),
    (48000, 24576000, 3000, 1.0),
])
def test_fir_matches_oracle(gpu, oracle, fs_in, fs_out, n, gain):
    x = SG.sine_noise(n, fs=fs_in, nch=2, seed=fs_out % 97)
    f = FIR.FirChain(fs_in, fs_out, gain, nch=2)
    y = f.process(x)
    for c in range(2):
        ref = ol.fir_chain(oracle, fs_in, fs_out, x[c], gain)
        assert np.array_equal(y[c], ref), np.max(np.abs(y[c] - ref))
        truth = ol.fir_chain(oracle, fs_in, fs_out, x[c], gain, long_double=True)
        assert np.max(np.abs(y[c] - truth)) < 1e-12


def test_fir_streaming_equals_one_shot(gpu):
    x = SG.sine_noise(40000, nch=2, seed=3)
    f1 = FIR.FirChain(44100, 2822400, 1.0, 2)
    whole = f1.process(x)
    f2 = FIR.FirChain(44100, 2822400, 1.0, 2)
    parts = [f2.process(x[:, a:b]) for a, b in [(0, 10), (10, 9000), (9000, 40000)]]
    assert np.array_equal(np.concatenate(parts, axis=1), whole)


DAS_CASES = [
    # K, L, B, rate, F, seg_len, nseg, seg0_first, kind
    (8, 16, 8, 2822400, 60000, 0, 8, 1, "sine"),
    (8, 16, 8, 2822400, 60000, 0, 8, 0, "sine"),
    (32, 64, 8, 11289600, 120000, 16384, 0, 0, "pink"),
    (32, 64, 8, 11289600, 120000, 16384, 0, 1, "pink"),
    (16, 32, 8, 5644800, 90000, 0, 5, 1, "loud"),
    (32, 64, 8, 11289600, 120000, 16384, 0, 0, "quiet"),   # boundaries moved out of silence
    (8, 16, 8, 2822400, 120000, 16384, 0, 1, "quiet"),
]


@pytest.mark.parametrize("K,L,B,rate,F,seg_len,nseg,seg0,kind", DAS_CASES)
def test_das_bit_exact(gpu, oracle, K, L, B, rate, F, seg_len, nseg, seg0, kind):
    nch = 2
    x = np.stack([_x(kind, F, rate, 11 + c) for c in range(nch)])
    fi = oracle.ntf_auto(rate)
    ref, n, sp, plan = ol.das(oracle, fi, B, K, L, x, seg_len=seg_len, nseg=nseg,
                              seg0_first=seg0, reposition=1)
    d = D.Das(fi, B, K, L, nch, seg_len=seg_len, nseg=nseg, seg0_first=bool(seg0), reposition=True)
    y, m = d.run(x)
    S, slow, pos = d.last_info(len(sp))
    assert S == len(sp)
    if kind == "quiet":   # the quiet-boundary reposition (src/dsp_plugin.c:1963-2034) took place
        assert any(plan[i, 0] != i * seg_len for i in range(2, S)), plan[:, 0].tolist()
    assert m == n
    assert list(pos[1:]) == list(sp[1:]) or S == 1
    for c in range(nch):
        bad = np.nonzero(y[c, :n] != ref[c, :n])[0]
        assert bad.size == 0, "ch%d first mismatch %d" % (c, bad[0])


@pytest.mark.parametrize("K", [16, 32])
def test_equal_metrics_in_parallel_prune(gpu, oracle, K):
    """Seeded survivors share one state and metric 0, so their children tie
    exactly for several steps.  The sequential sort orders equal metrics by
    index; the parallel prune decides those steps in that order (below
    K = 32 no step needs the literal prune) and matches the oracle."""
    import torch
    rate = 11289600
    fi = oracle.ntf_auto(rate)
    x = _x("pink", 16000, rate, 23)
    o_src = oracle.ctx(fi, 8, K, 64)
    o_src.process(x[:6000])
    st = o_src.best_state()
    o_dst = oracle.ctx(fi, 8, K, 64)
    o_dst.copy_from(o_src)
    o_dst.seed(st)
    ref = o_dst.process(x[6000:])
    p = T.SdmPool(fi, 8, K, 64, 0.0, 2)
    p.process(0, x[:6000])
    d_in = torch.from_numpy(np.ascontiguousarray(x[6000:])).cuda()
    n = d_in.numel()
    d_out = torch.zeros(n // 64 + 2, dtype=torch.int64, device="cuda")
    seed = np.full((1, 10), np.nan)
    seed[0, :] = st
    kept, slow = p.run_jobs([1], [d_in.data_ptr()], [n], [d_out.data_ptr()], copy_src=[0], seed=seed)
    torch.cuda.synchronize()
    assert kept[0] == len(ref)
    y = T.unpack_bits(d_out.cpu().numpy(), int(kept[0]))
    assert np.array_equal(y, ref)
    if K < 32:
        assert slow[0] == 0


def test_trellis_empty_and_tiny_calls(gpu, oracle):
    """Empty and one-sample calls between blocks change nothing: the output
    stream equals the oracle's over the concatenated input (the reference's
    sdm_process_block takes any count, src/trellis.c:1013-1040)."""
    fi = oracle.ntf_auto(11289600)
    x = _x("pink", 9000, 11289600, 13)
    ref = oracle.ctx(fi, 8, 32, 64).process(x)
    p = T.SdmPool(fi, 8, 32, 64, 0.0, 1)
    ys = []
    for a, b in [(0, 0), (0, 1), (1, 1), (1, 64), (64, 65), (65, 65), (65, 9000), (9000, 9000)]:
        ys.append(p.process(0, x[a:b]))
    assert [y.size for y in ys[:3]] == [0, 0, 0]
    assert np.array_equal(np.concatenate(ys), ref)


def test_seam_trellis_short_calls(gpu, oracle):
    """gpu_hip_trellis, like gpu_cuda_trellis (src/gpu_cuda.c:1434-1437),
    has no minimum count (the plugin checks GPU_MIN_SAMPLES before calling,
    src/dsp_plugin.c:1405): short and empty calls continue the stream
    exactly.  gpu_hip_trellis_das with no samples returns -1."""
    import ctypes as C
    from foo_dsd_trellis_amd._lib import lib
    L = lib()
    h = L.gpu_hip_create()
    assert h
    try:
        fi = oracle.ntf_auto(11289600)
        f = oracle.ntf_get(fi)
        a = np.ascontiguousarray(f["a"][:10])
        g = np.ascontiguousarray(f["g"][:10])
        dp = lambda v: v.ctypes.data_as(C.POINTER(C.c_double))
        fp = lambda v: v.ctypes.data_as(C.POINTER(C.c_float))
        assert L.gpu_hip_trellis_setup(h, 8, f["order"], 16, dp(a), dp(g), 0.0, 8) == 0
        x = _x("pink", 3000, 11289600, 17)
        ref = oracle.ctx(fi, 8, 8, 16).process(x)
        got = []
        for a0, b0 in [(0, 0), (0, 5), (5, 16), (16, 17), (17, 17), (17, 3000)]:
            part = np.ascontiguousarray(x[a0:b0])
            out = np.zeros(max(1, b0 - a0), np.float32)
            assert L.gpu_hip_trellis(h, dp(part if part.size else np.zeros(1)), fp(out), b0 - a0) == 0
            got.append(out[:max(0, b0 - max(a0, 16))])
        assert np.array_equal(np.concatenate(got), ref)
        z = np.zeros(1)
        o = np.zeros(1, np.float32)
        assert L.gpu_hip_trellis_das(h, dp(z), fp(o), 0, 1) == -1
    finally:
        L.gpu_hip_destroy(h)


@pytest.mark.parametrize("nch,F,nseg,seg0", [(2, 40000, 1, 1), (1, 40001, 1, 0), (2, 40000, 2, 1),
                                             (1, 65599, 2, 1)])
def test_das_reduces_to_sequential_reference(gpu, oracle, reference, nch, F, nseg, seg0):
    """One segment, or two with seg0-first seeding (segment 1 continues the
    persistent context exactly, src/dsp_plugin.c:2185-2250), is the plain
    sequential trellis: the GPU DAS output equals the reference's own
    trellis.c over each whole channel (ragged F and one channel included)."""
    rate = 2822400
    x = np.stack([_x("pink", F, rate, 31 + c) for c in range(nch)])
    fi = oracle.ntf_auto(rate)
    d = D.Das(fi, 8, 8, 16, nch, seg_len=0, nseg=nseg, seg0_first=bool(seg0), reposition=True)
    y, m = d.run(x)
    for c in range(nch):
        seq = reference.ctx(fi, 8, 8, 16).process(x[c])
        assert m == seq.size
        assert np.array_equal(y[c, :m], seq), c


def test_das_chunks_of_changing_size(gpu, oracle):
    """The host-buffer DAS entry over chunks that grow and shrink (its
    device staging only grows) continues the persistent contexts exactly
    like the oracle's chunked DAS."""
    rate, K, L, B, nch, seg = 11289600, 16, 32, 8, 2, 16384
    fi = oracle.ntf_auto(rate)
    d = D.Das(fi, B, K, L, nch, seg_len=seg, seg0_first=False, reposition=True)
    ref = ol.DasStream(oracle, fi, B, K, L, nch)
    for k, F in enumerate([70000, 20000, 90000, 16384]):
        x = np.stack([_x("pink", F, rate, 100 * k + c) for c in range(nch)])
        y, m = d.run(x)
        want = ref.chunk(x, seg, 0, 1)
        assert m == want.shape[1], k
        assert np.array_equal(y[:, :m], want), k


@pytest.mark.parametrize("fs_in,fs_out", [(44100, 11289600), (48000, 24576000), (44100, 2822400)])
def test_fir_fused_pairs_equal_stage_by_stage(gpu, oracle, fs_in, fs_out):
    """Adjacent 63-tap stages fused into one pass give exactly the
    stage-by-stage outputs and the oracle's, over chunks that cross tile
    edges, shorter than the fused pass's halo (16 inputs), and carry the
    delay lines of every stage (the intermediate one included) across
    chunks."""
    from foo_dsd_trellis_amd import fir, signals
    sizes = [1000, 7, 513, 2048, 15, 1, 4096, 333]
    n = sum(sizes)
    x = signals.pink_multitone(n, fs=fs_in, nch=2, seed=21)
    R = fs_out // fs_in
    outs = []
    for fused in (True, False):
        ch = fir.FirChain(fs_in, fs_out, fir.PCM_FIR_GAIN, 2)
        ch.set_fused(fused)
        parts = []
        c0 = 0
        for m in sizes:
            print("case: %d -> %d fused=%s chunk %d" % (fs_in, fs_out, fused, m), flush=True)
            parts.append(ch.process(np.ascontiguousarray(x[:, c0:c0 + m])))
            c0 += m
        outs.append(np.concatenate(parts, axis=1))
    assert outs[0].shape == (2, n * R)
    assert np.array_equal(outs[0], outs[1])
    for c in range(2):
        ref = ol.fir_chain(oracle, fs_in, fs_out, x[c], fir.PCM_FIR_GAIN)
        assert np.array_equal(outs[0][c], ref), c
