"""DAS oracle invariants (the reference's DAS driver, src/dsp_plugin.c,
needs Win32 and does not build here): one segment, or seg0-first with two
segments, is the sequential trellis; counts; stitch search on constructed
overlaps."""
import numpy as np
import pytest

import oracle_lib as ol


def _x(nch, F, seed):
    rng = np.random.default_rng(seed)
    t = np.arange(F)
    return 0.2 * np.sin(2 * np.pi * t * 1000 / 2822400 * 8)[None, :] + 0.03 * rng.standard_normal((nch, F))


@pytest.mark.parametrize("nseg,seg0", [(1, 1), (1, 0), (2, 1)])
def test_reduces_to_sequential(oracle, nseg, seg0):
    x = _x(2, 40000, 1)
    fi = oracle.ntf_auto(2822400)
    out, n, sp, plan = ol.das(oracle, fi, 8, 8, 16, x, nseg=nseg, seg0_first=seg0)
    for c in range(2):
        seq = oracle.ctx(fi, 8, 8, 16).process(x[c])
        assert n == seq.size
        assert np.array_equal(out[c, :n], seq)


def test_counts_and_plan(oracle):
    F = 60000
    x = _x(2, F, 2)
    fi = oracle.ntf_auto(2822400)
    out, n, sp, plan = ol.das(oracle, fi, 8, 8, 16, x, nseg=8, seg0_first=1)
    assert plan[:, 1].sum() == F and plan[0, 0] == 0
    assert n == F - 16                      # first chunk: latency fill on seg0 only
    out, n, sp, plan = ol.das(oracle, fi, 8, 8, 16, x, seg_len=8192, nseg=0, seg0_first=0)
    assert len(plan) == F // 8192 and plan[-1, 1] >= 8192
    assert set(np.unique(out[:, :n])) <= {-1.0, 1.0}


def test_stitch_identical_streams_picks_zero_cost(oracle):
    rng = np.random.default_rng(3)
    a = np.where(rng.random(4096) > 0.5, 1.0, -1.0).astype(np.float32)
    pos, q, nc = ol.stitch(oracle, a, a.copy(), 64)
    assert q == 0.0 or nc == 1
    assert 0 <= pos < 4096


def test_stitch_prefers_converged_region(oracle):
    """Streams that disagree except inside one window: the chosen position
    lies in that window (density-aligned)."""
    rng = np.random.default_rng(4)
    prev = np.where(rng.random(4096) > 0.5, 1.0, -1.0).astype(np.float32)
    cur = -prev
    cur[2000:2600] = prev[2000:2600]
    pos, q, nc = ol.stitch(oracle, prev, cur, 64)
    assert 2000 <= pos < 2600


def test_chunk_stream_first_chunk_equals_das(oracle):
    """The per-chunk oracle DAS on its first chunk is the one-shot DAS, and
    later chunks continue from the persistent contexts (full output)."""
    x = _x(2, 3 * 20000, 7)
    fi = oracle.ntf_auto(2822400)
    st = ol.DasStream(oracle, fi, 8, 8, 16, 2)
    a = st.chunk(x[:, :20000], 4096, 0, 1)
    ref, n, _, _ = ol.das(oracle, fi, 8, 8, 16, x[:, :20000], seg_len=4096, nseg=0, seg0_first=0)
    assert a.shape[1] == n and np.array_equal(a, ref[:, :n])
    b = st.chunk(x[:, 20000:40000], 4096, 0, 1)
    assert b.shape[1] > 0
    assert set(np.unique(b)) <= {-1.0, 1.0}
