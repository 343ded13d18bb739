"""Batched DAS (gpu_hip_plan_*): tracks of uneven length, chunk-streamed,
all segments of all tracks in one trellis launch per chunk, against the
oracle's per-track FIR chain and per-chunk DAS with persistent contexts.
The device-side graph-captured chunk step is disabled in this build (its
first GPU run faulted, cause not found); plans run the host-orchestrated
batched step, which these tests exercise."""
import numpy as np
import pytest
import torch

import oracle_lib as ol

pytestmark = pytest.mark.gpu

FS_IN, FS_OUT = 44100, 5644800          # x128, 7 half-band stages
K, L, DEPTH = 16, 32, 8
CF = 1024                               # PCM frames per chunk


def tracks(seed=3):
    from foo_dsd_trellis_amd import signals
    lens = [2500, 1100, 3333]
    out = []
    for i, n in enumerate(lens):
        x = signals.bandlimited_gauss(n, fs=FS_IN, nch=2, seed=seed + i, fmax=18000.0)
        if i == 0:                       # a quiet stretch: boundary reposition
            x[:, 300:900] *= 1e-3
        out.append(np.ascontiguousarray(x))
    return out


def oracle_bits(oracle, pcm, fi, seg_len):
    from foo_dsd_trellis_amd import fir
    nch, n = pcm.shape
    y = np.stack([ol.fir_chain(oracle, FS_IN, FS_OUT, pcm[c], fir.PCM_FIR_GAIN) for c in range(nch)])
    R = FS_OUT // FS_IN
    st = ol.DasStream(oracle, fi, DEPTH, K, L, nch)
    parts = []
    for f0 in range(0, n, CF):
        f1 = min(n, f0 + CF)
        # DSD128 is below DSD512: the reference's seg0-first seeding
        parts.append(st.chunk(y[:, f0 * R:f1 * R], seg_len, 1, 1))
    return np.concatenate(parts, axis=1)


def gpu_bits(plan, t, nbits):
    w = plan.outputs[t].cpu().numpy().view(np.uint64)
    bits = np.unpackbits(w.view(np.uint8), axis=1, bitorder="little")[:, :nbits]
    return np.where(bits > 0, 1.0, -1.0).astype(np.float32)


def test_batch_matches_oracle(gpu, oracle):
    from foo_dsd_trellis_amd import batch, trellis
    fi = trellis.ntf_auto(FS_OUT)
    pcm = tracks()
    plan = batch.BatchPlan(FS_IN, FS_OUT, fi, K=K, L=L, depth=DEPTH, nch=2, chunk_frames=CF,
                           seg_len=8192, use_graph=False)
    plan.enqueue([torch.from_numpy(x).cuda() for x in pcm])
    first = plan.run(max_chunks=1)
    bits = plan.run()
    info = plan.info()
    assert info["chunks_done"] == info["chunks_total"] == 4
    for t, x in enumerate(pcm):
        ref = oracle_bits(oracle, x, fi, 8192)
        assert bits[t] == ref.shape[1], (t, bits[t], ref.shape)
        assert first[t] <= bits[t]
        assert np.array_equal(gpu_bits(plan, t, int(bits[t])), ref), "track %d" % t


def test_batch_autotuned(gpu, oracle):
    """Launch shape and segment length from the autotuner; the bits equal
    the oracle run with the segment length it chose."""
    from foo_dsd_trellis_amd import batch, trellis
    fi = trellis.ntf_auto(FS_OUT)
    pcm = tracks(seed=11)
    plan = batch.BatchPlan(FS_IN, FS_OUT, fi, K=K, L=L, depth=DEPTH, nch=2, chunk_frames=CF)
    plan.enqueue([torch.from_numpy(x).cuda() for x in pcm])
    bits = plan.run()
    tune = plan.tune()
    print("tuned:", tune)
    # a full chunk's length over a whole number of segments, none shorter
    # than 4 overlaps (64 L samples each)
    assert 4 * 64 * L <= tune["seg_len"] <= CF * (FS_OUT // FS_IN)
    assert tune["waves_per_group"] == 1 and tune["waves_per_cu"] >= 1 and tune["trial_msps"] > 0
    assert tune["use_graph"] is False          # the graph-captured step is disabled in this build
    for t, x in enumerate(pcm):
        ref = oracle_bits(oracle, x, fi, tune["seg_len"])
        assert np.array_equal(gpu_bits(plan, t, int(bits[t])), ref), "track %d" % t


@pytest.mark.skip(reason="device-side graph-captured chunk step (csrc/batch_kernel.hip + tr_run_queue): "
                         "its first GPU run ended in a GPU memory fault whose cause has not been shown; "
                         "the step is switched off in the library (kPlanLaunchEnabled = 0) and must not "
                         "run on a GPU until the cause is found and fixed")
def test_batch_device_step_graph(gpu, oracle):
    """The device-side chunk step captured in a hipGraph and replayed per
    chunk (persistent work-queue trellis launch, two waves per workgroup,
    four per CU), against the oracle with every segment seeded from the
    persistent context (the step's seeding)."""
    from foo_dsd_trellis_amd import batch, fir, trellis
    fi = trellis.ntf_auto(FS_OUT)
    pcm = tracks()
    plan = batch.BatchPlan(FS_IN, FS_OUT, fi, K=K, L=L, depth=DEPTH, nch=2, chunk_frames=CF,
                           seg_len=8192, waves_per_group=2, waves_per_cu=4, use_graph=True)
    plan.enqueue([torch.from_numpy(x).cuda() for x in pcm])
    bits = plan.run()
    assert plan.tune()["use_graph"] is True
    R = FS_OUT // FS_IN
    for t, x in enumerate(pcm):
        y = np.stack([ol.fir_chain(oracle, FS_IN, FS_OUT, x[c], fir.PCM_FIR_GAIN) for c in range(2)])
        st = ol.DasStream(oracle, fi, DEPTH, K, L, 2)
        ref = np.concatenate([st.chunk(y[:, f0 * R:min(x.shape[1], f0 + CF) * R], 8192, 0, 1)
                              for f0 in range(0, x.shape[1], CF)], axis=1)
        assert bits[t] == ref.shape[1]
        assert np.array_equal(gpu_bits(plan, t, int(bits[t])), ref), "track %d" % t
