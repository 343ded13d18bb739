#!/usr/bin/env python3
"""Benchmark of the PCM->DSD hot path on one MI355X.

Workload (one "step"): 30 s of stereo 44.1 kHz float32 PCM (seeded pink
noise at -12 dBFS RMS clipped to +-0.5 plus 7 tones 100 Hz..15 kHz at
-20 dBFS) -> eight 2x half-band FIR stages to DSD256 (11.2896 MHz, gain
0.354) -> DAS parallel trellis (CLANS-8, K=32, L=64, B=8, 65536-sample
segments, overlap 4096, quiet-boundary reposition) -> stitched 1-bit
streams.  The headline seeds as the reference does at DSD256 (seg0-first:
seg0 on the persistent context, seg1 its continuation, the 16-level
estimate from seg1's start seeding every later segment); the same step with
every segment seeded from the persistent state (the reference's DSD512+
mode) is timed beside it.  Inputs are resident in HBM before the timed
region; each step starts from fresh FIR delay lines and fresh trellis
contexts.

Prints one JSON line: Msamples/s of DSD output over both channels, with
x-realtime, the trellis kernel's roofline position (HIP-event timed on the
all-segments launch), the plain per-channel GPU form on a bounded sample,
and the reference's DAS path (the oracle's C restatement, same plan, FIR
included) timed on all of this host's cores.
"""
import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FS_IN = 44100
DSD256 = 11289600
SECONDS = 30
NCH = 2
K, L, B = 32, 64, 8
SEG_LEN = 65536
FP64_PEAK_TFLOPS = 78.6


def flops_per_sample(P, K_):
    """FP64 operations of one trellis step per active candidate: loop filter
    (P-1 multiply-subtract-add rows, P multiply-adds for v), the two child
    metrics and the +-1 state update: K * (5P + 6)."""
    return K_ * (5 * P + 6)


def traffic_per_input_sample():
    """HBM bytes per trellis input sample measured with rocprofv3 FETCH_SIZE
    and WRITE_SIZE (scaled by a copy-kernel calibration), recorded in
    profiles/tr_run_jobs_dsd256_das_hbm_traffic.md."""
    import re
    path = os.path.join(ROOT, "profiles", "tr_run_jobs_dsd256_das_hbm_traffic.md")
    try:
        m = re.search(r"traffic_bytes_per_input_sample:\s*([0-9.]+)", open(path).read())
        return float(m.group(1)) if m else None
    except OSError:
        return None


def host_cpu():
    """Cores this process may use and the CPU model string."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return n, os.cpu_count(), model


def cpu_baseline(pcm, seconds, seg0_first, threads):
    """The reference's DAS path on host threads, bounded: the first `seconds`
    of the headline's PCM through the oracle's FP64 FIR chain (one thread per
    channel) and the oracle's DAS with the same plan (65,536-sample
    segments, reposition, the given seeding; its phases run concurrently
    over channels x segments on `threads` threads, as the reference's
    threadpool does).  The oracle is the C restatement of src/fir.c,
    src/trellis.c and src/dsp_plugin.c's DAS (gcc -O2 -ffp-contract=off),
    bit-exact with the reference's own trellis.c; the reference's DAS
    driver itself needs Win32 + IPP and does not build here."""
    import ctypes as C
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    from foo_dsd_trellis_amd import fir
    O = ol.oracle()
    n_in = int(seconds * FS_IN)
    x = np.ascontiguousarray(pcm[:, :n_in])
    fi = O.ntf_auto(DSD256)
    O.lib.ora_das_set_threads.argtypes = [C.c_int]
    O.lib.ora_das_set_threads(threads)
    t0 = time.perf_counter()
    ys = [None] * NCH

    def fir_ch(c):
        ys[c] = ol.fir_chain(O, FS_IN, DSD256, x[c], fir.PCM_FIR_GAIN)

    th = [threading.Thread(target=fir_ch, args=(c,)) for c in range(NCH)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    t1 = time.perf_counter()
    out, n, sp, plan = ol.das(O, fi, B, K, L, np.stack(ys), seg_len=SEG_LEN, nseg=0,
                              seg0_first=1 if seg0_first else 0, reposition=1)
    t2 = time.perf_counter()
    O.lib.ora_das_set_threads(1)
    cores, nproc, model = host_cpu()
    wall = t2 - t0
    return {"value": round(NCH * n / wall / 1e6, 4), "unit": "Msamples/s",
            "x_realtime": round(seconds / wall, 4), "cores": threads, "nproc": nproc,
            "affinity_cores": cores, "cpu_model": model, "kind": "port",
            "seed_mode": "seg0-first" if seg0_first else "all-parallel",
            "fir_s": round(t1 - t0, 3), "das_s": round(t2 - t1, 3),
            "sample": "first %.2f s of the headline PCM (%d DSD256 samples per channel, %d segments): "
                      "oracle FP64 FIR chain (1 thread per channel) + oracle DAS, same plan, phases over "
                      "%d threads, %.1f s wall" % (seconds, n_in * (DSD256 // FS_IN), len(plan), threads, wall)}


def pmc_utilisation():
    """VALU issue and LDS utilisation of the trellis launch from the SQ
    counters committed in profiles/tr_run_jobs_dsd256_pmc.md (formulas
    there)."""
    import re
    path = os.path.join(ROOT, "profiles", "tr_run_jobs_dsd256_pmc.md")
    out = {}
    try:
        txt = open(path).read()
    except OSError:
        return out
    for key in ("valu_issue_util", "lds_util", "fp64_valu_share"):
        m = re.search(key + r":\s*([0-9.]+)", txt)
        if m:
            out[key] = float(m.group(1))
    return out


LF_TRUNC_SPAN = 64


def diff_loop_filter(steps, check_rows):
    """Differentiable loop filter (64 x 262144 at DSD128, 63-tap
    pre-emphasis, CLANS-6 CRFB, tanh(v/0.1), in-band rFFT loss, state
    checkpoint every 4096): HIP forward + adjoint throughput, and the
    gradient error against the FP64 oracle on `check_rows` rows.  At
    temperature 10 the relaxed loop is chaotic (a perturbation grows about
    1.7x per sample) and the full-horizon adjoint overflows, so the adjoint
    is truncated: it restarts from zero every 64 samples (the state at each
    interval start held constant); the oracle's gradient is truncated the
    same way.  The full-horizon adjoint is checked at temperature 1, where
    the loop contracts."""
    import torch
    from foo_dsd_trellis_amd import loopfilter as LF, signals
    B, N, fs = 64, 262144, 5644800
    p = torch.from_numpy(signals.multitone_noise(B, N, fs=fs, seed=9)).cuda()
    pre = LF.FixedPreEmphasis(63)
    sdm = LF.DiffSDM(rate=fs, temperature=10.0, trunc_span=LF_TRUNC_SPAN)
    loss_fn = LF.AudioBandLoss(fs)
    for _ in range(2):
        LF.train_step(pre, sdm, loss_fn, p)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        LF.train_step(pre, sdm, loss_fn, p)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    finite = bool(torch.isfinite(sdm.a.grad).all() and torch.isfinite(pre.taps.grad).all())
    if not finite:
        raise RuntimeError("differentiable loop filter: non-finite gradient at temperature 10")
    out = {"value": round(B * N / (ms / 1e3) / 1e6, 3), "unit": "Msamples/s (fwd+bwd)",
           "ms_per_step": round(ms, 3), "temperature": 10.0, "grad_finite": finite,
           "adjoint": "truncated: restarts from zero every %d samples" % LF_TRUNC_SPAN,
           "config": "64 x 262144 f32 at DSD128, 63-tap pre-emphasis, %s CRFB, tanh(v/0.1), "
                     "checkpoint 4096, in-band (0-20 kHz) error power by rFFT" % sdm.ntf_name}
    if check_rows > 0:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as ol
        O = ol.oracle()
        x = p[:check_rows].contiguous()
        xn = x.cpu().numpy()
        wk = ol.inband_weights(N, fs)
        c = np.full(N // 2 + 1, 2.0)
        c[0] = c[-1] = 1.0
        rel = lambda v, r: float(np.max(np.abs(v - r)) / max(np.max(np.abs(r)), 1e-300))
        for T, span, key in ((10.0, LF_TRUNC_SPAN, "grad_rel_err_vs_fp64"), (1.0, 0, "grad_rel_err_vs_fp64_t1_full")):
            pre1 = LF.FixedPreEmphasis(63)
            sdm1 = LF.DiffSDM(rate=fs, temperature=T, trunc_span=span)
            LF.train_step(pre1, sdm1, LF.AudioBandLoss(fs), x)
            w = pre1.taps.detach().cpu().numpy()
            a = sdm1.a.detach().cpu().numpy()
            g = sdm1.g.detach().cpu().numpy()
            u = ol.pe_forward(O, w, xn)
            y = ol.lf_forward(O, a, g, T, u)
            E = np.fft.rfft(y - xn.astype(np.float64), axis=1)
            gy = 2.0 * np.fft.irfft(wk * E / c, N, axis=1) / (N * check_rows)
            gu, ga, gg = ol.lf_backward(O, a, g, T, u, gy, span)
            gw = ol.pe_backward(O, 63, xn, gu)
            out[key] = max(rel(sdm1.a.grad.cpu().numpy(), ga), rel(sdm1.g.grad.cpu().numpy(), gg),
                           rel(pre1.taps.grad.cpu().numpy(), gw))
        out["grad_check"] = "%d rows x %d, max-norm relative error of dL/da, dL/dg, dL/dtaps vs the FP64 " \
                            "oracle adjoint: at temperature 10 both truncated every %d samples; at " \
                            "temperature 1 over the full horizon" % (check_rows, N, LF_TRUNC_SPAN)
    return out


def cpu_das_rate(pcm, fs_in, fs_out, seg_len, seg0_first, threads):
    """The oracle's FP64 FIR chain (one thread per channel) and DAS (phases
    over `threads` threads) on host PCM [nch][n]: output Msamples/s (all
    channels) and the wall time."""
    import ctypes as C
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    from foo_dsd_trellis_amd import fir
    O = ol.oracle()
    nch = pcm.shape[0]
    fi = O.ntf_auto(fs_out)
    O.lib.ora_das_set_threads.argtypes = [C.c_int]
    O.lib.ora_das_set_threads(threads)
    t0 = time.perf_counter()
    ys = [None] * nch

    def fir_ch(c):
        ys[c] = ol.fir_chain(O, fs_in, fs_out, np.ascontiguousarray(pcm[c]), fir.PCM_FIR_GAIN)

    th = [threading.Thread(target=fir_ch, args=(c,)) for c in range(nch)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    out, n, sp, plan = ol.das(O, fi, B, K, L, np.stack(ys), seg_len=seg_len, nseg=0,
                              seg0_first=1 if seg0_first else 0, reposition=1)
    wall = time.perf_counter() - t0
    O.lib.ora_das_set_threads(1)
    return nch * n / wall / 1e6, wall, len(plan)


def batched_uneven(chunks, cf, plain_tracks, cpu_frames=0, cpu_threads=1):
    """The batched uneven workload, bounded: 16 tracks of 6-channel 48 kHz
    PCM (lengths drawn from U[5, 120] s, seeded band-limited Gaussian at
    -14 dBFS RMS) -> 9 half-band stages to DSD512/48 (24.576 MHz) -> DAS
    trellis K=32, L=64 with every segment of every channel and track in one
    launch per chunk.  The tracks are enqueued at their full lengths (the
    autotuner's segment length is for the whole workload); the first
    `chunks` chunks of `cf` frames are run (every track is longer than
    that), and the first chunk (tuning, set-up) is not timed.  `plain_tracks` tracks are also run one DAS launch per track on
    the same first chunk, for the per-track form's throughput."""
    import torch
    from foo_dsd_trellis_amd import batch, das, fir, signals, trellis
    fs_in, fs_out, nch = 48000, 24576000, 6
    R = fs_out // fs_in
    lens = signals.uneven_tracks(16, nch, fs_in, seed=4)
    n = chunks * cf
    # every track is enqueued at its full length, so the autotuner sizes the
    # segments for the whole workload; only the first n frames (the chunks
    # that are run) carry content, the rest is silence that is never reached
    pcm = []
    for t, L_ in enumerate(lens):
        x = torch.zeros((nch, L_), dtype=torch.float32, device="cuda")
        x[:, :min(L_, n)] = torch.from_numpy(signals.bandlimited_gauss(min(L_, n), fs=fs_in, nch=nch,
                                                                       seed=100 + t)).cuda()
        pcm.append(x)
    fi = trellis.ntf_auto(fs_out)
    plan = batch.BatchPlan(fs_in, fs_out, fi, K=K, L=L, depth=B, nch=nch, chunk_frames=cf)
    plan.enqueue(pcm)
    plan.run(max_chunks=1)
    torch.cuda.synchronize()
    b0 = plan.run(max_chunks=0)
    t0 = time.perf_counter()
    b1 = plan.run(max_chunks=chunks - 1)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    out_bits = int((b1 - b0).sum()) * nch
    tune = plan.tune()
    rate = out_bits / dt
    total = sum(lens) * R * nch                  # output samples of the whole workload
    res = {"value": round(rate / 1e6, 3), "unit": "Msamples/s",
           "timed_fraction": round(out_bits / total, 5),
           "full_job_s_extrapolated": round(total / rate, 2),
           "extrapolation": "whole workload (all tracks to their ends) at the timed chunks' rate; every "
                            "timed chunk has all 16 tracks live, later chunks have fewer tracks and "
                            "fewer jobs per launch, so this is the rate while the batch is full",
           "x_realtime_per_track": round((chunks - 1) * cf / fs_in / dt * len(pcm), 3),
           "timed_chunks": chunks - 1, "chunk_frames": cf, "tracks": len(pcm), "channels": nch,
           "track_seconds": [round(L_ / fs_in, 2) for L_ in lens], "tune": tune,
           "mode": "host-orchestrated batched step (device-side graph step disabled in this build)",
           "config": "16 tracks x 6 ch 48 kHz f32 -> DSD512/48 (24.576 MHz, 9 stages) -> DAS K=32 L=64 B=8, "
                     "all segments of all tracks in one trellis launch per chunk"}
    if plain_tracks > 0:
        # per-track form on the first chunk of plain_tracks tracks: FIR, then one DAS per track,
        # with the shortest segments the reference allows (4 overlaps), which give one track's
        # chunk the most jobs per launch
        plain_len = 4 * 64 * L
        ch = fir.FirChain(fs_in, fs_out, fir.PCM_FIR_GAIN, nch)
        d_f = torch.empty((nch, cf * R), dtype=torch.float64, device="cuda")
        words = cf * R // 64 + 2
        d_o = torch.zeros((nch, words), dtype=torch.int64, device="cuda")
        tot, tt = 0, 0.0
        for t in range(plain_tracks):
            x = pcm[t][:, :cf].contiguous()
            ch.reset()
            ch.process_dev(x.data_ptr(), cf, cf, d_f.data_ptr(), cf * R)
            D = das.Das(fi, B, K, L, nch, seg_len=plain_len, seg0_first=False, reposition=True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            nout = D.run_dev(d_f.data_ptr(), cf * R, cf * R, d_o.data_ptr(), words)
            torch.cuda.synchronize()
            tt += time.perf_counter() - t0
            tot += nout * nch
        res["plain_per_track"] = {"value": round(tot / tt / 1e6, 3), "unit": "Msamples/s",
                                  "seg_len": plain_len,
                                  "sample": "first chunk of %d tracks, one DAS launch per track "
                                            "(segments of 4 overlaps, the most jobs per launch)" % plain_tracks}
    if cpu_frames > 0:
        x = pcm[0][:, :cpu_frames].cpu().numpy().astype(np.float32)
        v, wall, nseg = cpu_das_rate(x, fs_in, fs_out, tune["seg_len"], False, cpu_threads)
        cores, nproc, model = host_cpu()
        res["cpu_baseline"] = {"value": round(v, 4), "unit": "Msamples/s", "cores": cpu_threads,
                               "nproc": nproc, "cpu_model": model, "kind": "port",
                               "sample": "first %d frames of track 0 (%d ch, %d segments per channel of %d): "
                                         "oracle FP64 FIR chain + oracle DAS, all-parallel seeding as at "
                                         "DSD512, %.1f s wall" % (cpu_frames, nch, nseg, tune["seg_len"], wall)}
    return res


def streaming_room(chunks, seg_len):
    """The streaming workload: 6-channel 44.1 kHz float32 PCM (seeded pink
    noise plus tones, as the headline workload, one seed per channel) in
    chunks of 8192 frames -> eight half-band stages to DSD256 (2,097,152
    samples per chunk per channel) -> UPOLS convolution with a seeded
    synthetic room IR per channel (2^20 taps, exponentially decaying Gaussian
    noise, RT60 0.4 s, unit energy; partitions of 8192, 128 partitions, FFT
    16384, complex FP32 spectra) -> DAS trellis K=32, L=64, B=8 over all
    channels, seeded as the reference seeds at DSD256.  Chunk latency is the
    wall time of one chunk call (FIR, convolution, DAS, bits in HBM), with
    the median host-side timeline of the calls (ms from the call's start;
    zero for points the seeding does not pass); the first
    `warm` chunks are not timed (set-up, and the clock ramp of a GPU that
    was idle)."""
    import torch
    from foo_dsd_trellis_amd import signals, stream, trellis
    nch, cf = 6, 8192
    ir = signals.room_ir(1 << 20, fs=DSD256, nch=nch, rt60=0.4, seed=5)
    fi = trellis.ntf_auto(DSD256)
    S = stream.Stream(FS_IN, DSD256, nch, cf, ir=ir, P=8192, ntf_index=fi, K=K, L=L, depth=B,
                      seg_len=seg_len)
    warm = 4
    pcm = signals.pink_multitone(cf * (chunks + warm), fs=FS_IN, nch=nch, seed=7)
    d_pcm = torch.from_numpy(np.ascontiguousarray(pcm)).cuda()
    torch.cuda.synchronize()
    for c in range(warm):
        S.process(d_pcm[:, c * cf:(c + 1) * cf])
    torch.cuda.synchronize()
    lat, tot, tls = [], 0, []
    for c in range(warm, chunks + warm):
        t0 = time.perf_counter()
        tot += S.process(d_pcm[:, c * cf:(c + 1) * cf])
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t0)
        tls.append(S.last_timeline())
    lat_ms = np.array(lat) * 1000.0
    wall = float(np.sum(lat))
    tl_names = ["das_start", "seg0_done", "estimate_start", "last_published", "estimate_joined",
                "trellis_read", "das_done"]
    tl_p50 = {k: round(float(np.median([t[i] for t in tls])), 3) for i, k in enumerate(tl_names)}
    return {"value": round(tot * nch / wall / 1e6, 3), "unit": "Msamples/s",
            "timeline_ms_p50": tl_p50,
            "x_realtime": round(chunks * cf / FS_IN / wall, 3),
            "chunk_ms_p50": round(float(np.percentile(lat_ms, 50)), 3),
            "chunk_ms_p99": round(float(np.percentile(lat_ms, 99)), 3),
            "chunk_ms_max": round(float(lat_ms.max()), 3),
            "chunk_audio_ms": round(cf / FS_IN * 1000.0, 3), "timed_chunks": chunks, "untimed_chunks": warm,
            "seg_len": seg_len, "seed_mode": "seg0-first" if S.seg0_first else "all-parallel",
            "config": "6 ch 44.1 kHz f32, 8192-frame chunks -> DSD256 (8 stages) -> UPOLS 2^20-tap "
                      "room IR (P=8192, c-f32) -> DAS K=32 L=64 B=8 seeded as the reference at DSD256, one stream, chunks in order"}


def next_rows(d_bits, n_bits, seconds):
    """The rows after the hot path, on data of the headline workload:
    PreCorr (greedy SDM, one sequential recurrence per channel) over 6
    channels of DSD256-rate input; the same-rate DSD256 re-encode of the
    headline's own output bits (the engine's boxcar of 64 taps and path
    settings CLANS-6, K=2, L=16, B=4), streamed in 2^20-sample chunks with
    2,048-sample segments (more segments: lower chunk latency) and offline
    in one 2^24-sample chunk with 4,096-sample segments; DoP
    packing and unpacking of all of those bits (HBM-bound byte work)."""
    import torch
    from foo_dsd_trellis_amd import dop, precorr, signals, stream, trellis
    res = {}
    # PreCorr: 6 channels x `seconds` at DSD256
    n = int(seconds * DSD256)
    t = torch.arange(n, dtype=torch.float64, device="cuda") / DSD256
    x = torch.stack([0.4 * torch.sin(2 * np.pi * (997.0 + 200 * c) * t) for c in range(6)]).contiguous()
    fi = trellis.ntf_auto(DSD256)
    pc = precorr.PreCorr(fi, 6, 0.0)
    pc.process(x[:, :65536])
    pc.reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pc.process(x)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    res["precorr"] = {"value": round(6 * n / dt / 1e6, 3), "unit": "Msamples/s", "channels": 6,
                      "samples_per_channel": n, "x_realtime": round(seconds / dt, 3),
                      "note": "greedy SDM is sequential per channel: one wave per channel"}
    del x, t
    # same-rate DSD256 re-encode of the headline output: streamed in chunks
    # of 2^20 samples (first `seconds` per channel), and converted offline in
    # one chunk of 2^24 samples (1.49 s); K=2 runs one lane per segment job
    fi6 = trellis.ntf_lookup(4, DSD256)            # clans-6 (the path table's DSD256 same-rate filter)
    for name, cf, m, seg in (("dsd_same_rate_reencode", 1 << 20, min(n_bits, n), 2048),
                             ("dsd_same_rate_reencode_offline", 1 << 24, min(n_bits, 1 << 24), 4096)):
        m = max(cf, m // cf * cf)
        if m > n_bits:
            continue
        words = m // 64
        src = d_bits[:, :words].contiguous()
        S = stream.Stream(DSD256, DSD256, 2, cf, ntf_index=fi6, K=2, L=16, depth=4, seg_len=seg,
                          gain=float(np.float32(0.708)), input=stream.IN_DSD)
        S.process_dsd(src[:, :cf // 64], cf)
        S.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tot = 0
        for k in range(m // cf):
            tot += S.process_dsd(src[:, k * cf // 64:(k + 1) * cf // 64], cf)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        res[name] = {"value": round(2 * tot / dt / 1e6, 3), "unit": "Msamples/s",
                     "x_realtime": round(m / DSD256 / dt, 3), "channels": 2, "chunk": cf,
                     "config": "DSD256 -> DSD256, boxcar 64, CLANS-6 K=2 L=16 B=4, DAS %d-sample segments" % seg}
        del S, src
    # DoP pack (i24, interleaved) and unpack of the whole headline output
    words = (n_bits // 64) // 1024 * 1024
    src = d_bits[:, :words].contiguous()
    nb = words * 64
    out, _ = dop.pack_i24(src, nb, 0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        out, _ = dop.pack_i24(src, nb, 0)
    torch.cuda.synchronize()
    dp = (time.perf_counter() - t0) / 5
    pcm = dop.pack(src, nb)
    back = dop.unpack(pcm)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        back = dop.unpack(pcm)
    torch.cuda.synchronize()
    du = (time.perf_counter() - t0) / 5
    ok = bool(torch.equal(back[:, :words], src))
    res["dop"] = {"pack_i24_GBps": round((src.numel() * 8 + out.numel()) / dp / 1e9, 2),
                  "unpack_GBps": round((pcm.numel() * 4 + back.numel() * 8) / du / 1e9, 2),
                  "roundtrip_exact": ok, "frames_per_channel": nb // 16,
                  "note": "bytes read + written per call over wall time; HBM peak 8000 GB/s"}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-audio-seconds", type=float, default=1.0,
                    help="audio seconds of the headline workload for the cpu_baseline leg (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: every core this process may use")
    ap.add_argument("--plain-seconds", type=float, default=0.05,
                    help="audio seconds per channel for the plain per-channel GPU form (0: skip)")
    ap.add_argument("--seg-len", type=int, default=SEG_LEN)
    ap.add_argument("--seed-mode", choices=["both", "seg0-first", "all-parallel"], default="both",
                    help="headline seeding (seg0-first is the reference's at DSD256); both: "
                         "all-parallel is timed beside it")
    ap.add_argument("--feed-waves", type=int, default=-1,
                    help="seg0-first: resident waves per CU of the host-fed launch (-1: library default)")
    ap.add_argument("--lf-steps", type=int, default=5,
                    help="differentiable loop filter fwd+bwd steps (0: skip)")
    ap.add_argument("--lf-check-rows", type=int, default=2)
    ap.add_argument("--batch-chunks", type=int, default=3,
                    help="chunks of the batched uneven workload (first untimed; 0: skip)")
    ap.add_argument("--batch-chunk-frames", type=int, default=12000)
    ap.add_argument("--batch-plain-tracks", type=int, default=2)
    ap.add_argument("--batch-cpu-frames", type=int, default=4000,
                    help="PCM frames of one track for the batched workload's CPU baseline (0: skip)")
    ap.add_argument("--stream-chunks", type=int, default=30,
                    help="timed chunks of the streaming room-correction workload (0: skip)")
    ap.add_argument("--stream-seg-len", type=int, default=8192)
    ap.add_argument("--next-seconds", type=float, default=0.5,
                    help="audio seconds per channel for the PreCorr / same-rate / DoP legs (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not supported "
              "(the path would shard whole channels/tracks across GPUs)" % args.gpus)
        sys.exit(2)

    import torch
    from foo_dsd_trellis_amd import das, fir, signals, trellis

    dev = torch.device("cuda:0")
    n_in = FS_IN * SECONDS
    F = n_in * (DSD256 // FS_IN)
    x = signals.pink_multitone(n_in, fs=FS_IN, nch=NCH, seed=2)
    d_x = torch.from_numpy(x).to(dev)
    d_f = torch.empty((NCH, F), dtype=torch.float64, device=dev)
    words = F // 64 + 1
    d_bits = torch.zeros((NCH, words), dtype=torch.int64, device=dev)

    chain = fir.FirChain(FS_IN, DSD256, fir.PCM_FIR_GAIN, NCH)
    fi = trellis.ntf_auto(DSD256)
    P = trellis.ntf_row(fi)["order"]
    fir_ev = []

    def run_mode(seg0_first):
        D = das.Das(fi, B, K, L, NCH, seg_len=args.seg_len, seg0_first=seg0_first, reposition=True)
        D.timing(True)
        if args.feed_waves >= 0:
            D.feed_waves(args.feed_waves)

        def step(timed=False):
            chain.reset()
            D.reset()
            if timed:
                ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                ev[0].record()
            chain.process_dev(d_x.data_ptr(), n_in, n_in, d_f.data_ptr(), F)
            if timed:
                ev[1].record()
                fir_ev.append(ev)
            return D.run_dev(d_f.data_ptr(), F, F, d_bits.data_ptr(), words)

        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        D.timing_read(clear=True)
        t0 = time.perf_counter()
        n_out = 0
        for _ in range(args.steps):
            n_out = step(timed=True)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        k_ms, k_launches, k_samples = D.timing_read(clear=True)
        ms_step = (t1 - t0) * 1000.0 / args.steps
        nseg, slow, _ = D.last_info(0)
        r = {"value": round(NCH * n_out / (ms_step / 1000.0) / 1e6, 3), "unit": "Msamples/s",
             "x_realtime": round(SECONDS / (ms_step / 1000.0), 3), "ms_per_step": round(ms_step, 3),
             "out_samples_per_channel": n_out, "segments_per_channel": nseg, "literal_prune_steps": slow,
             "trellis_ms_per_step": round(k_ms / args.steps, 3),
             "trellis_launches_per_step": k_launches // args.steps,
             "trellis_inputs_per_step": k_samples // args.steps}
        if seg0_first:
            est_ms, est_n, redo, nb = D.last_estimate()
            r["timeline_ms"] = dict(zip(["seg0_done", "estimate_start", "last_published", "estimate_joined",
                                          "trellis_read", "run_done"], D.last_timeline()))
            r["estimator"] = {"host_chain_ms": round(est_ms, 3), "samples_all_channels": est_n,
                              "ns_per_sample_per_chain": round(est_ms * 1e6 / max(1, est_n // NCH), 3),
                              "reference_form_blocks": redo, "batches": nb,
                              "note": "two host threads per channel (index guess; its check against "
                                      "the reference-order sum, which owns the exact state); every "
                                      "segment runs in one host-fed persistent launch whose wave "
                                      "starts as soon as the estimate for that segment is published"}
        return r, n_out

    modes = ["seg0-first", "all-parallel"] if args.seed_mode == "both" else [args.seed_mode]
    res = {}
    for m in modes:
        fir_ev.clear()
        res[m], n_m = run_mode(m == "seg0-first")
        if m == modes[0]:
            n_out = n_m
            fir_ms = sum(a.elapsed_time(b) for a, b in fir_ev) / max(1, len(fir_ev))
            bits0 = d_bits.clone()
    head = res[modes[0]]
    value, ms_step, xrt = head["value"], head["ms_per_step"], head["x_realtime"]

    # FIR chain (8 half-band stages; adjacent 63-tap stages fused into one
    # pass, so only every second stage's output reaches HBM): bytes = the
    # float input + each pass's double output + the double inputs of passes
    # 2..4; compulsory bytes = the float input + the final double output.
    # FP64 work: 63 taps per output pair (dense polyphase, taps in the
    # reference's form) = 63 flops per output of every stage.
    nst = int(round(np.log2(DSD256 // FS_IN)))
    written = [2 ** (2 * (k + 1)) for k in range(nst // 2)]          # pass outputs per input sample
    fir_bytes = NCH * n_in * (4 + 8 * sum(written) + 8 * sum(written[:-1]))
    fir_min = NCH * n_in * (4 + 8 * 2 ** nst)
    fir_flops = NCH * n_in * 63 * sum(2 ** (k + 1) for k in range(nst))
    fir_roof = {"bound": "hbm", "ms": round(fir_ms, 3), "bytes": int(fir_bytes),
                "achieved_GBps": round(fir_bytes / (fir_ms / 1e3) / 1e9, 1) if fir_ms > 0 else None,
                "peak_GBps": 8000.0,
                "frac": round(fir_bytes / (fir_ms / 1e3) / 8e12, 4) if fir_ms > 0 else None,
                "compulsory_bytes": int(fir_min),
                "frac_compulsory": round(fir_min / (fir_ms / 1e3) / 8e12, 4) if fir_ms > 0 else None,
                "fp64_tflops": round(fir_flops / (fir_ms / 1e3) / 1e12, 2) if fir_ms > 0 else None,
                "fp64_frac": round(fir_flops / (fir_ms / 1e3) / 1e12 / FP64_PEAK_TFLOPS, 4) if fir_ms > 0 else None,
                "stages": nst, "passes": nst // 2,
                "note": "whole chain per step on the default stream, HIP events; adjacent 63-tap stages fused "
                        "(the intermediate stream stays in LDS)"}

    # trellis roofline on the one-launch (all-segments) form: algorithmic
    # FP64 flops = K (5P + 6) per output sample, outputs of both channels
    fps = flops_per_sample(P, K)
    roof = None
    ap_ = res.get("all-parallel")
    if ap_ is not None and ap_["trellis_launches_per_step"] == 1:
        launch_ms = ap_["trellis_ms_per_step"]
        outs = NCH * ap_["out_samples_per_channel"]
        achieved = outs * fps / (launch_ms / 1000.0) / 1e12
        tb = traffic_per_input_sample()
        roof = {"bound": "valu", "achieved": round(achieved, 4), "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 6),
                "traffic": int(tb * ap_["trellis_inputs_per_step"]) if tb else None,
                "kernel": "tr_run_jobs<8,1>", "avg_launch_ms": launch_ms,
                "flops_per_output_sample": fps, "outputs_per_launch": outs,
                "inputs_per_launch": ap_["trellis_inputs_per_step"],
                "peak_note": "MI355X FP64 vector peak (no MFMA: the trellis is not a contraction)",
                "measured_on": "all-parallel launch (every segment of both channels in one launch); "
                               "the seg0-first headline is bound by its sequential estimate",
                "traffic_source": "profiles/tr_run_jobs_dsd256_das_hbm_traffic.md (FETCH_SIZE/WRITE_SIZE)"}
        roof.update(pmc_utilisation())

    plain = None
    if args.plain_seconds > 0:
        npl = int(args.plain_seconds * DSD256)
        xs = d_f[:, :npl].cpu().numpy()
        pool = trellis.SdmPool(fi, B, K, L, 0.0, NCH)
        d_in = torch.from_numpy(np.ascontiguousarray(xs)).to(dev)
        d_o = torch.zeros((NCH, npl // 64 + 2), dtype=torch.int64, device=dev)
        torch.cuda.synchronize()
        tp0 = time.perf_counter()
        kept, _ = pool.run_jobs(list(range(NCH)), [d_in[c].data_ptr() for c in range(NCH)],
                                [npl] * NCH, [d_o[c].data_ptr() for c in range(NCH)])
        torch.cuda.synchronize()
        tp = time.perf_counter() - tp0
        plain = {"value": round(float(sum(kept)) / tp / 1e6, 4), "unit": "Msamples/s",
                 "sample": "sequential trellis, one context per channel, %.3f s of DSD256 audio "
                           "per channel (%d samples)" % (args.plain_seconds, npl)}

    cpu = None
    if args.cpu_audio_seconds > 0:
        cores = host_cpu()[0]
        thr = args.cpu_threads if args.cpu_threads > 0 else cores
        cpu = cpu_baseline(x, args.cpu_audio_seconds, modes[0] == "seg0-first", thr)

    lf = diff_loop_filter(args.lf_steps, args.lf_check_rows) if args.lf_steps > 0 else None
    bu = (batched_uneven(args.batch_chunks, args.batch_chunk_frames, args.batch_plain_tracks,
                         args.batch_cpu_frames, args.cpu_threads if args.cpu_threads > 0 else host_cpu()[0])
          if args.batch_chunks > 1 else None)
    sr = streaming_room(args.stream_chunks, args.stream_seg_len) if args.stream_chunks > 0 else None
    nr = next_rows(bits0, n_out, args.next_seconds) if args.next_seconds > 0 else None

    line = {
        "metric": "Trellis SDM Msamples/s & xrealtime at DSD256 K=32, batched vs plain",
        "value": value, "unit": "Msamples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "x_realtime": xrt,
        "config": {"workload": "PCM 44.1k stereo 30 s -> 8x2 half-band FIR -> DSD256 DAS trellis",
                   "fs_out": DSD256, "seconds": SECONDS, "channels": NCH, "K": K, "L": L, "B": B,
                   "ntf": trellis.ntf_row(fi)["name"], "seg_len": args.seg_len,
                   "overlap": 64 * L, "segments_per_channel": head["segments_per_channel"],
                   "seed_mode": modes[0],
                   "literal_prune_steps": head["literal_prune_steps"], "out_samples_per_channel": n_out},
        "seed_modes": res,
        "roofline": roof,
        "fir_roofline": fir_roof,
        "plain_gpu": plain,
        "cpu_baseline": cpu,
        "diff_loop_filter": lf,
        "batched_uneven": bu,
        "streaming_room": sr,
        "next_rows": nr,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
