"""HIP differentiable loop filter and pre-emphasis against the oracle
(oracle/or_lf.c): forward bit-exact (both sides use the same tanh built from
correctly rounded operations), gradients to a stated relative tolerance."""
import numpy as np
import pytest
import torch

import oracle_lib as ol
from test_lf_oracle import inband_grad_np, ntf, signal

pytestmark = pytest.mark.gpu
DSD128 = 5644800


def rel(x, ref):
    return float(np.max(np.abs(x - ref)) / max(np.max(np.abs(ref)), 1e-300))


@pytest.mark.parametrize("B,N,T", [(3, 1000, 63), (70, 777, 63), (2, 5000, 7), (1, 64, 128)])
def test_preemphasis_forward_bit_exact(gpu, oracle, B, N, T):
    from foo_dsd_trellis_amd import loopfilter as LF
    rng = np.random.default_rng(B * N + T)
    p = signal(B, N, 11)
    w = rng.standard_normal(T) * 0.1
    w[0] = 1.0
    u = LF.PreEmphasisFunction.apply(torch.from_numpy(p).cuda(), torch.from_numpy(w).cuda())
    ref = ol.pe_forward(oracle, w, p)
    assert np.array_equal(u.cpu().numpy().T, ref)


@pytest.mark.parametrize("T,N,E,B", [(10.0, 3 * 4096 + 77, 4096, 5), (1.0, 3 * 4096 + 77, 4096, 5),
                                     (10.0, 5000, 512, 70), (1.0, 16, 16, 1)])
def test_loop_filter_forward_bit_exact(gpu, oracle, T, N, E, B):
    from foo_dsd_trellis_amd import loopfilter as LF
    a, g = ntf(oracle)
    u = signal(B, N, 3).astype(np.float64)
    ud = torch.from_numpy(np.ascontiguousarray(u.T)).cuda()
    y = LF.LoopFilterFunction.apply(ud, torch.from_numpy(a).cuda(), torch.from_numpy(g).cuda(), T,
                                    E, 0)
    ref = ol.lf_forward(oracle, a, g, T, u)
    assert np.array_equal(y.cpu().numpy().T, ref)


@pytest.mark.parametrize("T,N,E,B,G", [(1.0, 2 * 4096 + 300, 4096, 4, 0), (1.0, 5000, 512, 3, 1),
                                       (1.0, 5000, 512, 66, 3), (10.0, 20, 16, 2, 1),
                                       (10.0, 20, 16, 2, 0)])
def test_gradients_match_oracle(gpu, oracle, T, N, E, B, G):
    """Pre-emphasis (63 taps) -> loop filter -> in-band loss; gradients of
    taps, a, g through the HIP adjoint vs the FP64 oracle adjoint."""
    from foo_dsd_trellis_amd import loopfilter as LF
    torch.manual_seed(0)
    p = signal(B, N, 4)
    pre = LF.FixedPreEmphasis(63)
    with torch.no_grad():
        pre.taps[2:] = 0.003 * torch.randn(61, dtype=torch.float64, device="cuda")
    sdm = LF.DiffSDM(rate=DSD128, temperature=T, checkpoint=E, slab_chunks=G)
    loss_fn = LF.AudioBandLoss(DSD128)
    xd = torch.from_numpy(p).cuda()
    loss = LF.train_step(pre, sdm, loss_fn, xd)

    w = pre.taps.detach().cpu().numpy()
    a, g = ntf(oracle)
    u = ol.pe_forward(oracle, w, p)
    y = ol.lf_forward(oracle, a, g, T, u)
    gu, ga, gg = ol.lf_backward(oracle, a, g, T, u, inband_grad_np(y, p.astype(np.float64), DSD128))
    gw = ol.pe_backward(oracle, 63, p, gu)
    assert rel(sdm.a.grad.cpu().numpy(), ga) < 1e-9
    assert rel(sdm.g.grad.cpu().numpy(), gg) < 1e-9
    assert rel(pre.taps.grad.cpu().numpy(), gw) < 1e-9


def test_full_size_gradient_vs_fp64_oracle(gpu, oracle):
    """64 x 262144 at DSD128 (temperature 1, where the loop contracts): the
    HIP gradient is within 1e-4 (relative, max-norm) of the FP64 oracle."""
    from foo_dsd_trellis_amd import loopfilter as LF, signals
    B, N = 64, 262144
    p = signals.multitone_noise(B, N, fs=DSD128, seed=9)
    pre = LF.FixedPreEmphasis(63)
    sdm = LF.DiffSDM(rate=DSD128, temperature=1.0)
    loss_fn = LF.AudioBandLoss(DSD128)
    LF.train_step(pre, sdm, loss_fn, torch.from_numpy(p).cuda())
    w = pre.taps.detach().cpu().numpy()
    a, g = ntf(oracle)
    u = ol.pe_forward(oracle, w, p)
    y = ol.lf_forward(oracle, a, g, 1.0, u)
    gu, ga, gg = ol.lf_backward(oracle, a, g, 1.0, u, inband_grad_np(y, p.astype(np.float64), DSD128))
    gw = ol.pe_backward(oracle, 63, p, gu)
    errs = (rel(sdm.a.grad.cpu().numpy(), ga), rel(sdm.g.grad.cpu().numpy(), gg),
            rel(pre.taps.grad.cpu().numpy(), gw))
    print("full-size gradient rel err (a, g, taps):", errs)
    assert max(errs) < 1e-4


@pytest.mark.parametrize("T,N,E,B,G,H", [(10.0, 5000, 512, 3, 3, 64), (10.0, 2 * 4096 + 300, 4096, 4, 0, 64),
                                         (1.0, 5000, 512, 66, 0, 512), (10.0, 4096, 4096, 2, 0, 128)])
def test_truncated_gradients_match_oracle(gpu, oracle, T, N, E, B, G, H):
    """Adjoint restarted every H samples (intervals in parallel) against the
    FP64 oracle's truncated adjoint: finite at temperature 10, where the full
    adjoint overflows over these lengths."""
    from foo_dsd_trellis_amd import loopfilter as LF
    torch.manual_seed(1)
    p = signal(B, N, 6)
    pre = LF.FixedPreEmphasis(63)
    with torch.no_grad():
        pre.taps[2:] = 0.003 * torch.randn(61, dtype=torch.float64, device="cuda")
    sdm = LF.DiffSDM(rate=DSD128, temperature=T, checkpoint=E, slab_chunks=G, trunc_span=H)
    LF.train_step(pre, sdm, LF.AudioBandLoss(DSD128), torch.from_numpy(p).cuda())
    w = pre.taps.detach().cpu().numpy()
    a, g = ntf(oracle)
    u = ol.pe_forward(oracle, w, p)
    y = ol.lf_forward(oracle, a, g, T, u)
    gu, ga, gg = ol.lf_backward(oracle, a, g, T, u, inband_grad_np(y, p.astype(np.float64), DSD128), H)
    gw = ol.pe_backward(oracle, 63, p, gu)
    assert np.isfinite(ga).all() and np.isfinite(gw).all()
    assert rel(sdm.a.grad.cpu().numpy(), ga) < 1e-9
    assert rel(sdm.g.grad.cpu().numpy(), gg) < 1e-9
    assert rel(pre.taps.grad.cpu().numpy(), gw) < 1e-9
