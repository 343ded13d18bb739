"""Differentiable sigma-delta loop filter for training pre-emphasis FIR taps
and NTF coefficients with PyTorch (ROCm), on hand-written HIP kernels.

Mirrors the reference training pipeline's modules
(training_pipeline/diff_sdm.py:42-168, model_pre_sdm.py:22-53):

  FixedPreEmphasis(num_taps)   causal FIR y[n] = sum_k taps[k] x[n-k]
  DiffSDM(ntf, temperature)    CRFB loop filter of src/trellis.c:37-151 with
                               the quantizer relaxed to tanh(v * temperature)
  AudioBandLoss(dsd_rate)      in-band (0..20 kHz) error power via rFFT

Forward and adjoint of the loop filter and the pre-emphasis run in HIP
(libdsdhip.so); the loop filter keeps its state only every `checkpoint`
samples (default 4096) and recomputes the rest in backward.  Tensors live on
the GPU in float64; the kernels' native layout is time-major (N, B), and the
(B, N) tensors these modules return are transposed views of it.

The temperature is 1/tau (tanh(v/0.1) is temperature 10).  With the
reference's NTF coefficients the relaxed loop is chaotic at high
temperature (perturbations grow by about 1.4x per sample at temperature
10 with CLANS-6 at DSD128), so gradients over long sequences grow without
bound there; the reference anneals the temperature upward from 1
(training_pipeline/train_pre_sdm.py:130-140), where the loop contracts.
"""
import numpy as np
import torch

from . import _lib, trellis

DEFAULT_CHECKPOINT = 4096


def _stream():
    return torch.cuda.current_stream().cuda_stream


def _time_major(x):
    """(B, N) tensor -> contiguous (N, B) float64 on its device (no copy when
    x already is the transposed view of one)."""
    xt = x.t()
    if xt.dtype == torch.float64 and xt.is_contiguous():
        return xt
    return xt.to(torch.float64).contiguous()


class LoopFilterFunction(torch.autograd.Function):
    """y = loop_filter(u; a, g).  u_nb: (N, B) float64 contiguous; a, g: (P,)
    float64; returns y as (N, B)."""

    @staticmethod
    def forward(ctx, u_nb, a, g, temperature, checkpoint, slab_chunks, trunc_span=0):
        L = _lib.lib()
        N, B = u_nb.shape
        P = a.numel()
        a = a.detach().contiguous()
        g = g.detach().contiguous()
        nck = L.dsdhip_lf_checkpoints(N, checkpoint)
        y = torch.empty_like(u_nb)
        ck = torch.empty((max(nck, 1), P, B), dtype=torch.float64, device=u_nb.device)
        _lib.check(L.dsdhip_lf_forward(a.data_ptr(), g.data_ptr(), P, float(temperature),
                                       u_nb.data_ptr(), B, N, checkpoint, y.data_ptr(),
                                       ck.data_ptr(), _stream()), "dsdhip_lf_forward")
        ctx.save_for_backward(u_nb, a, g, ck)
        ctx.cfg = (float(temperature), int(checkpoint), int(slab_chunks), int(trunc_span))
        return y

    @staticmethod
    def backward(ctx, gy):
        L = _lib.lib()
        u_nb, a, g, ck = ctx.saved_tensors
        T, E, G, tr = ctx.cfg
        N, B = u_nb.shape
        P = a.numel()
        gy = gy.to(torch.float64).contiguous()
        ws_bytes = L.dsdhip_lf_workspace_bytes(B, N, P, E, G)
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=u_nb.device)
        gu = torch.empty_like(u_nb)
        ga = torch.empty(P, dtype=torch.float64, device=u_nb.device)
        gg = torch.empty(P, dtype=torch.float64, device=u_nb.device)
        _lib.check(L.dsdhip_lf_backward_ex(a.data_ptr(), g.data_ptr(), P, T, u_nb.data_ptr(), B, N,
                                           E, ck.data_ptr(), gy.data_ptr(), gu.data_ptr(),
                                           ga.data_ptr(), gg.data_ptr(), ws.data_ptr(), ws_bytes, G,
                                           tr, _stream()), "dsdhip_lf_backward_ex")
        return gu, ga, gg, None, None, None, None


class PreEmphasisFunction(torch.autograd.Function):
    """u = causal FIR of p with taps w.  p: (B, N) float32 contiguous; w:
    (T,) float64; returns u as (N, B) float64.  No gradient w.r.t. p."""

    @staticmethod
    def forward(ctx, p, w):
        L = _lib.lib()
        B, N = p.shape
        w = w.detach().contiguous()
        u = torch.empty((N, B), dtype=torch.float64, device=p.device)
        _lib.check(L.dsdhip_pe_forward(w.data_ptr(), w.numel(), p.data_ptr(), B, N, u.data_ptr(),
                                       _stream()), "dsdhip_pe_forward")
        ctx.save_for_backward(p)
        ctx.ntaps = w.numel()
        return u

    @staticmethod
    def backward(ctx, gu):
        L = _lib.lib()
        (p,) = ctx.saved_tensors
        B, N = p.shape
        T = ctx.ntaps
        gu = gu.to(torch.float64).contiguous()
        ws_bytes = L.dsdhip_pe_workspace_bytes(B, N, T)
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=p.device)
        gw = torch.empty(T, dtype=torch.float64, device=p.device)
        _lib.check(L.dsdhip_pe_backward(p.data_ptr(), B, N, T, gu.data_ptr(), gw.data_ptr(),
                                        ws.data_ptr(), ws_bytes, _stream()), "dsdhip_pe_backward")
        return None, gw


class FixedPreEmphasis(torch.nn.Module):
    """Learned causal FIR (model_pre_sdm.py:22-53): taps[0] = 1 (identity)
    and taps[1] = -0.007 at initialisation."""

    def __init__(self, num_taps=63, device="cuda"):
        super().__init__()
        t = torch.zeros(num_taps, dtype=torch.float64)
        t[0] = 1.0
        if num_taps > 1:
            t[1] = -0.007
        self.taps = torch.nn.Parameter(t.to(device))

    def forward(self, x):
        """x: (B, N) float32 on the GPU -> (B, N) float64 view."""
        x = x.contiguous()
        if x.dtype != torch.float32:
            x = x.to(torch.float32)
        return PreEmphasisFunction.apply(x, self.taps).t()


class DiffSDM(torch.nn.Module):
    """Relaxed sigma-delta modulator (diff_sdm.py:42-100) with trainable NTF
    coefficients taken from the reference's table (src/ntf.c)."""

    def __init__(self, ntf=None, rate=5644800, temperature=10.0, input_scale=1.0,
                 checkpoint=DEFAULT_CHECKPOINT, trainable=True, slab_chunks=0, trunc_span=0,
                 device="cuda"):
        """trunc_span > 0: backpropagate within intervals of that many
        samples only (the state at each interval start held constant;
        trunc_span divides the checkpoint interval and is a multiple of 64):
        bounded gradients at high temperature, where the full adjoint grows
        without bound."""
        super().__init__()
        if ntf is None:
            idx = trellis.ntf_auto(rate)
        elif isinstance(ntf, str):
            idx = next(i for i in range(trellis.lib().dsdhip_ntf_count())
                       if trellis.ntf_row(i)["name"] == ntf and trellis.ntf_row(i)["freq"] <= rate)
        else:
            idx = int(ntf)
        row = trellis.ntf_row(idx)
        P = row["order"]
        a = torch.tensor(np.asarray(row["a"][:P]), dtype=torch.float64, device=device)
        g = torch.tensor(np.asarray(row["g"][:P]), dtype=torch.float64, device=device)
        if trainable:
            self.a = torch.nn.Parameter(a)
            self.g = torch.nn.Parameter(g)
        else:
            self.register_buffer("a", a)
            self.register_buffer("g", g)
        self.ntf_name = row["name"]
        self.temperature = float(temperature)
        self.input_scale = float(input_scale)
        self.checkpoint = int(checkpoint)
        self.slab_chunks = int(slab_chunks)
        self.trunc_span = int(trunc_span)

    @property
    def order(self):
        return self.a.numel()

    def forward(self, x):
        """x: (B, N) float64 on the GPU -> soft DSD (B, N) float64 view."""
        if self.input_scale != 1.0:
            x = x * self.input_scale
        y = LoopFilterFunction.apply(_time_major(x), self.a, self.g, self.temperature,
                                     self.checkpoint, self.slab_chunks, self.trunc_span)
        return y.t()


def inband_weights(N, fs, f_hi=20000.0, device="cuda"):
    k = torch.arange(N // 2 + 1, dtype=torch.float64, device=device)
    w = torch.where(k * fs / N <= f_hi, 2.0, 0.0).to(torch.float64)
    w[0] = 1.0
    if N % 2 == 0 and float(w[-1]) != 0.0:
        w[-1] = 1.0
    return w


class AudioBandLoss(torch.nn.Module):
    """In-band error power between the relaxed SDM output and the reference
    signal: L = sum_k w_k |rFFT(y - r)_k|^2 / (B N^2) over 0..cutoff_hz,
    one-sided weights w (2 inside the band, 1 at DC/Nyquist)."""

    def __init__(self, dsd_rate=5644800, cutoff_hz=20000.0):
        super().__init__()
        self.fs = float(dsd_rate)
        self.cutoff = float(cutoff_hz)
        self._w = {}

    def forward(self, y, ref):
        B, N = y.shape
        key = (N, y.device)
        if key not in self._w:
            self._w[key] = inband_weights(N, self.fs, self.cutoff, y.device)
        E = torch.fft.rfft(y - ref.to(torch.float64), dim=1)
        return (self._w[key] * (E.real ** 2 + E.imag ** 2)).sum() / (N * N * B)


def train_step(pre, sdm, loss_fn, x, opt=None):
    """One forward + backward (+ optimizer step) over a batch x (B, N) float32."""
    u = pre(x)
    y = sdm(u)
    loss = loss_fn(y, x)
    if opt is not None:
        opt.zero_grad(set_to_none=True)
    loss.backward()
    if opt is not None:
        opt.step()
    return loss
