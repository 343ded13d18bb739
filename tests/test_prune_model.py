"""The kernel's wave-parallel prune formulation (tests/csrc/prune_model.c,
64-bit lane masks exactly as the kernel computes them) reproduces the
sequential reference prune on every step it decides itself."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as ol

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def model(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("pm") / "libprune.so")
    o = os.path.join(ROOT, "oracle")
    srcs = [os.path.join(HERE, "csrc", "prune_model.c")] + [
        os.path.join(o, f) for f in ("or_trellis.c", "or_fir.c", "or_das.c", "or_api.c")]
    subprocess.check_call(["gcc", "-std=gnu11", "-O2", "-fPIC", "-ffp-contract=off", "-shared",
                           "-o", so] + srcs + ["-lm"])
    lib = ol.TrellisLib(so, "ora")
    lib.lib.prune_model_run.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_size_t,
                                        C.POINTER(C.c_ulonglong)]
    return lib


@pytest.mark.parametrize("K,L,B,kind", [(32, 64, 8, "pink"), (32, 64, 8, "loud"), (8, 16, 8, "sine"),
                                       (16, 32, 8, "pink"), (32, 128, 8, "loud"), (4, 8, 8, "sine")])
def test_bulk_prune_equals_sequential(model, K, L, B, kind):
    rng = np.random.default_rng(K + L)
    n = 40000
    t = np.arange(n)
    if kind == "sine":
        x = 0.25 * np.sin(2 * np.pi * t * 997 / 11289600 * 16) + 1e-5 * rng.standard_normal(n)
    elif kind == "loud":
        x = 0.45 * np.sin(2 * np.pi * t * 997 / 11289600 * 16) + 0.05 * rng.standard_normal(n)
    else:
        w = np.fft.irfft(np.fft.rfft(rng.standard_normal(n)) / np.sqrt(np.arange(n // 2 + 1) + 1.0), n)
        x = np.clip(0.18 * w / np.std(w), -0.5, 0.5)
    c = model.ctx(model.ntf_auto(11289600), B, K, L)
    st = (C.c_ulonglong * 5)()
    x = np.ascontiguousarray(x)
    model.lib.prune_model_run(c.h, x.ctypes.data_as(C.POINTER(C.c_double)), n, st)
    steps, slow, bad = st[0], st[1], st[2]
    assert steps == n
    assert bad == 0
    assert slow < n // 50          # the literal fallback stays rare


@pytest.mark.parametrize("K,L,B,kind", [(32, 64, 8, "pink"), (32, 64, 8, "loud"), (8, 16, 8, "sine"),
                                       (16, 32, 8, "pink"), (32, 128, 8, "loud"), (4, 8, 8, "sine"),
                                       (32, 64, 8, "quiet"), (24, 64, 8, "pink")])
def test_pair_closed_form_equals_sequential(model, K, L, B, kind):
    """The closed form for groups of at most two same-path candidates (best
    member of each group, first K in (metric, index) order; registrations
    from the early skip with pairs leaving the list at their better
    partner) equals the sequential prune on every step it does not decline,
    and declines rarely."""
    rng = np.random.default_rng(7 * K + L)
    n = 60000
    t = np.arange(n)
    if kind == "sine":
        x = 0.25 * np.sin(2 * np.pi * t * 997 / 11289600 * 16) + 1e-5 * rng.standard_normal(n)
    elif kind == "loud":
        x = 0.45 * np.sin(2 * np.pi * t * 997 / 11289600 * 16) + 0.05 * rng.standard_normal(n)
    elif kind == "quiet":
        x = 1e-4 * rng.standard_normal(n)
    else:
        w = np.fft.irfft(np.fft.rfft(rng.standard_normal(n)) / np.sqrt(np.arange(n // 2 + 1) + 1.0), n)
        x = np.clip(0.18 * w / np.std(w), -0.5, 0.5)
    model.lib.pair_model_run.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_size_t,
                                         C.POINTER(C.c_ulonglong)]
    c = model.ctx(model.ntf_auto(11289600), B, K, L)
    st = (C.c_ulonglong * 3)()
    x = np.ascontiguousarray(x)
    model.lib.pair_model_run(c.h, x.ctypes.data_as(C.POINTER(C.c_double)), n, st)
    steps, declined, bad = st[0], st[1], st[2]
    print("steps %d declined %d mismatches %d" % (steps, declined, bad))
    assert steps == n
    assert bad == 0
    assert declined < n // 5
