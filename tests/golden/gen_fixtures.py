#!/usr/bin/env python3
"""Writes tests/golden/trellis_fixtures.npz from the reference's own
trellis.c / ntf.c, compiled from /root/reference by oracle/Makefile into
oracle/_ref/libdsdref.so (run `make -C oracle` first).

Every case stores its input (float32, widened exactly to double for the
trellis), its parameters, the reference's +-1 output as packed bits and
the reference's counters; estimator cases store final states.  Only the
reference is called here; the tests check the oracle and the HIP kernels
against these vectors.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as ol  # noqa: E402

DSD64, DSD128, DSD256, DSD512 = 2822400, 5644800, 11289600, 22579200
NTF_CLANS_4, NTF_CLANS_5, NTF_CLANS_6, NTF_CLANS_7, NTF_CLANS_8 = 0, 2, 4, 6, 8


def pinkish(n, rng):
    w = rng.standard_normal(n)
    f = np.fft.rfft(w)
    k = np.arange(f.size, dtype=np.float64)
    k[0] = 1.0
    f /= np.sqrt(k)
    y = np.fft.irfft(f, n)
    return y / np.sqrt(np.mean(y * y))


def make_input(kind, n, rate, seed):
    rng = np.random.default_rng(seed)
    t = np.arange(n) / rate
    if kind == "sine":
        x = 0.5012 * np.sin(2 * np.pi * 997 * 16 * t) + 1e-5 * rng.standard_normal(n)
    elif kind == "pink":
        x = np.clip(0.18 * pinkish(n, rng), -0.5, 0.5)
    elif kind == "loud":
        x = 0.45 * np.sin(2 * np.pi * 997 * 16 * t) + 0.05 * rng.standard_normal(n)
    elif kind == "dc":
        x = np.full(n, 0.3) + 1e-4 * rng.standard_normal(n)
    elif kind == "full":
        x = 0.95 * np.sign(np.sin(2 * np.pi * 50 * 64 * t))
    else:
        raise ValueError(kind)
    return x.astype(np.float32)


# name, ntf id, rate, K, L, B, state_limit, kind, n
CASES = [
    ("c1_dsd64_k8", NTF_CLANS_5, DSD64, 8, 16, 8, 0.0, "sine", 8000),
    ("dsd256_k32_pink", NTF_CLANS_8, DSD256, 32, 64, 8, 0.0, "pink", 6000),
    ("dsd256_k32_loud", NTF_CLANS_8, DSD256, 32, 64, 8, 0.0, "loud", 6000),
    ("dsd128_k16", NTF_CLANS_6, DSD128, 16, 32, 8, 0.0, "pink", 6000),
    ("k32_l128_b16", NTF_CLANS_8, DSD256, 32, 128, 16, 0.0, "pink", 4000),
    ("k4_l8_b4", NTF_CLANS_5, DSD64, 4, 8, 4, 0.0, "sine", 4000),
    ("k1_l1_b1", NTF_CLANS_5, DSD64, 1, 1, 1, 0.0, "sine", 2000),
    ("k32_b32", NTF_CLANS_8, DSD256, 32, 64, 32, 0.0, "pink", 4000),
    ("k32_limit", NTF_CLANS_8, DSD256, 32, 64, 8, 2.5, "loud", 4000),
    ("k2_l16_b4", NTF_CLANS_6, DSD256, 2, 16, 4, 0.0, "pink", 6000),
    ("order4_k8", NTF_CLANS_4, DSD64, 8, 32, 8, 0.0, "sine", 4000),
    ("order7_k16", NTF_CLANS_7, DSD256, 16, 64, 8, 0.0, "pink", 4000),
    ("dc_k8", NTF_CLANS_5, DSD64, 8, 16, 8, 0.0, "dc", 3000),
    ("fullscale_k8", NTF_CLANS_5, DSD64, 8, 16, 8, 4.0, "full", 3000),
    ("dsd512_k32_l200", NTF_CLANS_8, DSD512, 32, 200, 8, 0.0, "loud", 3000),
]


def main():
    R = ol.reference()
    if R is None:
        sys.exit("oracle/_ref/libdsdref.so is not built (needs the reference sources)")
    out = {}
    meta = {"cases": [], "copy": None, "drain": None, "estimate": [], "ntf": []}
    for i in range(R.ntf_count()):
        r = R.ntf_get(i)
        meta["ntf"].append({"name": r["name"], "order": r["order"], "freq": r["freq"]})
        out["ntf_a_%d" % i] = r["a"]
        out["ntf_g_%d" % i] = r["g"]
    lookups = []
    for nid in range(10):
        for rate in (DSD64, DSD128, DSD256, DSD512, 3072000, 6144000, 12288000, 24576000, 90316800):
            lookups.append([nid, rate, R.ntf_lookup(nid, rate)])
    meta["lookup"] = lookups
    meta["auto"] = [[rate, R.ntf_auto(rate)] for rate in (DSD64, DSD128, DSD256, DSD512, 24576000)]

    for idx, (name, nid, rate, K, L, B, lim, kind, n) in enumerate(CASES):
        fi = R.ntf_lookup(nid, rate)
        x = make_input(kind, n, rate, 1000 + idx)
        c = R.ctx(fi, B, K, L, lim)
        y = c.process(x.astype(np.float64))
        out["in_" + name] = x
        out["bits_" + name] = ol.pack_bits(y)
        meta["cases"].append({"name": name, "filter": fi, "K": K, "L": L, "B": B, "state_limit": lim,
                              "n": n, "nout": int(y.size), "counters": c.counters(),
                              "best_state": c.best_state().tolist()})

    # copy_state + DAS-style seeding, then both contexts continue
    fi = R.ntf_lookup(NTF_CLANS_8, DSD256)
    x = make_input("pink", 9000, DSD256, 77)
    a = R.ctx(fi, 8, 32, 64)
    a.process(x[:3000].astype(np.float64))
    b = R.ctx(fi, 8, 32, 64)
    b.copy_from(a)
    seed = a.best_state() * 0.5
    b.seed(seed)
    yb = b.process(x[3000:6000].astype(np.float64))
    ya = a.process(x[3000:6000].astype(np.float64))
    out["in_copy"] = x
    out["seed_copy"] = seed
    out["bits_copy_b"] = ol.pack_bits(yb)
    out["bits_copy_a"] = ol.pack_bits(ya)
    meta["copy"] = {"filter": fi, "K": 32, "L": 64, "B": 8, "split": [3000, 6000],
                    "nout_a": int(ya.size), "nout_b": int(yb.size)}

    # drain after a partial block
    fi5 = R.ntf_lookup(NTF_CLANS_5, DSD64)
    xd = make_input("sine", 1000, DSD64, 78)
    d = R.ctx(fi5, 8, 8, 16)
    y1 = d.process(xd.astype(np.float64))
    y2 = d.drain(64)
    out["in_drain"] = xd
    out["bits_drain"] = ol.pack_bits(np.concatenate([y1, y2]))
    meta["drain"] = {"filter": fi5, "K": 8, "L": 16, "B": 8, "n1": int(y1.size), "n2": int(y2.size)}

    # 16-level and 2-level greedy estimator
    xe32 = make_input("pink", 20000, DSD256, 79)
    xe = xe32.astype(np.float64)
    for levels, lim in ((16, 0.0), (16, 3.0), (2, 0.0), (8, 0.0)):
        init = np.linspace(-0.01, 0.01, 10)
        st = R.estimate(fi, init, xe, lim, levels)
        meta["estimate"].append({"levels": levels, "state_limit": lim, "filter": fi,
                                 "state": st.tolist(), "init": init.tolist()})
    out["in_estimate"] = xe32

    out["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    path = os.path.join(HERE, "trellis_fixtures.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
