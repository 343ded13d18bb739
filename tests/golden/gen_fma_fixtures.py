"""Writes tests/golden/trellis_fma_fixtures.npz from the reference's own
src/trellis.c compiled with its shipped AVX2/FMA code path selected
(oracle/_ref/libdsdref_fma.so, built by oracle/Makefile where the reference
sources are present).  Order-8 filters only (the path exists only for
order 8); candidate counts that are and are not multiples of 4, the
state limiter, and copy + seeding.

    python tests/golden/gen_fma_fixtures.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)
import oracle_lib as ol  # noqa: E402
from gen_fixtures import make_input  # noqa: E402

NTF_CLANS_8, NTF_SDM_8 = 8, 9
DSD256, DSD512 = 11289600, 22579200

CASES = [
    ("k32_pink", NTF_CLANS_8, DSD256, 32, 64, 8, 0.0, "pink", 6000),
    ("k32_loud", NTF_CLANS_8, DSD256, 32, 64, 8, 0.0, "loud", 5000),
    ("k30_pink", NTF_CLANS_8, DSD256, 30, 64, 8, 0.0, "pink", 4000),
    ("k13_sine", NTF_CLANS_8, DSD256, 13, 32, 8, 0.0, "sine", 4000),
    ("k3_sine", NTF_CLANS_8, DSD256, 3, 16, 8, 0.0, "sine", 4000),
    ("k8_sdm8", NTF_SDM_8, DSD512, 8, 64, 8, 0.0, "pink", 4000),
    ("k32_limit", NTF_CLANS_8, DSD256, 32, 64, 8, 2.5, "loud", 4000),
    ("k32_b16_l128", NTF_CLANS_8, DSD512, 32, 128, 16, 0.0, "pink", 3000),
]


def main():
    R = ol.reference_fma()
    if R is None:
        sys.exit("oracle/_ref/libdsdref_fma.so is not built (needs the reference sources)")
    out, meta = {}, {"cases": [], "copy": None}
    for idx, (name, nid, rate, K, L, B, lim, kind, n) in enumerate(CASES):
        fi = R.ntf_lookup(nid, rate)
        x = make_input(kind, n, rate, 2000 + idx)
        c = R.ctx(fi, B, K, L, lim)
        y = c.process(x.astype(np.float64))
        out["in_" + name] = x
        out["bits_" + name] = ol.pack_bits(y)
        meta["cases"].append({"name": name, "filter": fi, "K": K, "L": L, "B": B, "state_limit": lim,
                              "n": n, "nout": int(y.size), "counters": c.counters(),
                              "best_state": c.best_state().tolist()})
    fi = R.ntf_lookup(NTF_CLANS_8, DSD256)
    x = make_input("pink", 9000, DSD256, 2077)
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
    out["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    path = os.path.join(HERE, "trellis_fma_fixtures.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes,", len(meta["cases"]), "cases")


if __name__ == "__main__":
    main()
