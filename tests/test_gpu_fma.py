"""The order-8 FMA filter mode (the reference AVX2 build's arithmetic) on
the GPU: the reference's own vectors from its AVX2/FMA code path
(tests/golden/trellis_fma_fixtures.npz) bit for bit, including the
loop-filter states, and DAS against the oracle in the same mode."""
import json
import os

import numpy as np
import pytest

import oracle_lib as ol
from foo_dsd_trellis_amd import das as D
from foo_dsd_trellis_amd import signals as SG
from foo_dsd_trellis_amd import trellis as T

pytestmark = pytest.mark.gpu
FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "trellis_fma_fixtures.npz")


def _load():
    z = np.load(FIX, allow_pickle=False)
    return z, json.loads(bytes(z["meta"]).decode())


def test_gpu_fma_reproduces_reference_vectors(gpu):
    z, meta = _load()
    for c in meta["cases"]:
        p = T.SdmPool(c["filter"], c["B"], c["K"], c["L"], c["state_limit"], 1)
        p.fma_o8(True)
        y = p.process(0, z["in_" + c["name"]].astype(np.float64))
        assert y.size == c["nout"], c["name"]
        assert np.array_equal(ol.pack_bits(y), z["bits_" + c["name"]]), c["name"]
        assert np.array_equal(p.best_state(0), np.array(c["best_state"])), c["name"]


def test_gpu_fma_copy_and_seed_vectors(gpu):
    z, meta = _load()
    m = meta["copy"]
    x = z["in_copy"].astype(np.float64)
    s0, s1 = m["split"]
    p = T.SdmPool(m["filter"], m["B"], m["K"], m["L"], 0.0, 2)
    p.fma_o8(True)
    p.process(0, x[:s0])
    p.copy(1, 0)
    p.seed(1, z["seed_copy"])
    assert np.array_equal(ol.pack_bits(p.process(1, x[s0:s1])), z["bits_copy_b"])
    assert np.array_equal(ol.pack_bits(p.process(0, x[s0:s1])), z["bits_copy_a"])


@pytest.mark.parametrize("seg0", [0, 1])
def test_gpu_fma_das_matches_oracle(gpu, oracle, seg0):
    rate, K, L, B, F, nch = 11289600, 32, 64, 8, 100000, 2
    rng = np.random.default_rng(21)
    x = np.stack([np.clip(0.18 * SG.pink(F, rng), -0.5, 0.5) for _ in range(nch)])
    fi = oracle.ntf_auto(rate)
    with ol.fma_o8(oracle):
        ref, n, sp, _ = ol.das(oracle, fi, B, K, L, x, seg_len=16384, nseg=0, seg0_first=seg0,
                               reposition=1)
    d = D.Das(fi, B, K, L, nch, seg_len=16384, nseg=0, seg0_first=bool(seg0), reposition=True)
    d.fma_o8(True)
    y, m = d.run(x)
    assert m == n
    for c in range(nch):
        bad = np.nonzero(y[c, :n] != ref[c, :n])[0]
        assert bad.size == 0, "ch%d first mismatch %d" % (c, bad[0])
