"""The HIP trellis against the reference's own vectors (tests/golden)."""
import json
import os

import numpy as np
import pytest

import oracle_lib as ol
from foo_dsd_trellis_amd import trellis as T

pytestmark = pytest.mark.gpu
FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "trellis_fixtures.npz")


def _load():
    z = np.load(FIX, allow_pickle=False)
    return z, json.loads(bytes(z["meta"]).decode())


def test_gpu_reproduces_reference_vectors(gpu):
    z, meta = _load()
    for c in meta["cases"]:
        p = T.SdmPool(c["filter"], c["B"], c["K"], c["L"], c["state_limit"], 1)
        y = p.process(0, z["in_" + c["name"]].astype(np.float64))
        assert y.size == c["nout"], c["name"]
        assert np.array_equal(ol.pack_bits(y), z["bits_" + c["name"]]), c["name"]
        assert np.array_equal(p.best_state(0), np.array(c["best_state"])), c["name"]


def test_gpu_copy_and_seed_vectors(gpu):
    z, meta = _load()
    m = meta["copy"]
    x = z["in_copy"].astype(np.float64)
    s0, s1 = m["split"]
    p = T.SdmPool(m["filter"], m["B"], m["K"], m["L"], 0.0, 2)
    p.process(0, x[:s0])
    p.copy(1, 0)
    p.seed(1, z["seed_copy"])
    assert np.array_equal(ol.pack_bits(p.process(1, x[s0:s1])), z["bits_copy_b"])
    assert np.array_equal(ol.pack_bits(p.process(0, x[s0:s1])), z["bits_copy_a"])
