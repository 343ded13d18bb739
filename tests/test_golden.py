"""The oracle against the reference's own vectors (tests/golden, written by
gen_fixtures.py from the compiled reference trellis.c / ntf.c)."""
import json
import os

import numpy as np
import pytest

import oracle_lib as ol

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "trellis_fixtures.npz")


@pytest.fixture(scope="module")
def fx():
    z = np.load(FIX, allow_pickle=False)
    meta = json.loads(bytes(z["meta"]).decode())
    return z, meta


def test_ntf_table(oracle, fx):
    z, meta = fx
    assert oracle.ntf_count() == len(meta["ntf"])
    for i, row in enumerate(meta["ntf"]):
        r = oracle.ntf_get(i)
        assert (r["name"], r["order"], r["freq"]) == (row["name"], row["order"], row["freq"])
        assert np.array_equal(r["a"], z["ntf_a_%d" % i]) and np.array_equal(r["g"], z["ntf_g_%d" % i])
    for nid, rate, idx in meta["lookup"]:
        assert oracle.ntf_lookup(nid, rate) == idx
    for rate, idx in meta["auto"]:
        assert oracle.ntf_auto(rate) == idx


def _cases(fx):
    z, meta = fx
    return [(c, z["in_" + c["name"]], z["bits_" + c["name"]]) for c in meta["cases"]]


def test_trellis_cases(oracle, fx):
    for c, x, bits in _cases(fx):
        ctx = oracle.ctx(c["filter"], c["B"], c["K"], c["L"], c["state_limit"])
        y = ctx.process(x.astype(np.float64))
        assert y.size == c["nout"], c["name"]
        assert np.array_equal(ol.pack_bits(y), bits), c["name"]
        assert ctx.counters() == c["counters"], c["name"]
        assert np.array_equal(ctx.best_state(), np.array(c["best_state"])), c["name"]


def test_latency_count_and_pm1(oracle, fx):
    """Fill produces no output; then 1 in -> 1 out of +-1 (test/test_trellis.c:78-118)."""
    for c, x, _ in _cases(fx):
        ctx = oracle.ctx(c["filter"], c["B"], c["K"], c["L"], c["state_limit"])
        y = ctx.process(x.astype(np.float64))
        assert y.size == max(0, x.size - c["L"])
        assert set(np.unique(y)) <= {-1.0, 1.0}


def test_copy_and_seed(oracle, fx):
    z, meta = fx
    m = meta["copy"]
    x = z["in_copy"].astype(np.float64)
    a = oracle.ctx(m["filter"], m["B"], m["K"], m["L"])
    a.process(x[:m["split"][0]])
    b = oracle.ctx(m["filter"], m["B"], m["K"], m["L"])
    b.copy_from(a)
    b.seed(z["seed_copy"])
    yb = b.process(x[m["split"][0]:m["split"][1]])
    ya = a.process(x[m["split"][0]:m["split"][1]])
    assert np.array_equal(ol.pack_bits(yb), z["bits_copy_b"])
    assert np.array_equal(ol.pack_bits(ya), z["bits_copy_a"])


def test_drain(oracle, fx):
    z, meta = fx
    m = meta["drain"]
    d = oracle.ctx(m["filter"], m["B"], m["K"], m["L"])
    y1 = d.process(z["in_drain"].astype(np.float64))
    y2 = d.drain(64)
    assert (y1.size, y2.size) == (m["n1"], m["n2"])
    assert np.array_equal(ol.pack_bits(np.concatenate([y1, y2])), z["bits_drain"])


def test_estimator(oracle, fx):
    z, meta = fx
    xe = z["in_estimate"].astype(np.float64)
    for e in meta["estimate"]:
        st = oracle.estimate(e["filter"], np.array(e["init"]), xe, e["state_limit"], e["levels"])
        assert np.array_equal(st, np.array(e["state"])), e
