"""The C ABI library loads and exports every function include/dsd_hip.h
declares (no GPU calls)."""
import ctypes as C
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "dsd_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b((?:gpu_hip|dsdhip)_[a-z0-9_]+)\s*\(", txt)
    return sorted(set(names))


def test_library_exports_header():
    so = os.path.join(ROOT, "foo_dsd_trellis_amd", "libdsdhip.so")
    assert os.path.exists(so), "libdsdhip.so not built"
    lib = C.CDLL(so)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert len(declared()) > 40


def test_ntf_table_matches_oracle(oracle):
    from foo_dsd_trellis_amd import trellis as T
    assert T.lib().dsdhip_ntf_count() == oracle.ntf_count()
    for i in range(oracle.ntf_count()):
        a, b = T.ntf_row(i), oracle.ntf_get(i)
        assert (a["name"], a["order"], a["freq"]) == (b["name"], b["order"], b["freq"])
        assert (a["a"] == b["a"]).all() and (a["g"] == b["g"]).all()
    for rate in (2822400, 5644800, 11289600, 22579200, 24576000):
        assert T.ntf_auto(rate) == oracle.ntf_auto(rate)


def test_fir_plan_and_taps_match_oracle(oracle):
    import numpy as np
    import oracle_lib as ol
    from foo_dsd_trellis_amd import fir
    assert fir.plan(44100, 2822400) == [63] * 6
    assert fir.plan(48000, 24576000) == [127] + [63] * 8
    for T in (63, 127):
        assert np.array_equal(fir.halfband_taps(T), ol.halfband(oracle, T))
