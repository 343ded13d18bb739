import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


def gpu_available():
    try:
        from foo_dsd_trellis_amd import lib
        return bool(lib().gpu_hip_probe())
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib
    return oracle_lib.oracle()


@pytest.fixture(scope="session")
def reference():
    import oracle_lib
    r = oracle_lib.reference()
    if r is None:
        pytest.skip("reference sources not built here (oracle/_ref)")
    return r


@pytest.fixture(scope="session")
def reference_fma():
    import oracle_lib
    r = oracle_lib.reference_fma()
    if r is None:
        pytest.skip("reference sources not built here (oracle/_ref)")
    return r


@pytest.fixture(scope="session")
def gpu():
    if not gpu_available():
        pytest.fail("GPU test run without a usable GPU / libdsdhip.so")
    return True
