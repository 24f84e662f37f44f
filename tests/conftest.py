import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def st():
    import thesis_amd
    thesis_amd.lib()
    return thesis_amd


@pytest.fixture(scope="session")
def gpu(st):
    n = st.device_count()
    if n < 1:
        raise RuntimeError("GPU test collected but no HIP device is visible")
    name = st.device_name(0)
    print(f"[gpu] {name}", flush=True)
    return st
