"""Shared test setup: package path, markers, GPU availability."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "dmipy-sim_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

D = 2e-9

# The autotuner persists its choices; a test session keeps them in a file of its own
# so that tuning runs in tests never steer later walks of the same user.
if "DMIPY_AUTOTUNE_CACHE" not in os.environ:
    import tempfile
    os.environ["DMIPY_AUTOTUNE_CACHE"] = os.path.join(
        tempfile.mkdtemp(prefix="dmipy_autotune_"), "autotune.json")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long CPU run")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as orc
    orc.lib()
    return orc


@pytest.fixture(scope="session")
def gpu():
    import dmipy_sim_amd as ds
    n = ds._native.device_count()
    if n < 1:
        pytest.fail("GPU test collected but no HIP device is visible")
    return ds
