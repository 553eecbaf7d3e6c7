import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
# the autotuner's cache file stays out of the tests (test_gpu_autotune.py
# gives its child processes their own)
os.environ["RRG_TUNE_CACHE"] = ""


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs the library's kernels")


@pytest.fixture(scope="session")
def rrg():
    import rrgalaxsim_amd

    rrgalaxsim_amd.lib()
    return rrgalaxsim_amd


@pytest.fixture(scope="session")
def orc():
    import oracle

    oracle.lib()
    return oracle
