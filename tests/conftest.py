import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (run with -m gpu)")
    config.addinivalue_line("markers", "slow: long-running")
    # the C oracle is test infrastructure; build it if it is missing
    if not os.path.exists(os.path.join(ROOT, "oracle", "liborc.so")):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True,
                       capture_output=True)


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP library, initialised on cuda:0. Fails loudly when it is missing."""
    import torch
    assert torch.cuda.is_available(), "GPU test without a GPU"
    from rrtmg_sw_amd import api
    api.rrtmg_sw_ini(1004.64)
    return api
