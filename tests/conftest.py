import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libkiwi_hip on the device)")


@pytest.fixture(scope="session")
def hip_ctx():
    # PyTorch-ROCm bundles its own libamdhip64.so.7; the process shares whichever copy loads first, and
    # torch's device initialisation needs its own.  Tests that stage device tensors with torch therefore
    # bring torch's GPU up before the library's context.
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.init()
    except ImportError:
        pass
    from kiwi_amd import hip
    ctx = hip.Context(0)
    yield ctx
    ctx.close()
