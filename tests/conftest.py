import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def gpu_ctx():
    """One liar_amd Context on cuda:0 for the whole GPU session."""
    import torch

    import liar_amd

    if not torch.cuda.is_available():
        pytest.fail("GPU test collected but no GPU is visible")
    torch.cuda.init()
    ctx = liar_amd.Context(0)
    yield ctx
    ctx.close()
