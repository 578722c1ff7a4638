import os
import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(REPO / "ecl-suite_amd"))
sys.path.insert(0, str(REPO))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libeclcc.so on cuda:0)")


@pytest.fixture(scope="session")
def cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("gpu test selected but no GPU is visible")
    return torch.device("cuda:0")


def announce(msg: str) -> None:
    """Names the case before each GPU step (flushed), so a hard failure
    still shows what was running."""
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()
