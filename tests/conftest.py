import os
import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(REPO / "gas-atenuator_amd"))
sys.path.insert(0, str(REPO / "oracle"))
sys.path.insert(0, str(REPO))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long-running statistical test")


@pytest.fixture(scope="session")
def raw():
    from gasx.tables import raw_tables
    return raw_tables()


@pytest.fixture(scope="session")
def otab(raw):
    import oracle
    return oracle.build_tables(raw)


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda:0")


def log(msg: str):
    """Name the case before GPU work (flushed, so a fault still shows it)."""
    print(f"[case] {msg}", flush=True)
