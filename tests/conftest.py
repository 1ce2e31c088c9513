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


def deposit_error_bound(op, st, vel):
    """Per-cell bound on |J(FP32 fixed point) - J(FP64 EM7)|.

    FP32 positions resolve r/dr and z/dz to an absolute max(nr, nz) * 2^-24,
    so every area weight of a particle is off by at most a few times that
    (absolute, not relative to the weight); v_r and the products add a few
    2^-24 of |v|.  The bound sums FN*e*|v|/V(i) over the particles whose 3x3
    cell neighbourhood contains the cell."""
    import numpy as np
    nr, nz = op.nr, op.nz
    ok = ~np.isnan(st[2])
    r = np.sqrt(st[0][ok].astype(np.float64) ** 2 + st[1][ok].astype(np.float64) ** 2)
    ir = np.clip(np.floor(r / op.dr - 0.5), -1, nr - 1).astype(int) + 1
    iz = np.clip(np.floor(st[2][ok] / op.dz - 0.5), -1, nz - 1).astype(int) + 1
    h = np.zeros((nz + 2, nr + 2))
    np.add.at(h, (iz, ir), np.abs(vel[ok]))
    box = np.zeros_like(h)
    for dk in (-1, 0, 1):
        for di in (-1, 0, 1):
            box += np.roll(np.roll(h, dk, 0), di, 1)
    box = box[1:nz + 1, 1:nr + 1]
    i = np.arange(nr)
    V = op.phip / 2 * ((i + 1) ** 2 - i ** 2) * op.dr ** 2 * op.dz
    tol = (16 + 4 * max(nr, nz)) * 2.0 ** -24
    return (tol * op.fn * 1.6e-19 * box / V).ravel()


def log(msg: str):
    """Name the case before GPU work (flushed, so a fault still shows it)."""
    print(f"[case] {msg}", flush=True)
