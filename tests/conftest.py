"""Shared test setup: repo root on sys.path and the `gpu` marker."""
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
# Every projector built by the tests replays the operators' index ranges on the host first.
os.environ.setdefault("PRC_AUDIT", "1")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle
    return oracle


@pytest.fixture(scope="session")
def prc():
    import prc_mlem_amd
    prc_mlem_amd.lib()  # fails loudly when the HIP library is missing
    return prc_mlem_amd


def rel_l2(a, b):
    import numpy as np
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def announce(msg):
    """Name the case on stdout before a GPU step (a fault may end the process)."""
    print(msg, flush=True)
