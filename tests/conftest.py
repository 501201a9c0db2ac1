import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))
sys.path.insert(0, str(ROOT / "oracle"))  # the oracle's NumPy equation layer (oracle/equ.py)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    # Under pytest-xdist every worker loads the OpenMP oracle; split the cores
    # between them, since oversubscribed OpenMP barriers spin for a long time.
    workers = int(os.environ.get("PYTEST_XDIST_WORKER_COUNT", "0"))
    if workers > 1 and "OMP_NUM_THREADS" not in os.environ:
        os.environ["OMP_NUM_THREADS"] = str(max(1, (os.cpu_count() or 1) // workers))


def pytest_collection_modifyitems(config, items):
    # GPU tests need the HIP runtime and a device; everything else must pass on CPU.
    pass
