import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT / "hpr-qp_amd", ROOT):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

REF = Path("/root/reference")
FIXTURES = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libhprqp.so")
    config.addinivalue_line("markers", "slow: longer CPU oracle runs")


def fixture_path(name):
    """Reference data files are mirrored under tests/golden/ (the GPU box has
    no /root/reference)."""
    p = FIXTURES / name
    if p.exists():
        return str(p)
    q = REF / name
    if q.exists():
        return str(q)
    pytest.skip(f"fixture {name} not available")
