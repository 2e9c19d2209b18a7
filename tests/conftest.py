import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (runs the HIP library through the C-ABI)")
    config.addinivalue_line(
        "markers", "slow: CPU test that takes more than ~20 s")


def pytest_report_header(config):
    return f"cubff repo root: {ROOT}"


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.build()
    return O
