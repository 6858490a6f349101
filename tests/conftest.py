import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (runs through the C ABI)")


@pytest.fixture(scope="session")
def O():
    from oracle import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def C():
    from legionorigins_amd import circuit
    circuit.lib()
    return circuit


def pytest_report_header(config):
    return [f"threads for the oracle: {min(16, os.cpu_count() or 1)}"]
