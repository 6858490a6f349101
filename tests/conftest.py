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


@pytest.fixture(autouse=True)
def _no_bounds_violations(request):
    """In a run with the bounds-checked library, every GPU test must leave the
    device's count of out-of-range kernel indices at 0."""
    yield
    if os.environ.get("CKT_CHECKED_RUN") != "1" or request.node.get_closest_marker("gpu") is None:
        return
    import ctypes
    from legionorigins_amd import circuit
    n = ctypes.c_uint64()
    st = circuit.lib().ckt_debug_bounds_violations(0, ctypes.byref(n))
    assert st == 0, circuit.lib().ckt_last_error()
    assert n.value == 0, f"{n.value} out-of-range kernel indices"
