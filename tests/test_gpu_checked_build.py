"""The GPU tests again with the bounds-checked library (libckt_checked.so:
every index the hot-path kernels derive is checked against its array before
the access, CKT_BOUNDS_CHECKS).  The run must pass as with the release
library, and after each of its tests the device's count of out-of-range
indices must be 0 (tests/conftest.py checks it)."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECKED = os.path.join(ROOT, "legionorigins_amd", "libckt_checked.so")
INSIDE = os.environ.get("CKT_CHECKED_RUN") == "1"


def test_release_build_has_no_bounds_checks(C):
    import ctypes
    n = ctypes.c_uint64(7)
    assert C.lib().ckt_debug_bounds_violations(0, ctypes.byref(n)) == 8  # CKT_ERR_UNSUPPORTED
    assert n.value == 0


@pytest.mark.skipif(INSIDE, reason="this is the checked run itself")
def test_gpu_tests_with_bounds_checks():
    assert os.path.exists(CHECKED), f"{CHECKED} is missing: run __graft_entry__.build()"
    env = dict(os.environ, CKT_CHECKED_RUN="1", CKT_LIBRARY=CHECKED)
    files = ["test_gpu_parity.py", "test_gpu_tail.py", "test_gpu_upv.py", "test_gpu_devgen.py",
             "test_gpu_batch.py", "test_gpu_serve.py", "test_gpu_graph.py"]
    r = subprocess.run([sys.executable, "-u", "-m", "pytest", *[os.path.join("tests", f) for f in files],
                        "-m", "gpu", "-x", "-q", "-p", "no:cacheprovider",
                        "-k", "not 512_circuits and not poisson and not growth_regimes"],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert " passed" in r.stdout
