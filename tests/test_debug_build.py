"""Debug builds, on the CPU.

  * The kernels' index arithmetic (legionorigins_amd/csrc/ckt_index.h: the
    XCD tile order, the piece-table workgroup map, the LDS reduction's wire
    subranges, update_voltages' node groups), compiled for the host with
    AddressSanitizer and UndefinedBehaviorSanitizer and checked exhaustively:
    every element of a launch is covered by exactly one lane.
  * The CPU suite run again with the sanitized library (libckt_san.so: host
    code under AddressSanitizer and UndefinedBehaviorSanitizer, kernels with
    bounds checks) and the sanitized oracle, with the sanitizer runtime
    preloaded; any report fails it.
"""
import os
import subprocess
import sys
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN_LIB = os.path.join(ROOT, "legionorigins_amd", "libckt_san.so")
SAN_ORACLE = os.path.join(ROOT, "oracle", "libcircuit_oracle_san.so")
INSIDE = os.environ.get("CKT_SANITIZED_RUN") == "1"


def test_kernel_index_arithmetic_exhaustive():
    from legionorigins_amd import build as B
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "index_math")
        cmd = [B.hipcc(), "-O1", "-g", "-std=c++17", "--offload-arch=gfx950", "-x", "hip",
               "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
               "-Xarch_host", "-fno-sanitize-recover=undefined",
               "-I" + os.path.join(ROOT, "legionorigins_amd", "csrc"),
               os.path.join(ROOT, "tests", "native", "index_math.cpp"), "-o", exe]
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-3000:]
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
        env.pop("LD_PRELOAD", None)
        r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
        assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
        assert "index arithmetic: ok" in r.stdout


@pytest.mark.skipif(INSIDE, reason="this is the sanitized run itself")
def test_cpu_suite_under_sanitizers():
    from legionorigins_amd import build as B
    for f in (SAN_LIB, SAN_ORACLE):
        assert os.path.exists(f), f"{f} is missing: run __graft_entry__.build()"
    env = dict(os.environ, CKT_SANITIZED_RUN="1", CKT_LIBRARY=SAN_LIB, ORACLE_LIBRARY=SAN_ORACLE,
               LD_PRELOAD=B.sanitizer_runtime(),
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([sys.executable, "-m", "pytest", os.path.join(ROOT, "tests"), "-m", "not gpu",
                        "-q", "-x", "-p", "no:cacheprovider",
                        "--deselect", "tests/test_debug_build.py::test_kernel_index_arithmetic_exhaustive"],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert " passed" in r.stdout


@pytest.mark.skipif(not INSIDE, reason="only inside the sanitized run")
def test_sanitized_libraries_are_loaded(C, O):
    maps = open("/proc/self/maps").read()
    assert "libckt_san.so" in maps and "libcircuit_oracle_san.so" in maps
    assert "libclang_rt.asan" in maps
