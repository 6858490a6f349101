"""The host driver ckt_sim (region_main replacement): command line of
circuit.cc:991-1040 and report lines of circuit.cc:144-146,165,314-334."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SIM = os.path.join(ROOT, "legionorigins_amd", "ckt_sim")


def run(*args, timeout=120):
    return subprocess.run([SIM, *map(str, args)], capture_output=True, text=True, timeout=timeout)


@pytest.mark.skipif(os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK),
                    reason="a GPU is visible")
def test_driver_without_gpu_reports_error():
    r = run("-l", "1")
    assert r.returncode == 1
    assert "no usable device" in r.stderr
    assert "circuit settings: loops=1 pieces=4" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("extra", [[], ["-sync", "1"], ["-policy", "fast", "-graph"],
                                   ["-mode", "csr", "-per_piece", "-profile", "stable"],
                                   ["-devgen", "-policy", "fast"],
                                   ["-devgen", "-policy", "scaled", "-graph"]])
def test_driver_report_format(extra):
    r = run("-l", "2", "-p", "4", "-npp", "1024", "-wpp", "4096", "-pct", "95", "-s", "12345",
            *extra)
    assert r.returncode == 0, r.stderr
    out = r.stdout
    assert "circuit settings: loops=2 pieces=4 nodes/piece=1024 wires/piece=4096 " \
           "pct_in_piece=95 seed=12345" in out
    assert "Starting main simulation loop" in out and "SUCCESS!" in out
    t = float(re.search(r"ELAPSED TIME = +([0-9.]+) s", out).group(1))
    g = float(re.search(r"GFLOPS = +([0-9.]+) GFLOPS", out).group(1))
    assert g > 0
    # GFLOPS follows the reference formula (circuit.cc:320-334); the printed
    # time is rounded to 1 ms
    ops = (16384 * (10 * 6 + 9 * 4) * 10000 + 16384 * 4 + 4096 * 4) * 2
    assert 1e-9 * ops / (t + 0.0005) <= g * 1.001
    if t > 0.0005:
        assert g <= 1.001 * 1e-9 * ops / (t - 0.0005)
    assert re.search(r"KERNEL calc_new_currents +[0-9.]+ ms total", out)
