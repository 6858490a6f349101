"""bench.py's command-line contract, checked without a GPU: more than one GPU
is refused with a message (the library runs on one GPU), and the options the
measurement runs use are accepted."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = run("--gpus", "2", "--steps", "1", "--warmup", "0")
    assert r.returncode == 2
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["n_gpus"] == 2 and "one GPU" in out["error"]


def test_help_lists_the_measurement_modes():
    r = run("--help")
    assert r.returncode == 0
    for flag in ("--steps", "--warmup", "--gpus", "--batched", "--served", "--upv-sweep",
                 "--dsc-sweep", "--gen", "--accessor", "--profile", "--copy-calibration"):
        assert flag in r.stdout, flag
