"""The two correctly rounded divisions of the EXACT step, checked on the CPU.

The FP32 split division (legionorigins_amd/csrc/ckt_divsplit.h), the one the
step loop uses: tests/native/div_split.cpp, built for the host from the same
header the kernels use, checks every dividend significand against IEEE
division for the step's divisors, edge divisors and random ones (including
divisors that need the moved split), random operand pairs over the whole
exponent range, the special dividends the kernel's guard has to catch, and
that at least 99.8% of divisors get a verified split.

The FP64 reciprocal, which takes what the split does not cover and the short
reference-order tails of the other policies, checked with IEEE float32/float64
arithmetic (numpy): for float a and b, with rb = RN64(1/b),

    RN32(RN64(a * rb)) == RN32(a / b)   whenever a / b is a normal float.

A quotient of two floats is never exactly halfway between two normal floats
and stays at least 2^-49 (relative) from every such midpoint, while the FP64
product is within 2^-52 of a / b (S. Figueroa, "When is double rounding
innocuous?", 1995).  The kernel sends zero and subnormal quotients to the IEEE
division; this test checks the normal range, with random operands, with the
divisors the circuit uses (dt, R in [1, 11], C in [0.001, 0.011]) and with
dividends placed within two ulps of b*m for random midpoints m.
"""
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

FLT_MIN = np.float32(2.0 ** -126)


def method(a, b):
    rb = 1.0 / b.astype(np.float64)
    return (a.astype(np.float64) * rb).astype(np.float32)


def check(a, b):
    with np.errstate(all="ignore"):
        want = a / b
        got = method(a, b)
    normal = np.isfinite(want) & (np.abs(want) >= FLT_MIN)
    assert normal.mean() > 0.5
    bad = got[normal].view(np.uint32) != want[normal].view(np.uint32)
    assert not bad.any(), (a[normal][bad][:3], b[normal][bad][:3])


def test_random_operands():
    rng = np.random.default_rng(0)
    for _ in range(10):
        a = rng.integers(0, 2**32, 2_000_000, dtype=np.uint64).astype(np.uint32).view(np.float32)
        b = rng.integers(0, 2**32, 2_000_000, dtype=np.uint64).astype(np.uint32).view(np.float32)
        check(a, b)


def test_circuit_divisors():
    rng = np.random.default_rng(1)
    n = 3_000_000
    a = (rng.standard_normal(n) * 10.0 ** rng.uniform(-30, 30, n)).astype(np.float32)
    for b in (np.full(n, 1e-6, np.float32),
              rng.uniform(1, 11, n).astype(np.float32),
              rng.uniform(0.001, 0.011, n).astype(np.float32)):
        check(a, b)


def test_dividends_next_to_midpoints():
    rng = np.random.default_rng(2)
    n = 2_000_000
    b = np.ldexp(1.0 + rng.integers(0, 2**23, n) / 2.0**23, rng.integers(-30, 30, n))
    b = b.astype(np.float32)
    # midpoint m = (2^24 + k + 1/2) * 2^e: 25 significant bits
    m = np.ldexp((2.0**24 + rng.integers(0, 2**24, n) + 0.5), rng.integers(-54, 6, n))
    a0 = (m * b.astype(np.float64)).astype(np.float32)
    for d in (-2, -1, 0, 1, 2):
        a = a0.copy()
        for _ in range(abs(d)):
            a = np.nextafter(a, np.float32(np.inf) if d > 0 else np.float32(-np.inf))
        check(a, b)


@pytest.fixture(scope="module")
def div_split_exe(tmp_path_factory):
    from legionorigins_amd import build as B
    exe = str(tmp_path_factory.mktemp("div_split") / "div_split")
    cmd = [B.hipcc(), "-O2", "-std=c++17", "-ffp-contract=off", "--offload-arch=gfx950", "-x", "hip",
           "-I" + os.path.join(ROOT, "legionorigins_amd", "csrc"),
           os.path.join(ROOT, "tests", "native", "div_split.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


def _run(exe, *args):
    env = dict(os.environ)
    env.pop("LD_PRELOAD", None)
    return subprocess.run([exe, *args], capture_output=True, text=True, env=env, timeout=600)


def test_split_division_exhaustive(div_split_exe):
    r = _run(div_split_exe)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "split division: ok" in r.stdout


def test_planted_divisors_are_classified_as_recorded(div_split_exe):
    """The GPU test that plants these divisors in circuits relies on their
    kinds: refused ones take the FP64 path, moved ones the moved split."""
    fix = json.load(open(os.path.join(ROOT, "tests", "golden", "division_divisors.json")))
    want = {"out_of_range": "refused", "power_of_two": "base", "zero": "refused"}
    args = [(k, v) for k in ("base", "moved", "refused", "power_of_two", "out_of_range", "zero")
            for v in fix[k]]
    negs = [v[1:] for v in fix["negative"]]
    r = _run(div_split_exe, *[v for _, v in args], *fix["negative"], *negs)
    assert r.returncode == 0, r.stderr[-2000:]
    got = dict(line.split() for line in r.stdout.split("\n") if line.strip())
    for k, v in args:
        assert got[v] == want.get(k, k), (k, v, got[v])
    for v, m in zip(fix["negative"], negs):
        assert got[v] == got[m], (v, got[v], got[m])
    assert {got[v] for v in fix["negative"]} == {"base", "moved", "refused"}
