"""Comparison helpers for the parity tests.

Parity follows the reference's own divergence (its parameters drive the
relaxation to inf/NaN within ~10 inner steps): per element the non-finite
class must match exactly (NaN, +inf, -inf); where both are finite the values
are compared bit for bit (EXACT policy with the CSR reduction) or to a
relative tolerance (1e-5, the stated bar for node voltages).
"""
import numpy as np

TOL = 1e-5


def classes(a):
    a = np.asarray(a)
    return np.where(np.isnan(a), 3, np.where(np.isposinf(a), 1, np.where(np.isneginf(a), 2, 0)))


def assert_same_class(got, want, what=""):
    cg, cw = classes(got), classes(want)
    bad = cg != cw
    assert not bad.any(), f"{what}: {bad.sum()} elements differ in finite/inf/NaN class"


def assert_bit_identical(got, want, what=""):
    """Classes equal and finite values bit-identical (NaN payloads ignored)."""
    got, want = np.asarray(got), np.asarray(want)
    assert_same_class(got, want, what)
    fin = np.isfinite(want)
    g = got[fin].view(np.uint32) if got.dtype == np.float32 else got[fin]
    w = want[fin].view(np.uint32) if want.dtype == np.float32 else want[fin]
    diff = g != w
    if diff.any():
        i = np.flatnonzero(diff)[0]
        raise AssertionError(f"{what}: {diff.sum()} finite values differ, e.g. "
                             f"{got[fin][i]!r} vs {want[fin][i]!r}")


def max_rel_err(got, want, floor=0.0):
    """Max |got-want| / max(|want|, floor) over elements finite in both."""
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    fin = np.isfinite(got) & np.isfinite(want)
    if not fin.any():
        return 0.0
    den = np.maximum(np.abs(want[fin]), floor)
    den = np.where(den == 0, 1.0, den)
    return float(np.max(np.abs(got[fin] - want[fin]) / den))


def max_wire_rel_err(got, want):
    """Per wire, the error of its segment vector relative to the vector's
    largest magnitude; max over wires (arrays of shape [W, S])."""
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    ok = np.isfinite(got).all(1) & np.isfinite(want).all(1)
    if not ok.any():
        return 0.0
    scale = np.abs(want[ok]).max(1)
    scale = np.where(scale == 0, 1.0, scale)
    return float((np.abs(got[ok] - want[ok]).max(1) / scale).max())
