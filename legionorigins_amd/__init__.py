"""MI355X (gfx950) implementation of the Legion circuit hot path.

The compute lives in ``libckt.so`` (HIP kernels behind the C ABI of
``include/circuit_abi.h``); :mod:`legionorigins_amd.circuit` is the host-side
binding.
"""
from . import circuit  # noqa: F401

__all__ = ["circuit"]
