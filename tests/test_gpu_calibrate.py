"""bullseye_fp64_atomic_rate: the device's FP64 atomic-add rate (the grid flushes' ceiling).

The entry point checks on the host that every one of its 2^28 adds landed, so a return of 0
means the count was exact; the rate must be positive and physically plausible (between
1e8 and 1e12 adds per second), and the call needs no initLibrary and can be repeated.
"""
import ctypes

import pytest

from bullseye_amd.library_loader import load_library

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["mixed", "double"])
def test_fp64_atomic_rate(precision):
    print("case fp64 atomic rate", precision, flush=True)
    lib = load_library("GPU", precision)
    rates = []
    for _ in range(2):
        r = ctypes.c_double(0.0)
        assert lib.bullseye_fp64_atomic_rate(ctypes.byref(r)) == 0, lib.bullseye_last_error()
        rates.append(r.value)
    print("rates %.3e %.3e" % tuple(rates))
    assert all(1e8 < r < 1e12 for r in rates), rates


def test_fp64_atomic_rate_rejects_null():
    lib = load_library("GPU", "mixed")
    assert lib.bullseye_fp64_atomic_rate(None) != 0
    assert b"NULL" in lib.bullseye_last_error()
