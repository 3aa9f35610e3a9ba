"""Shared helpers for the parity tests."""

import numpy as np


def relerr(a, b):
    """The reference's binding-test norm, max |a - b| / (1 + |b|)
    (code/cpp/bindings/bindings_tests/test_functions.py:20-29)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.size == 0:
        return 0.0
    na, nb = np.isnan(a), np.isnan(b)
    if np.any(na != nb):
        return float("inf")  # a NaN on one side only
    ok = ~na
    if not np.any(ok):
        return 0.0
    return float(np.max(np.abs(a[ok] - b[ok]) / (1.0 + np.abs(b[ok]))))
