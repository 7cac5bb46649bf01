"""Shared comparison helpers for the parity tests."""
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def relerr(a, b):
    """max |a-b| / max |b| (0 when both are zero)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.max(np.abs(b)) if b.size else 0.0
    num = np.max(np.abs(a - b)) if b.size else 0.0
    if den == 0.0:
        return float(num)
    return float(num / den)


def assert_rel(a, b, tol, what=""):
    e = relerr(a, b)
    assert np.all(np.isfinite(a)) == np.all(np.isfinite(b)), what + " finiteness differs"
    assert e <= tol, "%s: relative error %.3e > %.1e" % (what, e, tol)


def have_gpu():
    try:
        import torch  # noqa: F401  (only to count devices; HIP itself is used through libfvk)
        return torch.cuda.device_count() > 0
    except Exception:
        return False


def require_lib():
    from programs_amd import fvk
    if not os.path.exists(fvk.LIB_PATH):
        pytest.fail("libfvk.so is not built: run python -m programs_amd.build")
    return fvk
