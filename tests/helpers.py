"""Shared helpers of the test suite."""
import numpy as np


def rel(a, b):
    """Norm-wise relative difference ||a - b|| / ||b|| (0 when both vanish)."""
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    nb = np.linalg.norm(b)
    na = np.linalg.norm(a - b)
    if nb == 0.0:
        return na
    return na / nb


def entry_rel(a, b):
    """max |a - b| / max |b| (entrywise error scaled by the largest entry)."""
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    s = np.abs(b).max()
    return np.abs(a - b).max() / (s if s > 0 else 1.0)
