"""Shared helpers for the tests (fixture loading, error measures)."""
import glob
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def fixtures(prefix):
    out = []
    for p in sorted(glob.glob(os.path.join(GOLDEN, prefix + "_*.npz"))):
        with np.load(p, allow_pickle=False) as z:
            out.append((os.path.basename(p), {k: z[k] for k in z.files}))
    return out


def rel(a, b):
    """Relative L2 error ||a - b|| / ||b||."""
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def maxrel(a, b):
    """max |a - b| / max |b|."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)
