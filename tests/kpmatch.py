"""Keypoint-set comparison helpers used by the parity tests."""
import numpy as np


def agreement(a, b, tol=1.0):
    """One-to-one matches (same scale bin via SIZE, Euclidean distance <= tol) divided by
    max(|a|, |b|).  OCTAVE is not unique across scales, so SIZE identifies the bin."""
    if len(a) == 0 and len(b) == 0:
        return 1.0
    matched = 0
    for size in np.union1d(a["size"], b["size"]):
        pa = a[a["size"] == size]
        pb = b[b["size"] == size]
        if len(pa) == 0 or len(pb) == 0:
            continue
        d = np.hypot(pa["x"][:, None] - pb["x"][None, :], pa["y"][:, None] - pb["y"][None, :])
        used = np.zeros(len(pb), bool)
        order = np.argsort(d.min(axis=1))
        for i in order:
            row = np.where(used, np.inf, d[i])
            j = int(np.argmin(row))
            if row[j] <= tol:
                used[j] = True
                matched += 1
    return matched / max(len(a), len(b))


def rel_err(a, b):
    """max|a-b| / max|b|."""
    den = float(np.abs(b).max())
    return float(np.abs(a.astype(np.float64) - b.astype(np.float64)).max()) / (den if den else 1.0)


def bits(kps):
    return np.ascontiguousarray(kps).view(np.uint32).reshape(len(kps), -1)
