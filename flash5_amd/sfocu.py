"""sfocu-style comparison norms (tools/sfocu/sfocu.c:31-33, 812-884).

For each variable over interior leaf cells, unweighted:
  * L1 (sfocu's default ``-L1``, benchmark B): sum|B - A| / max(sum|B|, 1e-99)
  * Linf ("magErr"): sup|A - B| / max(max(|A|, |B|) over cells, 1e-99)
  * per-cell norm 2|a - b| / |a + b| (sfocu.c:31-33), reported as its max
"""
from __future__ import annotations

import numpy as np


def interior(U: np.ndarray, ndim: int, ng: int) -> np.ndarray:
    """View of the interior cells of (..., nz, ny, nx) block data."""
    sl = [slice(None)] * U.ndim
    for a in range(ndim):
        sl[U.ndim - 1 - a] = slice(ng, U.shape[U.ndim - 1 - a] - ng)
    return U[tuple(sl)]


def l1_rel(a: np.ndarray, b: np.ndarray) -> float:
    return float(np.abs(b - a).sum() / max(np.abs(b).sum(), 1e-99))


def linf_rel(a: np.ndarray, b: np.ndarray) -> float:
    scale = max(float(np.abs(a).max(initial=0.0)), float(np.abs(b).max(initial=0.0)), 1e-99)
    return float(np.abs(a - b).max(initial=0.0) / scale)


def cell_norm_max(a: np.ndarray, b: np.ndarray) -> float:
    den = np.abs(a + b)
    num = 2.0 * np.abs(a - b)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.where(den > 0, num / np.where(den > 0, den, 1.0), np.where(num > 0, np.inf, 0.0))
    return float(r.max(initial=0.0))


def compare(A: np.ndarray, B: np.ndarray, var_axis: int, names, ndim: int, ng: int):
    """Per-variable (L1, Linf) of block data A against benchmark B."""
    out = {}
    for n, v in names.items():
        a = interior(np.take(A, v, axis=var_axis), ndim, ng)
        b = interior(np.take(B, v, axis=var_axis), ndim, ng)
        out[n] = (l1_rel(a, b), linf_rel(a, b))
    return out
