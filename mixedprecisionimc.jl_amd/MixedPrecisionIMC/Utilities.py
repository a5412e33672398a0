"""Utilities -- src/imc_utilities.jl.

``sorter`` is the reference's product reorderer (imc_utilities.jl:23-54), vectorised over
cells: each row of ``vals`` is sorted together with the scale and the smallest*largest pairs
are multiplied.  The host runs it in FP64 (Fleck factor, initial material energy); the GPU
sourcing kernel runs the same rule in the tracking weight precision.
"""
from __future__ import annotations

import numpy as np


def tointeger(x: float) -> int:
    """imc_utilities.jl:6-21.  Counts are 64-bit here (the Float16 -> UInt16 overflow
    of the reference for values above 65504 does not occur)."""
    xi = int(round(float(x)))
    if float(xi) != float(x):
        raise ValueError(f"InexactError: tointeger({x})")
    return xi


def sorter(vals, scales, precision=np.float64):
    """Product of ``vals`` (last axis) in sorted smallest*largest pairs, for the first scale
    (descending list) that keeps it finite.  Returns (product, scale, 1-based index); index 0
    (product 0) where no scale works, as the reference's fallback."""
    v = np.asarray(vals, dtype=np.float64)
    single = v.ndim == 1
    if single:
        v = v[None, :]
    scales = np.atleast_1d(np.asarray(scales, dtype=np.float64))
    nrow = v.shape[0]
    prod_out = np.zeros(nrow)
    scale_out = np.zeros(nrow)
    idx_out = np.zeros(nrow, dtype=np.int64)
    todo = np.ones(nrow, dtype=bool)
    for j, s in enumerate(scales):
        if not todo.any():
            break
        b = np.sort(np.concatenate([v[todo], np.full((todo.sum(), 1), s)], axis=1), axis=1)
        m = b.shape[1]
        p = np.ones(b.shape[0], dtype=precision)
        with np.errstate(over="ignore", invalid="ignore"):
            for i in range(m // 2):
                p = p * (b[:, i] * b[:, m - 1 - i]).astype(precision)
            if m % 2:
                p = p * b[:, m // 2].astype(precision)
        ok = np.isfinite(p)
        rows = np.flatnonzero(todo)
        prod_out[rows[ok]] = p[ok]
        scale_out[rows[ok]] = s
        idx_out[rows[ok]] = j + 1
        todo[rows[ok]] = False
    if single:
        return float(prod_out[0]), float(scale_out[0]), int(idx_out[0])
    return prod_out, scale_out, idx_out
