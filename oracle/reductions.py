"""Grid reductions (oracle; test infrastructure only).

Restates ``core/utils.py:201-279,321-373,402-425``: a 1024-wide shared-memory
tree per block (``sdata[tid] += sdata[tid + stride]`` for stride = 512, 256,
..., 1), one partial per block, then a host ``ndarray.sum()`` / ``.max()`` of
the partials.
"""

from __future__ import annotations

import numpy as np

TPB = 1024  # core/utils.py:343


def _block_tree(values: np.ndarray, fill: float, op) -> np.ndarray:
    v = np.asarray(values, dtype=np.float64).ravel()
    nblocks = (v.size + TPB - 1) // TPB
    buf = np.full(nblocks * TPB, fill, dtype=np.float64)
    buf[: v.size] = v
    buf = buf.reshape(nblocks, TPB)
    stride = TPB // 2
    while stride > 0:
        buf[:, :stride] = op(buf[:, :stride], buf[:, stride: 2 * stride])
        stride //= 2
    return buf[:, 0].copy()


def ref_sum(values: np.ndarray) -> float:
    """``compute_sum`` (``core/utils.py:321-346``): block tree, then host sum."""
    partial = _block_tree(values, 0.0, np.add)
    return float(partial.sum())


def ref_max(values: np.ndarray) -> float:
    """``compute_max`` (``core/utils.py:348-373``): signed maximum.

    Empty slots hold ``-1.0e300`` (``core/utils.py:267``); the block tree keeps
    the larger of two entries, which is an exact maximum in any order.
    """
    partial = _block_tree(values, -1.0e300, np.maximum)
    return float(partial.max())
