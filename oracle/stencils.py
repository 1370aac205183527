"""4th-order central-difference stencils (oracle; test infrastructure only).

Restates ``core/derivatives.py:17-173`` of the reference on whole NumPy grids.
Field arrays are shaped ``(nf, X, Y, Z)``: the C-order reshape of the
reference's flat layout ``idx = z + Z*(y + Y*x) + N*a`` (``core/utils.py:48``).

Boundary semantics reproduced exactly:

* first and pure second derivatives along an axis are computed only where
  ``halo <= i < n - halo`` and are 0 elsewhere (``core/derivatives.py:54,106``);
* mixed derivatives use ``i > halo`` on the low side but ``i < n - halo`` on
  the high side for both axes (``core/derivatives.py:121,139,157``), so index
  ``halo`` itself is zeroed on the low side only.

Arithmetic follows the reference's left-to-right expression order (no fused
multiply-add), so that the oracle equals a strict IEEE evaluation of the
reference's scalar code.
"""

from __future__ import annotations

import numpy as np

C12 = 1.0 / 12.0
C23 = 2.0 / 3.0
C43 = 4.0 / 3.0
C52 = 5.0 / 2.0
# mixed-derivative tensor stencil taps, core/derivatives.py:124-125
MIXED_OFFSETS = (-2, -1, 0, 1, 2)
MIXED_COEFFS = (1.0, -8.0, 0.0, 8.0, -1.0)


def _sl(axis: int, lo: int, hi: int, shift: int, ndim: int = 3):
    """Slice tuple selecting indices ``lo+shift .. hi+shift-1`` along ``axis``."""
    s = [slice(None)] * ndim
    s[axis] = slice(lo + shift, hi + shift)
    return tuple(s)


def d1(f: np.ndarray, axis: int, ls: float, halo: int = 2) -> np.ndarray:
    """First derivative of a scalar grid ``f`` (X, Y, Z) along ``axis``.

    ``core/derivatives.py:54-67``:
    ``((1/12) f[-2] - (2/3) f[-1] + (2/3) f[+1] - (1/12) f[+2]) / ls``
    where ``halo <= i < n-halo``, else 0.
    """
    if halo < 2:
        raise ValueError("halo must be >= 2 for a radius-2 stencil")
    n = f.shape[axis]
    out = np.zeros_like(f)
    lo, hi = halo, n - halo
    if hi <= lo:
        return out
    fm2 = f[_sl(axis, lo, hi, -2)]
    fm1 = f[_sl(axis, lo, hi, -1)]
    fp1 = f[_sl(axis, lo, hi, +1)]
    fp2 = f[_sl(axis, lo, hi, +2)]
    out[_sl(axis, lo, hi, 0)] = (C12 * fm2 - C23 * fm1 + C23 * fp1 - C12 * fp2) / ls
    return out


def d2_pure(f: np.ndarray, axis: int, ls: float, halo: int = 2) -> np.ndarray:
    """Pure second derivative along ``axis`` (``core/derivatives.py:106-119``).

    ``(-(1/12) f[-2] + (4/3) f[-1] - (5/2) f[0] + (4/3) f[+1] - (1/12) f[+2]) / (ls*ls)``
    where ``halo <= i < n-halo``, else 0.
    """
    if halo < 2:
        raise ValueError("halo must be >= 2 for a radius-2 stencil")
    n = f.shape[axis]
    out = np.zeros_like(f)
    lo, hi = halo, n - halo
    if hi <= lo:
        return out
    fm2 = f[_sl(axis, lo, hi, -2)]
    fm1 = f[_sl(axis, lo, hi, -1)]
    f0 = f[_sl(axis, lo, hi, 0)]
    fp1 = f[_sl(axis, lo, hi, +1)]
    fp2 = f[_sl(axis, lo, hi, +2)]
    out[_sl(axis, lo, hi, 0)] = ((-C12) * fm2 + C43 * fm1 - C52 * f0 + C43 * fp1 - C12 * fp2) / (ls * ls)
    return out


def d2_mixed(f: np.ndarray, ax1: int, ax2: int, ls1: float, ls2: float, halo: int = 2) -> np.ndarray:
    """Mixed second derivative (``core/derivatives.py:121-173``).

    ``(1/(144*ls1*ls2)) * sum_{p,q} c_p c_q f(i1+p, i2+q)`` with
    ``c = (1, -8, 0, 8, -1)``, accumulated in the reference's tap order
    (outer loop over the first axis), computed only where
    ``halo < i < n-halo`` on both axes (asymmetric band), else 0.
    Zero-weight taps are skipped: adding an exact zero leaves the partial sum
    bit-identical.
    """
    if halo < 2:
        raise ValueError("halo must be >= 2 for a radius-2 stencil")
    n1, n2 = f.shape[ax1], f.shape[ax2]
    out = np.zeros_like(f)
    lo1, hi1 = halo + 1, n1 - halo
    lo2, hi2 = halo + 1, n2 - halo
    if hi1 <= lo1 or hi2 <= lo2:
        return out
    acc = None
    for oi, ci in zip(MIXED_OFFSETS, MIXED_COEFFS):
        if ci == 0.0:
            continue
        for oj, cj in zip(MIXED_OFFSETS, MIXED_COEFFS):
            if cj == 0.0:
                continue
            s = [slice(None)] * 3
            s[ax1] = slice(lo1 + oi, hi1 + oi)
            s[ax2] = slice(lo2 + oj, hi2 + oj)
            term = (ci * cj) * f[tuple(s)]
            acc = term if acc is None else acc + term
    coeff = 1.0 / (144.0 * ls1 * ls2)
    s = [slice(None)] * 3
    s[ax1] = slice(lo1, hi1)
    s[ax2] = slice(lo2, hi2)
    out[tuple(s)] = coeff * acc
    return out


def first_derivatives(F: np.ndarray, ls, halo: int = 2, comps=None) -> np.ndarray:
    """D1 of every component: array ``d1[i, a, x, y, z]`` (i = axis).

    Equivalent to the reference's ``d1fd1x`` buffer with
    ``idx_d1(i, a, ...)`` (``core/utils.py:82``).
    """
    nf = F.shape[0]
    comps = range(nf) if comps is None else comps
    out = np.zeros((3,) + F.shape, dtype=np.float64)
    for a in comps:
        for i in range(3):
            out[i, a] = d1(F[a], i, ls[i], halo)
    return out


def second_derivatives(F: np.ndarray, ls, halo: int = 2, comps=None) -> np.ndarray:
    """D2 of every component: symmetric array ``d2[i, j, a, x, y, z]``.

    Pure on the diagonal, mixed off the diagonal with the symmetric copy the
    reference makes (``core/derivatives.py:137,155,173``).
    """
    nf = F.shape[0]
    comps = range(nf) if comps is None else comps
    out = np.zeros((3, 3) + F.shape, dtype=np.float64)
    for a in comps:
        for i in range(3):
            out[i, i, a] = d2_pure(F[a], i, ls[i], halo)
        # reference evaluates (1,0) as x-y, (2,0) as x-z, (2,1) as y-z
        mxy = d2_mixed(F[a], 0, 1, ls[0], ls[1], halo)
        mxz = d2_mixed(F[a], 0, 2, ls[0], ls[2], halo)
        myz = d2_mixed(F[a], 1, 2, ls[1], ls[2], halo)
        out[0, 1, a] = out[1, 0, a] = mxy
        out[0, 2, a] = out[2, 0, a] = mxz
        out[1, 2, a] = out[2, 1, a] = myz
    return out
