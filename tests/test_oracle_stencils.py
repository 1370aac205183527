"""Pin the oracle's stencils (core/derivatives.py:17-173) by closed forms.

* D1 is exact on polynomials of degree <= 4, pure D2 on degree <= 5, mixed D2
  on products p(x) q(y) with degree <= 4 each (away from the zeroed band);
* the error is O(h^4) on smooth non-polynomial fields;
* the zeroed band is exactly `i < halo or i >= n - halo` for D1 / pure D2 and
  `i <= halo or i >= n - halo` for the mixed derivative (asymmetric guard);
* a planted mistake (a wrong coefficient) fails the closed-form check.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import stencils as OS


def grid(n=(14, 13, 12), ls=(0.3, 0.25, 0.2)):
    x = ls[0] * np.arange(n[0])
    y = ls[1] * np.arange(n[1])
    z = ls[2] * np.arange(n[2])
    return np.meshgrid(x, y, z, indexing="ij"), ls


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("deg", [1, 2, 3, 4])
def test_d1_exact_on_polynomials(axis, deg):
    (X, Y, Z), ls = grid()
    c = [X, Y, Z][axis]
    others = [X, Y, Z]
    others.pop(axis)
    f = 0.3 + 1.7 * c ** deg - 0.4 * c ** max(deg - 1, 0) + 0.05 * others[0] * others[1]
    exact = deg * 1.7 * c ** (deg - 1) - 0.4 * max(deg - 1, 0) * c ** max(deg - 2, 0)
    d = OS.d1(f, axis, ls[axis], 2)
    n = f.shape[axis]
    sl = [slice(None)] * 3
    sl[axis] = slice(2, n - 2)
    np.testing.assert_allclose(d[tuple(sl)], exact[tuple(sl)], rtol=1e-11, atol=1e-10)


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("deg", [2, 3, 4, 5])
def test_d2_pure_exact_on_polynomials(axis, deg):
    (X, Y, Z), ls = grid()
    c = [X, Y, Z][axis]
    f = 1.3 * c ** deg + 0.7 * c + 0.1 * ([Y * Z, X * Z, X * Y][axis])
    exact = 1.3 * deg * (deg - 1) * c ** (deg - 2)
    d = OS.d2_pure(f, axis, ls[axis], 2)
    n = f.shape[axis]
    sl = [slice(None)] * 3
    sl[axis] = slice(2, n - 2)
    np.testing.assert_allclose(d[tuple(sl)], exact[tuple(sl)], rtol=1e-9, atol=1e-8)


@pytest.mark.parametrize("axes", [(0, 1), (0, 2), (1, 2)])
def test_mixed_exact_on_products(axes):
    (X, Y, Z), ls = grid()
    C = [X, Y, Z]
    a, b = axes
    p = C[a] ** 4 - 2 * C[a] ** 3 + C[a]
    q = 0.5 * C[b] ** 4 + C[b] ** 2 - C[b]
    f = p * q + 3.0 * C[a] ** 2 + C[b] ** 3  # separable extras have zero mixed derivative
    dp = 4 * C[a] ** 3 - 6 * C[a] ** 2 + 1
    dq = 2 * C[b] ** 3 + 2 * C[b] - 1
    d = OS.d2_mixed(f, a, b, ls[a], ls[b], 2)
    sl = [slice(None)] * 3
    sl[a] = slice(3, f.shape[a] - 2)
    sl[b] = slice(3, f.shape[b] - 2)
    np.testing.assert_allclose(d[tuple(sl)], (dp * dq)[tuple(sl)], rtol=1e-9, atol=1e-8)


def test_guard_pattern_d1_d2_and_asymmetric_mixed():
    rng = np.random.default_rng(0)
    f = rng.normal(size=(11, 10, 9)) + 5.0
    ls = (0.2, 0.2, 0.2)
    for axis in range(3):
        n = f.shape[axis]
        nz = np.abs(OS.d1(f, axis, ls[axis], 2)).max(axis=tuple(i for i in range(3) if i != axis)) > 0
        assert list(np.nonzero(nz)[0]) == list(range(2, n - 2))
        nz = np.abs(OS.d2_pure(f, axis, ls[axis], 2)).max(axis=tuple(i for i in range(3) if i != axis)) > 0
        assert list(np.nonzero(nz)[0]) == list(range(2, n - 2))
    m = OS.d2_mixed(f, 0, 1, 0.2, 0.2, 2)
    nzx = np.nonzero(np.abs(m).max(axis=(1, 2)) > 0)[0]
    nzy = np.nonzero(np.abs(m).max(axis=(0, 2)) > 0)[0]
    # core/derivatives.py:121 -- x > halo (low) but x < n - halo (high)
    assert list(nzx) == list(range(3, 11 - 2))
    assert list(nzy) == list(range(3, 10 - 2))
    # a larger halo moves both edges
    m3 = OS.d2_mixed(f, 0, 2, 0.2, 0.2, 3)
    assert list(np.nonzero(np.abs(m3).max(axis=(1, 2)) > 0)[0]) == list(range(4, 11 - 3))


def test_fourth_order_convergence():
    errs = []
    for n in (24, 48):
        h = 2.0 / (n - 1)
        x = h * np.arange(n)
        X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
        f = np.sin(1.3 * X) * np.cos(0.7 * Y) + np.sin(Z)
        d = OS.d2_mixed(f, 0, 1, h, h, 2)
        ex = -1.3 * 0.7 * np.cos(1.3 * X) * np.sin(0.7 * Y)
        sl = (slice(3, n - 2), slice(3, n - 2), slice(2, n - 2))
        errs.append(np.abs(d[sl] - ex[sl]).max())
    rate = np.log2(errs[0] / errs[1])
    assert 3.5 < rate < 4.6, rate


def test_planted_coefficient_mistake_is_caught(monkeypatch):
    """Replacing 2/3 by 0.66 must break the degree-<=4 exactness test."""
    (X, Y, Z), ls = grid()
    f = X ** 3
    monkeypatch.setattr(OS, "C23", 0.66)
    d = OS.d1(f, 0, ls[0], 2)
    err = np.abs(d[2:-2] - 3 * X[2:-2] ** 2).max()
    assert err > 1e-3


def test_full_derivative_buffers_are_symmetric():
    rng = np.random.default_rng(1)
    F = rng.normal(size=(4, 9, 8, 10))
    d2 = OS.second_derivatives(F, (0.2, 0.3, 0.25))
    for i in range(3):
        for j in range(3):
            assert np.array_equal(d2[i, j], d2[j, i])
