"""Post-convergence observables on the engine: the device sums the per-site
integrands in FP64 (``skx_observable``) and these functions apply the
reference's host formulas (``theories/nuclear_skyrme/observables.py:182-486``).
"""

from __future__ import annotations

import math

import numpy as np

from skyrmion_solver_amd import _engine


def compute_center_of_mass(sim) -> np.ndarray:
    """Baryon-density centre of mass (``compute_center_of_mass``, ``observables.py:182-222``).

    ``(sum x b, sum y b, sum z b) / sum b`` on the grid ``x = lsx * i``;
    zeros when ``|sum b| < 1e-30``.
    """
    m = sim.engine.observable(_engine.OBS_COM)
    if abs(m[0]) < 1.0e-30:
        return np.zeros(3, dtype=np.float64)
    return m[1:4] / m[0]


def compute_rms_radius(sim) -> float:
    """``sqrt(sum |r|^2 |b| / B)`` about the centre of mass (``compute_rms_radius``, ``:224-271``)."""
    com = compute_center_of_mass(sim)
    B = sim.engine.baryon()
    if abs(B) < 1.0e-30:
        return 0.0
    num = sim.engine.observable(_engine.OBS_RMS, com)[0]
    return math.sqrt(max(num / B, 0.0))


def _tensor(sim, kind, com=None) -> np.ndarray:
    return sim.engine.observable(kind, com).reshape(3, 3)


def compute_U_tensor(sim) -> np.ndarray:
    """Isospin inertia tensor ``U_ij = sum 2 U0_ij vol`` (``compute_U_tensor``, ``:273-306``)."""
    return _tensor(sim, _engine.OBS_U)


def compute_V_tensor(sim) -> np.ndarray:
    """Spin inertia tensor about the centre of mass (``compute_V_tensor``, ``:308-350``)."""
    return _tensor(sim, _engine.OBS_V, compute_center_of_mass(sim))


def compute_W_tensor(sim) -> np.ndarray:
    """Mixed inertia tensor about the centre of mass (``compute_W_tensor``, ``:352-394``)."""
    return _tensor(sim, _engine.OBS_W, compute_center_of_mass(sim))


def compute_quadrupole_tensor(sim) -> np.ndarray:
    """Isospin-0 electric quadrupole tensor (``compute_quadrupole_tensor``, ``:396-438``)."""
    return _tensor(sim, _engine.OBS_QUADRUPOLE, compute_center_of_mass(sim))


def compute_D_term(sim) -> float:
    """Monopole D-term ``-(2/5) E sum (r_i r_j T_ij - r^2 T_kk / 3) vol`` (``compute_D_term``, ``:440-486``)."""
    com = compute_center_of_mass(sim)
    E = sim.engine.energy()
    integral = sim.engine.observable(_engine.OBS_D, com)[0]
    return -(2.0 / 5.0) * E * integral
