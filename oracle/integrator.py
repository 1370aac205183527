"""RK4 / arrested-Newton-flow step (oracle; test infrastructure only).

Restates ``core/integrator.py:50-92`` (RK4 stage), ``:143-177`` (constrained
RK4 finalise), ``:284-410`` (``do_arrested_newton_flow``) and the
``Simulation`` buffer set (``core/simulation.py:94-118``) on NumPy grids.

The gradient of a stage is a plug-in (``gradient_fn(Temp) -> (EG, l)``), so
the same integrator can be driven by a linear test problem whose RK4
amplification is known in closed form.
"""

from __future__ import annotations

import numpy as np

from . import reductions, stencils, theories


class OracleState:
    """Host mirror of the reference's device buffers for one lattice."""

    def __init__(self, theory: str, p_i, p_f, F: np.ndarray, V: np.ndarray | None = None):
        self.theory = theory
        self.p_i = np.asarray(p_i, dtype=np.int32)
        self.p_f = np.asarray(p_f, dtype=np.float64)
        X, Y, Z = int(self.p_i[0]), int(self.p_i[1]), int(self.p_i[2])
        nf = int(self.p_i[5])
        self.shape = (nf, X, Y, Z)
        self.F = np.array(F, dtype=np.float64).reshape(self.shape)
        self.V = np.zeros(self.shape) if V is None else np.array(V, dtype=np.float64).reshape(self.shape)
        self.EG = np.zeros(self.shape)

    @property
    def halo(self) -> int:
        return int(self.p_i[3])

    @property
    def ls(self):
        return (self.p_f[3], self.p_f[4], self.p_f[5])

    @property
    def nsk(self) -> int:
        return theories.n_sigma_fields(self.theory, self.p_i)


def derivatives(state: OracleState, T: np.ndarray):
    """All first and second derivatives of ``T`` (``core/integrator.py:42-45``)."""
    d1 = stencils.first_derivatives(T, state.ls, state.halo)
    d2 = stencils.second_derivatives(T, state.ls, state.halo)
    return d1, d2


def gradient(state: OracleState, T: np.ndarray):
    """One ``gradient_step_kernel`` launch: returns ``(EG, l)`` for field ``T``."""
    d1, d2 = derivatives(state, T)
    return theories.GRADIENTS[state.theory](T, d1, d2, state.p_i, state.p_f)


def energy_density(state: OracleState, F: np.ndarray | None = None) -> np.ndarray:
    """``compute_energy_kernel``: per-site energy times vol."""
    F = state.F if F is None else F
    d1 = stencils.first_derivatives(F, state.ls, state.halo)
    return theories.ENERGIES[state.theory](F, d1, state.p_i, state.p_f)


def energy(state: OracleState, F: np.ndarray | None = None) -> float:
    """``observables.compute_energy`` (``theories/nuclear_skyrme/observables.py:114-147``)."""
    return reductions.ref_sum(energy_density(state, F))


def baryon_density(state: OracleState, F: np.ndarray | None = None) -> np.ndarray:
    """``compute_skyrmion_number_kernel``: per-site baryon density times vol."""
    F = state.F if F is None else F
    d1 = stencils.first_derivatives(F, state.ls, state.halo, comps=range(4))
    return theories.baryon_density(state.theory, F, d1, state.p_i, state.p_f)


def baryon_number(state: OracleState, F: np.ndarray | None = None) -> float:
    """``observables.compute_skyrmion_number`` (``observables.py:149-180``)."""
    return reductions.ref_sum(baryon_density(state, F))


def rk4_step(F, V, dt, gradient_fn, finalise_fn=None):
    """The four RK4 stages and the finalise of ``do_arrested_newton_flow``.

    ``core/integrator.py:382-398``; ``finalise_fn(F, V)`` applies the
    theory's normalisation and projection (``:172-175``).  Returns the
    stage-4 ``EnergyGradient`` (the array the reference's ``err`` reads).
    """
    k1 = dt * (V + 0.0 * V)
    T = F + 0.0 * F
    EG, l1 = gradient_fn(T)
    k2 = dt * (V + 0.5 * l1)
    T = F + 0.5 * k1
    EG, l2 = gradient_fn(T)
    k3 = dt * (V + 0.5 * l2)
    T = F + 0.5 * k2
    EG, l3 = gradient_fn(T)
    k4 = dt * (V + 1.0 * l3)
    T = F + 1.0 * k3
    EG, l4 = gradient_fn(T)
    V += (1.0 / 6.0) * (l1 + 2.0 * l2 + 2.0 * l3 + l4)
    F += (1.0 / 6.0) * (k1 + 2.0 * k2 + 2.0 * k3 + k4)
    if finalise_fn is not None:
        finalise_fn(F, V)
    return EG


def anf_step(state: OracleState, prev_energy: float):
    """``Simulation.step`` -> ``do_arrested_newton_flow`` (``core/integrator.py:284-410``).

    Returns ``(new_energy, err)`` where ``err`` is the reference's signed
    maximum of the stage-4 energy gradient over all ``nf*N`` entries.
    Mutates ``state.F`` / ``state.V`` / ``state.EG``.
    """
    dt = state.p_f[7]
    unit_length = int(state.p_i[10])
    killkinen = int(state.p_i[8])
    nsk = state.nsk

    def finalise(F, V):
        if unit_length:
            theories.normalise(F, nsk)
            theories.project(V, F, nsk)

    EG = rk4_step(state.F, state.V, dt, lambda T: gradient(state, T), finalise)
    state.EG = EG
    new_energy = energy(state)
    if killkinen and (new_energy > prev_energy):
        state.V[...] = 0.0
    err = reductions.ref_max(EG)
    return new_energy, err


def max_abs_force(state: OracleState) -> float:
    """True ``max |g|`` of the last stage-4 gradient (convergence measure)."""
    return float(np.max(np.abs(state.EG)))
