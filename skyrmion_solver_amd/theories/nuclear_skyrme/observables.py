"""Host observables of the Nuclear Skyrme model: energy and baryon number, plus the post-convergence
centre of mass, RMS radius, U / V / W inertia tensors, quadrupole tensor and D-term
(``theories/nuclear_skyrme/observables.py:182-486``), all reduced on the device in FP64."""

from __future__ import annotations

from skyrmion_solver_amd.theories._observables import (  # noqa: F401
    compute_center_of_mass, compute_D_term, compute_quadrupole_tensor, compute_rms_radius, compute_U_tensor,
    compute_V_tensor, compute_W_tensor)


def compute_energy(sim) -> float:
    """Total energy: sum over sites of e * vol (reference ``observables.compute_energy``)."""
    return sim.engine.energy()


def compute_skyrmion_number(sim) -> float:
    """Baryon number (reference ``observables.compute_skyrmion_number``)."""
    return sim.engine.baryon()
