"""Host observables of the Lightly bound Skyrme model: energy and baryon number, plus the centre of
mass and RMS radius (``theories/lightly_bound_skyrme/observables.py:176-258``), all reduced on the device in FP64."""

from __future__ import annotations

from skyrmion_solver_amd.theories._observables import compute_center_of_mass, compute_rms_radius  # noqa: F401


def compute_energy(sim) -> float:
    """Total energy: sum over sites of e * vol (reference ``observables.compute_energy``)."""
    return sim.engine.energy()


def compute_skyrmion_number(sim) -> float:
    """Baryon number (reference ``observables.compute_skyrmion_number``)."""
    return sim.engine.baryon()
