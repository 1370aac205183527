"""Host observables of the Coulomb-Skyrme model (energy and baryon number reduced on the device in FP64)."""

from __future__ import annotations


def compute_energy(sim) -> float:
    """Total energy: sum over sites of e * vol (reference ``observables.compute_energy``)."""
    return sim.engine.energy()


def compute_skyrmion_number(sim) -> float:
    """Baryon number (reference ``observables.compute_skyrmion_number``)."""
    return sim.engine.baryon()
