"""Initial data of the Coulomb-Skyrme model (V starts at 0, as in the reference) (host NumPy, uploaded once).

Modes follow the reference's ``initialize`` (``theories/coulomb_skyrme/initial_config.py``):
``ground`` / ``uniform`` / ``vacuum``, ``initial`` (the parameters' ansatz),
``hedgehog``, ``rational_map``, plus ``product`` (see
:func:`skyrmion_solver_amd.initial_data.build_field`).
"""

from __future__ import annotations

from skyrmion_solver_amd import initial_data


def initialize(*, sim, config=None):
    return initial_data.build_field(sim.rp, int(sim.rp.number_total_fields), config)
