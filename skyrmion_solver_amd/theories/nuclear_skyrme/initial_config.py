"""Initial data of the nuclear Skyrme model (host NumPy, uploaded once).

Modes follow the reference's ``initialize`` (``theories/nuclear_skyrme/initial_config.py:1410-1564``):
``ground`` / ``uniform`` / ``vacuum``, ``initial`` (the parameters' ansatz),
``hedgehog``, ``rational_map``, plus ``product`` (see
:func:`skyrmion_solver_amd.initial_data.build_field`).
"""

from __future__ import annotations

from skyrmion_solver_amd import initial_data


def initialize(*, sim, config=None):
    return initial_data.build_field(sim.rp, int(sim.rp.number_total_fields), config)
