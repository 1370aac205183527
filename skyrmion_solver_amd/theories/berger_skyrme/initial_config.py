"""Initial data of the Berger Skyrme model.

The reference's ``theories/berger_skyrme/initial_config.py`` is the nuclear
model's initial-data module unchanged, so the nuclear builder is used as is.
"""

from __future__ import annotations

from skyrmion_solver_amd.theories.nuclear_skyrme.initial_config import initialize  # noqa: F401
