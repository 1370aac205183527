"""Initial data of the omega-Skyrme model.

The pion components come from :func:`skyrmion_solver_amd.initial_data.build_field`;
the omega component is then set to ``-c_omega * CS`` from the engine's baryon
density (``CS * vol``), as ``create_omega_field_kernel`` does
(``theories/omega_skyrme/initial_config.py:1495-1532``).
"""

from __future__ import annotations

import numpy as np

from skyrmion_solver_amd import _engine, initial_data


def initialize(*, sim, config=None):
    return initial_data.build_field(sim.rp, int(sim.rp.number_total_fields), config)


def post_initialize(*, sim, config=None):
    if (config or {}).get("omega_init", "chern_simons") != "chern_simons":
        return
    F = sim.engine.download(_engine.FIELD)
    cs_vol = sim.engine.density(_engine.DENSITY_BARYON)
    N = cs_vol.size
    F[4 * N:5 * N] = -float(sim.rp.c_omega) * (cs_vol / float(sim.rp.grid_volume))
    sim.engine.upload(_engine.FIELD, F)
