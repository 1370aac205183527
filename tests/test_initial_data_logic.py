"""Host-side initial data: the smorgasbord lumps of the reference's ``_initialize_smorgasbord``
(theories/nuclear_skyrme/initial_config.py:1267-1407) and the topology of their product."""

from __future__ import annotations

import numpy as np

from helpers import make_params
from oracle import integrator as OI
from skyrmion_solver_amd import initial_data


def test_smorgasbord_draws_follow_the_reference_order():
    _, rp, _, _ = make_params("pion", 20, 20, 20)
    lumps, _ = initial_data.lump_specs(rp, dict(mode="smorgasbord", count=4, lump_baryon=2, seed=9,
                                                 position_scale=0.2, profile_m0=0.7))
    rng = np.random.default_rng(9)
    d = 0.2 * np.array([rp.xsize, rp.ysize, rp.zsize])
    pos = rng.uniform(low=-d, high=d, size=(4, 3))
    pos -= pos.mean(axis=0, keepdims=True)
    alpha = rng.uniform(0.0, 2.0 * np.pi, size=(4, 3))
    beta = rng.uniform(0.0, 2.0 * np.pi, size=(4, 3))
    assert len(lumps) == 4
    for n, sp in enumerate(lumps):
        assert sp["mode"] == "rational_map" and sp["B"] == 2.0 and sp["profile"] == ("fit", 0.7)
        np.testing.assert_array_equal(sp["center"], pos[n])
        # rotate_xyz(v, a0, a1, a2) = Rz(a2) Ry(a1) Rz(a0) v with the reference's sign conventions
        v = np.array([0.3, -0.5, 0.8])
        np.testing.assert_allclose(sp["space_rotation"] @ v, initial_data.rotate_xyz(*v, *alpha[n]), atol=1e-15)
        np.testing.assert_allclose(sp["rotation"] @ v, initial_data.rotate_xyz(*v, *beta[n]), atol=1e-15)
    np.testing.assert_allclose(np.mean([sp["center"] for sp in lumps], axis=0), 0.0, atol=1e-15)


def test_smorgasbord_count_defaults_to_the_baryon_number():
    _, rp, _, _ = make_params("pion", 12, 12, 12)
    lumps, _ = initial_data.lump_specs(rp, dict(mode="smorgasbord", baryon_number=3.4))
    assert len(lumps) == 3


def test_smorgasbord_product_carries_the_summed_baryon_number():
    # the degree of a pointwise quaternion product is the sum of the degrees
    _, rp, p_i, p_f = make_params("pion", 40, 40, 40, lsx=0.2, lsy=0.2, lsz=0.2)
    F = initial_data.build_field(rp, 4, dict(mode="smorgasbord", count=3, lump_baryon=1, seed=5,
                                             position_scale=0.15, vacuum_band=True))
    B = OI.baryon_number(OI.OracleState("pion", p_i, p_f, F))
    assert abs(B - 3.0) < 0.05
