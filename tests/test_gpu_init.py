"""GPU: device-side initial data equals the oracle's restatement of the
reference's initial configurations (oracle/initial_config.py, pinned by
tests/test_oracle_initial_config.py) on the same request.

The request's random draws (centres, rotation matrices) come from
initial_data.lump_specs; the oracle evaluates the lumps, their product and the
vacuum band on its own.  Transcendentals (cos, sin, exp, acos, atan2, tan,
tanh) come from the device math library, so the FP64 device field matches
the oracle to 1e-12 absolute rather than bitwise; the FP32 form to FP32
rounding.
"""

from __future__ import annotations

import numpy as np
import pytest

from helpers import make_params, oracle_field
from oracle import initial_config as OIC
from skyrmion_solver_amd import _engine, initial_data

pytestmark = pytest.mark.gpu

CONFIGS = [
    dict(mode="hedgehog"),
    dict(mode="hedgehog", profile=("exp", 0.9), isorotation="haar", jitter=0.2, seed=4, vacuum_band=True),
    dict(mode="rational_map", baryon_number=4, vacuum_band=True),
    dict(mode="product", seed=2, vacuum_band=True,
         lumps=[dict(mode="rational_map", baryon_number=2, center=(-0.8, 0.0, 0.0), profile="fit"),
                dict(mode="hedgehog", center=(0.9, 0.1, 0.0), profile=("gauss", 0.7), isorotation="haar")]),
] + [dict(mode="rational_map", baryon_number=B, profile=("exp", 1.2), isorotation="haar", space_rotation="haar",
          seed=B, isospin_rotation=0.3) for B in range(1, 10)] + [
    # the reference's fitted profiles (massive, and mixed with the massless fit), every B, and beyond the fits
    dict(mode="rational_map", baryon_number=B, profile=("fit", m0), isospin_rotation=0.2 * B)
    for B in range(1, 11) for m0 in (1.0, 0.35)] + [
    dict(mode="smorgasbord", baryon_number=5, seed=11, vacuum_band=True),
    dict(mode="smorgasbord", count=16, lump_baryon=1, profile_m0=0.6, position_scale=0.25, seed=3),
    dict(mode="smorgasbord", count=3, lump_baryon=2, seed=7),
]


def _device_field(theory, prec, rp, p_i, p_f, cfg, omega_cs=0.0):
    with _engine.Engine({"pion": 0, "omega": 1, "rho": 3}[theory], prec, p_i, p_f) as e:
        e.init_field(initial_data.device_init_desc(rp, cfg, omega_cs=omega_cs))
        return e.download(_engine.FIELD), e.download(_engine.VELOCITY)


@pytest.mark.parametrize("k", range(len(CONFIGS)))
def test_device_init_matches_oracle_fp64(k):
    cfg = CONFIGS[k]
    print(f"[case] device init {cfg}", flush=True)
    mod, rp, p_i, p_f = make_params("pion", 24, 22, 20)
    F_ref = oracle_field(rp, p_i, p_f, 4, cfg).ravel()
    F, V = _device_field("pion", _engine.FP64, rp, p_i, p_f, cfg)
    assert np.max(np.abs(F - F_ref)) <= 1e-12
    assert np.max(np.abs(V)) == 0.0


def test_device_init_fp32_and_extra_components():
    cfg = dict(mode="rational_map", baryon_number=3, profile=("exp", 1.1), isorotation="haar", seed=7,
               vacuum_band=True)
    mod, rp, p_i, p_f = make_params("rho", 20, 18, 16)
    F_ref = oracle_field(rp, p_i, p_f, 13, cfg).ravel()
    F, _ = _device_field("rho", _engine.FP32, rp, p_i, p_f, cfg)
    assert np.max(np.abs(F - F_ref)) <= 2e-7  # FP32 storage rounding
    N = 20 * 18 * 16
    assert np.max(np.abs(F[4 * N:])) == 0.0  # rho components start at 0


def test_device_omega_from_chern_simons():
    cfg = dict(mode="rational_map", baryon_number=2, profile=("exp", 1.0), vacuum_band=True)
    mod, rp, p_i, p_f = make_params("omega", 22, 20, 18)
    c = 0.8
    F_ref = OIC.create_omega_field(oracle_field(rp, p_i, p_f, 5, cfg), p_i, p_f, c)
    F, _ = _device_field("omega", _engine.FP64, rp, p_i, p_f, cfg, omega_cs=c)
    assert np.max(np.abs(F[4 * F_ref[0].size:])) > 0.1 * c  # omega is not trivially zero
    assert np.max(np.abs(F - F_ref.ravel())) <= 1e-11 * max(1.0, np.max(np.abs(F_ref)))


def test_batch_init_equals_engine_init():
    cfg = dict(mode="rational_map", baryon_number=2, profile=("exp", 1.0), isorotation="haar", seed=3,
               vacuum_band=True)
    mod, rp, p_i, p_f = make_params("pion", 18, 16, 20)
    desc = initial_data.device_init_desc(rp, cfg)
    with _engine.Batch(_engine.THEORY_PION, _engine.FP64, max_slots=2, max_blocks=256) as b:
        s = b.add(p_i, p_f, init=desc, max_steps=8)
        b.run(8)
        st, Fb, _ = b.remove(s)
    with _engine.Engine(_engine.THEORY_PION, _engine.FP64, p_i, p_f, tile_z=32, x_chunk=int(p_i[0])) as e:
        e.init_field(desc)
        ref = e.anf_run(8, -1.0, 8)
        Fe = e.download(_engine.FIELD)
    assert st["steps"] == ref.steps == 8
    assert st["energy"] == ref.energy
    assert np.array_equal(Fb, Fe)


def test_device_noise_is_seeded_and_sized():
    mod, rp, p_i, p_f = make_params("pion", 20, 20, 20)
    base = dict(mode="rational_map", baryon_number=1, profile=("exp", 1.0))
    F0, _ = _device_field("pion", _engine.FP64, rp, p_i, p_f, base)
    F1, _ = _device_field("pion", _engine.FP64, rp, p_i, p_f, dict(base, noise=1e-3, seed=5))
    F2, _ = _device_field("pion", _engine.FP64, rp, p_i, p_f, dict(base, noise=1e-3, seed=5))
    F3, _ = _device_field("pion", _engine.FP64, rp, p_i, p_f, dict(base, noise=1e-3, seed=6))
    assert np.array_equal(F1, F2)
    assert not np.array_equal(F1, F3)
    d = F1 - F0
    assert 0.4e-3 < np.std(d) < 1.2e-3  # N(0, 1e-3) per component, then renormalised
    N = 20 ** 3
    s = np.sqrt(np.sum(F1.reshape(4, N) ** 2, axis=0))
    assert np.max(np.abs(s - 1.0)) < 1e-14
