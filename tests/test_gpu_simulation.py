"""``Simulation.minimize`` on the GPU engine: the plain device loop, and the FP32 run
refined in FP64 against the same split made by hand."""

from __future__ import annotations

import numpy as np
import pytest

from helpers import THEORY_NAMES, make_params, smooth_field
from skyrmion_solver_amd import _engine
from skyrmion_solver_amd.core.simulation import Simulation
from skyrmion_solver_amd.theories import load_theory

pytestmark = pytest.mark.gpu


def _setup(shape=(20, 18, 22), seed=5):
    mod = load_theory(THEORY_NAMES["pion"])
    _, rp, p_i, p_f = make_params("pion", *shape, mpi=0.8)
    F = smooth_field(rp, 4, seed=seed, normalise=True)
    return mod, rp, p_i, p_f, F


def test_minimize_is_the_device_loop():
    print("[case] simulation minimize fp64", flush=True)
    mod, rp, p_i, p_f, F = _setup()
    sim = Simulation(rp, mod)
    sim.engine.upload(_engine.FIELD, F)
    res = sim.minimize(tol=1.0, max_steps=300)
    with _engine.Engine(_engine.THEORY_PION, _engine.FP64, p_i, p_f) as e:
        e.upload(_engine.FIELD, F)
        st = e.anf_run(300, 1.0, 0)
        Fe = e.download(_engine.FIELD)
    assert res["steps"] == st.steps and res["energy"] == st.energy and res["converged"] == st.converged
    assert np.array_equal(sim.engine.download(_engine.FIELD), Fe)


def test_fp32_minimize_refined_in_fp64_equals_the_split_by_hand():
    print("[case] simulation minimize fp32 -> fp64", flush=True)
    mod, rp, p_i, p_f, F = _setup()
    tol, factor, cap = 0.5, 4.0, 2000
    sim = Simulation(rp, mod, precision="fp32")
    sim.engine.upload(_engine.FIELD, F)
    res = sim.minimize(tol=tol, max_steps=cap, refine_fp64=True, refine_factor=factor)
    with _engine.Engine(_engine.THEORY_PION, _engine.FP32, p_i, p_f) as e:
        e.upload(_engine.FIELD, F)
        a = e.anf_run(cap, factor * tol, 0)
        Fa, Va = e.download(_engine.FIELD), e.download(_engine.VELOCITY)
    assert a.converged and a.err_abs >= tol, "the case must need the FP64 part"
    with _engine.Engine(_engine.THEORY_PION, _engine.FP64, p_i, p_f) as e:
        e.upload(_engine.FIELD, Fa)
        e.upload(_engine.VELOCITY, Va)
        b = e.anf_run(cap - a.steps, tol, 0)
        Fb = e.download(_engine.FIELD)
    assert res["refined"] and res["steps_fp32"] == a.steps and res["steps"] == a.steps + b.steps
    assert res["energy"] == b.energy and res["converged"] == b.converged == 1 and res["err_abs"] < tol
    assert sim.precision == "fp64"
    assert np.array_equal(sim.engine.download(_engine.FIELD), Fb)
