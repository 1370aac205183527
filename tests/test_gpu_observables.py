"""GPU parity of the post-convergence observables (``skx_observable``) against the oracle.

The engine sums the per-site integrands in FP64 on the device; the oracle
forms the same per-site arrays in NumPy.  The sums are compared to
``1e-12 * sum |per-site value|`` (the scale of summation-order rounding);
the Simulation-level results (centre of mass, RMS radius, tensors, D-term)
to 1e-11 relative.  An FP32 engine is checked on its own stored field
(downloaded exactly), so the same bound applies.
"""

from __future__ import annotations

import numpy as np
import pytest

from helpers import THEORY_NAMES, make_params, smooth_field
from oracle import integrator as OI
from oracle import observables as OO
from skyrmion_solver_amd import _engine
from skyrmion_solver_amd.core.simulation import Simulation

pytestmark = pytest.mark.gpu

THEORY_IDS = {"pion": _engine.THEORY_PION, "omega": _engine.THEORY_OMEGA, "coulomb": _engine.THEORY_COULOMB,
              "rho": _engine.THEORY_RHO, "berger": _engine.THEORY_BERGER, "lightly": _engine.THEORY_LIGHTLY_BOUND}
KINDS_ALL = ("com", "rms", "quadrupole", "U", "V", "W", "D")
SKX_EINVAL = -1
KIND_IDS = {"com": _engine.OBS_COM, "rms": _engine.OBS_RMS, "quadrupole": _engine.OBS_QUADRUPOLE,
            "U": _engine.OBS_U, "V": _engine.OBS_V, "W": _engine.OBS_W, "D": _engine.OBS_D}


def oracle_sites(kind, theory, F, p_i, p_f, com):
    if kind == "com":
        return OO.moments_sites(theory, F, p_i, p_f)
    if kind == "rms":
        return OO.rms_sites(theory, F, p_i, p_f, com)
    fn = {"quadrupole": OO.quadrupole_sites, "U": OO.U_sites, "V": OO.V_sites, "W": OO.W_sites,
          "D": OO.D_sites}[kind]
    return fn(F, p_i, p_f, com)


def check_sums(eng, theory, F, p_i, p_f, kinds):
    com = OO.center_of_mass(theory, F, p_i, p_f) + np.array([0.013, -0.021, 0.008])
    for kind in kinds:
        got = eng.observable(KIND_IDS[kind], com)
        sites = oracle_sites(kind, theory, F, p_i, p_f, com)
        nv = got.size
        sites = sites.reshape(nv, -1)
        ref = sites.sum(axis=1)
        scale = np.abs(sites).sum(axis=1)
        err = np.abs(got - ref)
        assert np.all(err <= 1e-12 * scale + 1e-300), (kind, got, ref, err / np.maximum(scale, 1e-300))


@pytest.mark.parametrize("theory", ["pion", "omega", "coulomb", "rho", "berger", "lightly"])
@pytest.mark.parametrize("shape", [(20, 18, 16), (9, 13, 24)])
def test_observable_sums_fp64(theory, shape):
    _, rp, p_i, p_f = make_params(theory, *shape)
    F = smooth_field(rp, int(p_i[5]), seed=5 + shape[2], normalise=True)
    kinds = KINDS_ALL if theory == "pion" else ("com", "rms")
    print(f"[case] observables theory={theory} lattice={shape} fp64", flush=True)
    with _engine.Engine(THEORY_IDS[theory], _engine.FP64, p_i, p_f) as eng:
        eng.upload(_engine.FIELD, F)
        check_sums(eng, theory, F, p_i, p_f, kinds)


@pytest.mark.parametrize("theory", ["pion", "rho"])
def test_observable_sums_fp32_storage(theory):
    shape = (22, 17, 19)
    _, rp, p_i, p_f = make_params(theory, *shape)
    F = smooth_field(rp, int(p_i[5]), seed=11, normalise=True)
    kinds = KINDS_ALL if theory == "pion" else ("com", "rms")
    print(f"[case] observables theory={theory} lattice={shape} fp32", flush=True)
    with _engine.Engine(THEORY_IDS[theory], _engine.FP32, p_i, p_f) as eng:
        eng.upload(_engine.FIELD, F)
        F32 = eng.download(_engine.FIELD).reshape(F.shape)
        assert np.max(np.abs(F32 - F)) < 1e-6
        check_sums(eng, theory, F32, p_i, p_f, kinds)


def _sim(theory, shape, **kw):
    from skyrmion_solver_amd.theories import load_theory

    mod = load_theory(THEORY_NAMES[theory])
    _, rp, p_i, p_f = make_params(theory, *shape, **kw)
    return Simulation(rp, mod), rp, p_i, p_f


def test_simulation_observables_pion():
    sim, rp, p_i, p_f = _sim("pion", (26, 24, 28))
    try:
        F = smooth_field(rp, 4, seed=21, normalise=True)
        sim.Field.copy_to_device(F)
        st = OI.OracleState("pion", p_i, p_f, F)
        print("[case] Simulation observables pion 26x24x28", flush=True)
        np.testing.assert_allclose(sim.compute_center_of_mass(), OO.center_of_mass("pion", F, p_i, p_f),
                                   rtol=1e-11, atol=1e-13)
        rms = sim.compute_rms_radius()
        assert rms == pytest.approx(OO.rms_radius("pion", F, p_i, p_f), rel=1e-11)
        for name, fn in (("compute_U_tensor", OO.U_tensor), ("compute_V_tensor", OO.V_tensor),
                         ("compute_W_tensor", OO.W_tensor), ("compute_quadrupole_tensor", OO.quadrupole)):
            got = getattr(sim, name)()
            ref = fn(F, p_i, p_f)
            assert got.shape == (3, 3)
            np.testing.assert_allclose(got, ref, rtol=0, atol=1e-11 * np.abs(ref).max(), err_msg=name)
        D = sim.compute_D_term()
        assert D == pytest.approx(OO.D_term(F, p_i, p_f, OI.energy(st)), rel=1e-11)
    finally:
        sim.close()


@pytest.mark.parametrize("theory", ["omega", "coulomb", "rho", "berger", "lightly"])
def test_simulation_observables_other_theories(theory):
    sim, rp, p_i, p_f = _sim(theory, (20, 22, 18))
    try:
        F = smooth_field(rp, int(p_i[5]), seed=4, normalise=True)
        sim.Field.copy_to_device(F)
        print(f"[case] Simulation observables {theory} 20x22x18", flush=True)
        np.testing.assert_allclose(sim.compute_center_of_mass(), OO.center_of_mass(theory, F, p_i, p_f),
                                   rtol=1e-11, atol=1e-13)
        assert sim.compute_rms_radius() == pytest.approx(OO.rms_radius(theory, F, p_i, p_f), rel=1e-11)
        # the reference's other observables modules have no inertia tensors or D-term
        with pytest.raises(AttributeError):
            sim.compute_U_tensor()
        with pytest.raises(AttributeError):
            sim.compute_D_term()
        # and the engine refuses the nuclear-only kinds for this theory
        with pytest.raises(_engine.SkxError):
            sim.engine.observable(_engine.OBS_V, np.zeros(3))
    finally:
        sim.close()


def test_observable_argument_errors():
    _, rp, p_i, p_f = make_params("pion", 12, 12, 12)
    with _engine.Engine(_engine.THEORY_PION, _engine.FP64, p_i, p_f) as eng:
        out = np.zeros(9)
        ptr = out.ctypes.data_as(_engine.C.POINTER(_engine.C.c_double))
        lib, h = eng._lib, eng._h
        assert lib.skx_observable(h, 7, None, ptr, 9) == SKX_EINVAL
        assert lib.skx_observable(h, _engine.OBS_V, None, ptr, 9) == SKX_EINVAL  # V needs the com
        assert lib.skx_observable(h, _engine.OBS_U, None, ptr, 8) == SKX_EINVAL  # U has 9 values
        assert lib.skx_observable(h, _engine.OBS_U, None, ptr, 9) == 0  # U reads no com


def test_translation_moves_the_centre_of_mass_exactly():
    # size-independent property at a production-sized box: an integer roll of a
    # field that is exact vacuum near the faces shifts the centre of mass by the step
    shape = (96, 96, 96)
    _, rp, p_i, p_f = make_params("pion", *shape, lsx=0.15, lsy=0.15, lsz=0.15)
    F = smooth_field(rp, 4, seed=2, noise=0.0, normalise=True)
    idx = np.indices(shape)
    far = ((idx[0] - 47.5) ** 2 + (idx[1] - 47.5) ** 2 + (idx[2] - 47.5) ** 2) > 40 ** 2
    F[:, far] = np.array([1.0, 0.0, 0.0, 0.0])[:, None]
    shift = (3, -2, 4)
    print("[case] CoM translation 96^3 fp64", flush=True)
    with _engine.Engine(_engine.THEORY_PION, _engine.FP64, p_i, p_f) as eng:
        eng.upload(_engine.FIELD, F)
        m0 = eng.observable(_engine.OBS_COM)
        eng.upload(_engine.FIELD, np.roll(F, shift, axis=(1, 2, 3)))
        m1 = eng.observable(_engine.OBS_COM)
    assert abs(m1[0] - m0[0]) < 1e-12 * abs(m0[0])
    np.testing.assert_allclose(m1[1:4] / m1[0] - m0[1:4] / m0[0], np.array(shift) * 0.15, rtol=0, atol=1e-11)
