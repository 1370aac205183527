"""GPU parity of the HIP engine against the FP64 oracle (through the C ABI).

Tolerances: the FP64 engine must match the oracle to 1e-12 relative
(norm-wise) on gradients, fields and energies; the FP32 form is checked
against the FP64 oracle with the per-result tolerances stated in each test.
"""

from __future__ import annotations

import numpy as np
import pytest

from helpers import make_params, rel_l2, smooth_field
from oracle import integrator as OI
from oracle import stencils as OS
from oracle import theories as OT
from skyrmion_solver_amd import _engine

pytestmark = pytest.mark.gpu

THEORY_IDS = {"pion": _engine.THEORY_PION, "omega": _engine.THEORY_OMEGA, "coulomb": _engine.THEORY_COULOMB,
              "rho": _engine.THEORY_RHO, "berger": _engine.THEORY_BERGER, "lightly": _engine.THEORY_LIGHTLY_BOUND}
SHAPES = [(20, 18, 16), (9, 7, 13), (24, 40, 70)]


def _engine_for(theory, rp, p_i, p_f, prec=_engine.FP64, **kw):
    print(f"[case] theory={theory} lattice={rp.xlen}x{rp.ylen}x{rp.zlen} prec={prec} opts={kw}", flush=True)
    return _engine.Engine(THEORY_IDS[theory], prec, p_i, p_f, **kw)


@pytest.mark.parametrize("theory", ["pion", "omega", "coulomb", "rho", "berger", "lightly"])
@pytest.mark.parametrize("shape", SHAPES)
def test_gradient_probe_fp64(theory, shape):
    mod, rp, p_i, p_f = make_params(theory, *shape)
    nf = int(p_i[5])
    T = smooth_field(rp, nf, seed=17 * len(theory) + shape[0] + 3 * shape[2])
    st = OI.OracleState(theory, p_i, p_f, T)
    EG_ref, l_ref = OI.gradient(st, st.F)
    with _engine_for(theory, rp, p_i, p_f) as eng:
        eg, l = eng.eval_gradient(T)
    assert rel_l2(eg, EG_ref) < 1e-12
    assert rel_l2(l, l_ref) < 1e-12
    # pointwise: every site within 1e-10 of the largest gradient entry
    assert np.max(np.abs(eg - EG_ref.ravel())) <= 1e-10 * np.max(np.abs(EG_ref))


# thinner than the stencil along some axes, down to one site (every site in the boundary band)
DEGENERATE = [(1, 1, 1), (2, 3, 4), (4, 4, 4), (5, 1, 6), (6, 5, 1), (1, 33, 9), (7, 2, 40)]


@pytest.mark.parametrize("prec", [_engine.FP64, _engine.FP32])
@pytest.mark.parametrize("theory", ["pion", "omega", "rho"])
@pytest.mark.parametrize("shape", DEGENERATE)
def test_degenerate_lattices(shape, theory, prec):
    """Lattices with axes shorter than the five-point stencil, down to 1x1x1: the gradient, and three ANF
    steps of the device loop, against the FP64 oracle (FP64: 1e-12; FP32 storage: 1e-5 of the run's scale)."""
    mod, rp, p_i, p_f = make_params(theory, *shape)
    nf = int(p_i[5])
    T = smooth_field(rp, nf, seed=5, normalise=True)
    st = OI.OracleState(theory, p_i, p_f, T)
    EG_ref, _ = OI.gradient(st, st.F)
    E = E0 = OI.energy(st)
    ref = []
    for _ in range(3):
        E, err = OI.anf_step(st, E)
        ref.append((E, err))
    with _engine_for(theory, rp, p_i, p_f, prec=prec) as eng:
        eg, _ = eng.eval_gradient(T)
        eng.upload(_engine.FIELD, T)
        run = eng.anf_run(3, -1.0, 3)
        Fg = eng.download(_engine.FIELD)
    tol = 1e-12 if prec == _engine.FP64 else 1e-5
    assert np.max(np.abs(eg - EG_ref.ravel())) <= tol * np.max(np.abs(EG_ref))
    assert run.steps == 3
    escale = max(abs(E0), max(abs(e) for e, _ in ref))
    assert abs(run.energy - ref[-1][0]) <= tol * escale
    assert abs(run.err_signed - ref[-1][1]) <= (1e-10 if prec == _engine.FP64 else 1e-5) * abs(ref[-1][1])
    assert rel_l2(Fg, st.F) < (1e-12 if prec == _engine.FP64 else 1e-5)


@pytest.mark.parametrize("theory", ["pion", "rho"])
@pytest.mark.parametrize("tile_z", [16, 32, 64])
def test_gradient_probe_tiles(tile_z, theory):
    mod, rp, p_i, p_f = make_params(theory, 19, 21, 67)
    nf = int(p_i[5])
    T = smooth_field(rp, nf, seed=7)
    st = OI.OracleState(theory, p_i, p_f, T)
    EG_ref, _ = OI.gradient(st, st.F)
    for xc in (3, 8, 19):
        with _engine_for(theory, rp, p_i, p_f, tile_z=tile_z, x_chunk=xc) as eng:
            eg, _ = eng.eval_gradient(T)
        assert rel_l2(eg, EG_ref) < 1e-12, (tile_z, xc)


@pytest.mark.parametrize("theory", ["pion", "omega", "coulomb", "rho", "berger", "lightly"])
def test_energy_baryon_fp64(theory):
    mod, rp, p_i, p_f = make_params(theory, 22, 20, 18)
    nf = int(p_i[5])
    F = smooth_field(rp, nf, seed=3, normalise=True)
    st = OI.OracleState(theory, p_i, p_f, F)
    with _engine_for(theory, rp, p_i, p_f) as eng:
        eng.upload(_engine.FIELD, F)
        E = eng.energy()
        B = eng.baryon()
        de = eng.density(_engine.DENSITY_ENERGY)
        db = eng.density(_engine.DENSITY_BARYON)
    E_ref = OI.energy(st)
    B_ref = OI.baryon_number(st)
    assert abs(E - E_ref) <= 1e-12 * abs(E_ref)
    assert abs(B - B_ref) <= 1e-12 * max(1.0, abs(B_ref))
    assert rel_l2(de, OI.energy_density(st)) < 1e-12
    assert rel_l2(db, OI.baryon_density(st)) < 1e-12


@pytest.mark.parametrize("theory", ["pion", "omega", "coulomb", "rho", "berger", "lightly"])
def test_anf_steps_fp64(theory):
    mod, rp, p_i, p_f = make_params(theory, 20, 18, 16)
    nf = int(p_i[5])
    F = smooth_field(rp, nf, seed=11, normalise=True)
    st = OI.OracleState(theory, p_i, p_f, F)
    E_prev = OI.energy(st)
    with _engine_for(theory, rp, p_i, p_f) as eng:
        eng.upload(_engine.FIELD, F)
        eng.upload(_engine.VELOCITY, np.zeros_like(F))
        Eg = E_prev
        for k in range(6):
            E_ref, err_ref = OI.anf_step(st, E_prev)
            E_gpu, err_gpu, errabs_gpu = eng.anf_step(Eg)
            assert abs(E_gpu - E_ref) <= 1e-12 * abs(E_ref), k
            assert abs(err_gpu - err_ref) <= 1e-10 * max(abs(err_ref), 1e-300), k
            assert abs(errabs_gpu - OI.max_abs_force(st)) <= 1e-10 * OI.max_abs_force(st), k
            E_prev, Eg = E_ref, E_gpu
        Fg = eng.download(_engine.FIELD)
        Vg = eng.download(_engine.VELOCITY)
    assert rel_l2(Fg, st.F) < 1e-12
    assert rel_l2(Vg, st.V) < 1e-10 or np.max(np.abs(st.V)) == 0.0


@pytest.mark.parametrize("theory", ["pion", "omega", "coulomb", "rho", "berger", "lightly"])
def test_long_trajectory_fp64(theory):
    """40 steps of the device-resident loop (arrest decided on the device) against 40 oracle steps on a
    lattice that is no multiple of any tile: every step's energy to 1e-12 of the run's energy scale
    (norm-wise: on this small box an energy can pass close to zero), the final field to 1e-11."""
    mod, rp, p_i, p_f = make_params(theory, 14, 12, 16)
    F = smooth_field(rp, int(p_i[5]), seed=23, normalise=True)
    st = OI.OracleState(theory, p_i, p_f, F)
    E = E0 = OI.energy(st)
    ref, arrests = [], 0
    for _ in range(40):
        E, _err = OI.anf_step(st, E)
        ref.append(E)
        arrests += int(np.max(np.abs(st.V)) == 0.0)
    with _engine_for(theory, rp, p_i, p_f) as eng:
        eng.upload(_engine.FIELD, F)
        eng.history_attach(48)
        run = eng.anf_run(40, -1.0, 10)
        hist = eng.history_read(48)
        Fg = eng.download(_engine.FIELD)
    print(f"  {theory}: arrests {arrests}, E0 {hist[0][0]:.12g} -> {run.energy:.12g}", flush=True)
    assert run.steps == 40 and len(hist) == 40
    scale = max(abs(E0), max(abs(e) for e in ref))
    for k in range(39):  # the history holds the energy entering each step
        assert abs(hist[k + 1][0] - ref[k]) <= 1e-12 * scale, k
    assert abs(run.energy - ref[-1]) <= 1e-12 * scale
    assert rel_l2(Fg, st.F) < 1e-11


def test_arrest_applied_like_reference():
    """Large dt makes the energy rise: the velocity must be zeroed exactly when E_new > E_prev."""
    mod, rp, p_i, p_f = make_params("pion", 16, 16, 16, time_step=0.04)
    F = smooth_field(rp, 4, seed=5, normalise=True)
    st = OI.OracleState("pion", p_i, p_f, F)
    with _engine_for("pion", rp, p_i, p_f) as eng:
        eng.upload(_engine.FIELD, F)
        E_prev = Eg = OI.energy(st)
        arrests = 0
        for k in range(12):
            E_ref, _ = OI.anf_step(st, E_prev)
            E_gpu, _, _ = eng.anf_step(Eg)
            Vg = eng.download(_engine.VELOCITY)
            if np.max(np.abs(st.V)) == 0.0:
                arrests += 1
                assert np.max(np.abs(Vg)) == 0.0
            assert rel_l2(Vg, st.V) < 1e-9 or np.max(np.abs(st.V)) == 0.0
            E_prev, Eg = E_ref, E_gpu
    print("arrests:", arrests)


@pytest.mark.parametrize("theory", ["pion", "coulomb", "rho", "berger"])
def test_run_equals_steps_bitwise(theory):
    """The device-resident loop (fused energy, deferred arrest, hipGraph) equals repeated step()."""
    mod, rp, p_i, p_f = make_params(theory, 18, 17, 20, time_step=0.04)
    nf = int(p_i[5])
    F = smooth_field(rp, nf, seed=2, normalise=True)
    with _engine_for(theory, rp, p_i, p_f) as a, _engine_for(theory, rp, p_i, p_f) as b:
        a.upload(_engine.FIELD, F)
        b.upload(_engine.FIELD, F)
        E = a.energy()
        for _ in range(9):
            E, err, errabs = a.anf_step(E)
        stats = b.anf_run(9, -1.0, 4)
        assert stats.steps == 9
        assert stats.energy == E
        assert stats.err_signed == err
        assert np.array_equal(a.download(_engine.FIELD), b.download(_engine.FIELD))
        assert np.array_equal(a.download(_engine.VELOCITY), b.download(_engine.VELOCITY))


def test_run_no_graph_equals_graph():
    mod, rp, p_i, p_f = make_params("pion", 16, 18, 20, time_step=0.04)
    F = smooth_field(rp, 4, seed=9, normalise=True)
    with _engine_for("pion", rp, p_i, p_f) as a, \
            _engine_for("pion", rp, p_i, p_f, flags=_engine.FLAG_NO_GRAPH) as b:
        a.upload(_engine.FIELD, F)
        b.upload(_engine.FIELD, F)
        sa = a.anf_run(11, -1.0, 3)
        sb = b.anf_run(11, -1.0, 5)
        assert sa.steps == sb.steps == 11
        assert sa.energy == sb.energy
        assert np.array_equal(a.download(_engine.FIELD), b.download(_engine.FIELD))


@pytest.mark.parametrize("theory", ["pion", "omega", "coulomb", "rho", "berger", "lightly"])
def test_fp32_form_against_fp64_oracle(theory):
    """FP32 storage/stencils: after 10 steps, |dE|/E <= 1e-5, |dB| <= 1e-4, field rel-L2 <= 1e-5."""
    mod, rp, p_i, p_f = make_params(theory, 20, 18, 16)
    nf = int(p_i[5])
    F = smooth_field(rp, nf, seed=4, normalise=True)
    st = OI.OracleState(theory, p_i, p_f, F)
    E_prev = OI.energy(st)
    with _engine_for(theory, rp, p_i, p_f, prec=_engine.FP32) as eng:
        eng.upload(_engine.FIELD, F)
        Eg = E_prev
        for _ in range(10):
            E_ref, _ = OI.anf_step(st, E_prev)
            Eg, _, _ = eng.anf_step(Eg)
            E_prev = E_ref
        Fg = eng.download(_engine.FIELD)
        B = eng.baryon()
    B_ref = OI.baryon_number(st)
    assert abs(Eg - E_ref) <= 1e-5 * abs(E_ref)
    assert abs(B - B_ref) <= 1e-4
    assert rel_l2(Fg, st.F) <= 1e-5


WIDE_STAGES = {"pion": {4}, "omega": {4}, "coulomb": {4}, "berger": {0, 1, 4}, "lightly": {0, 1, 4}}


@pytest.mark.parametrize("theory", ["pion", "omega", "coulomb", "berger", "lightly"])
def test_fp32_default_mixed_tiles_equal_uniform_tile(theory):
    """FP32 lattices with Y >= 32 run some stages on a 32x16 tile and the others on 32x8 (the runtime's
    stage_tile, reported by skx_stage_shape); the per-site arithmetic does not depend on the tile, so the
    gradient and the fields after the same steps are bit-identical to an all-32x8 run (tile_z=32) and
    the energies agree to reduction order."""
    mod, rp, p_i, p_f = make_params(theory, 22, 40, 36)
    F = smooth_field(rp, int(p_i[5]), seed=8, normalise=True)
    st = OI.OracleState(theory, p_i, p_f, F)
    with _engine_for(theory, rp, p_i, p_f, prec=_engine.FP32) as a, \
            _engine_for(theory, rp, p_i, p_f, prec=_engine.FP32, tile_z=32) as b:
        for k in range(5):
            assert a.stage_shape(k)["tile_y"] == (16 if k in WIDE_STAGES[theory] else 8), k
            assert b.stage_shape(k)["tile_y"] == 8, k
        a.upload(_engine.FIELD, F)
        b.upload(_engine.FIELD, F)
        ea, eb = a.energy(), b.energy()
        assert abs(ea - eb) <= 1e-13 * abs(eb)
        assert abs(ea - OI.energy(st)) <= 1e-5 * abs(OI.energy(st))
        ga, _ = a.eval_gradient(F)
        gb, _ = b.eval_gradient(F)
        assert np.array_equal(ga, gb)
        sa = a.anf_run(12, -1.0, 4)
        sb = b.anf_run(12, -1.0, 4)
        assert sa.steps == sb.steps == 12
        assert sa.err_signed == sb.err_signed and sa.err_abs == sb.err_abs
        assert abs(sa.energy - sb.energy) <= 1e-13 * abs(sb.energy)
        assert np.array_equal(a.download(_engine.FIELD), b.download(_engine.FIELD))
        assert np.array_equal(a.download(_engine.VELOCITY), b.download(_engine.VELOCITY))

