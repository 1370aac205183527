"""GPU: the FP64 pion stage kernels with two lanes per site against the one-lane kernels and the oracle.

RK stages 2-4 of the FP64 pion theory run two threads per site by default
(lane h holds components h and h + 2, contractions summed across the pair);
``FLAG_ONE_LANE`` selects the one-thread-per-site kernels.  Both are the
reference's per-site gradient and RK4 update
(theories/nuclear_skyrme/kernels.py:1077-1150, core/integrator.py:50-177);
they differ only in the order in which the four components enter a
contraction, so they agree to rounding and each matches the oracle to 1e-12.
"""

from __future__ import annotations

import numpy as np
import pytest

from helpers import make_params, rel_l2, smooth_field
from oracle import integrator as OI
from skyrmion_solver_amd import _engine

pytestmark = pytest.mark.gpu

SHAPES = [(20, 18, 16), (21, 40, 37), (12, 9, 70)]


def _steps(p_i, p_f, F, n, **kw):
    with _engine.Engine(_engine.THEORY_PION, _engine.FP64, p_i, p_f, **kw) as eng:
        eng.upload(_engine.FIELD, F)
        E = eng.energy()
        out = []
        for _ in range(n):
            E, err, errabs = eng.anf_step(E)
            out.append((E, err, errabs))
        return out, eng.download(_engine.FIELD), eng.download(_engine.VELOCITY)


@pytest.mark.parametrize("shape", SHAPES)
def test_two_lanes_and_one_lane_match_the_oracle(shape):
    """Six ANF steps (host arrest) with each kernel family: energy, signed err and max|g| of every step
    against the oracle to 1e-12 / 1e-10, the fields to 1e-12; the two families agree to 1e-13."""
    print(f"[case] lane pairs vs one lane vs oracle, shape={shape}", flush=True)
    mod, rp, p_i, p_f = make_params("pion", *shape)
    F = smooth_field(rp, 4, seed=5 + shape[0], normalise=True)
    st = OI.OracleState("pion", p_i, p_f, F)
    E = OI.energy(st)
    ref = []
    for _ in range(6):
        E, err = OI.anf_step(st, E)
        ref.append((E, err, OI.max_abs_force(st)))
    two, F2, V2 = _steps(p_i, p_f, F, 6)
    one, F1, V1 = _steps(p_i, p_f, F, 6, flags=_engine.FLAG_ONE_LANE)
    for k, ((Er, er, ar), (E2, e2, a2), (E1, e1, a1)) in enumerate(zip(ref, two, one)):
        for Eg, eg, ag in ((E2, e2, a2), (E1, e1, a1)):
            assert abs(Eg - Er) <= 1e-12 * abs(Er), k
            assert abs(eg - er) <= 1e-10 * abs(er), k
            assert abs(ag - ar) <= 1e-10 * ar, k
    assert rel_l2(F2, st.F) < 1e-12 and rel_l2(F1, st.F) < 1e-12
    assert rel_l2(F2, F1) < 1e-13
    assert rel_l2(V2, V1) < 1e-10


@pytest.mark.parametrize("tile_z", [16, 32, 64])
def test_two_lanes_on_every_tile(tile_z):
    """The device-resident loop on each tile (the 64 x 4 tile's stage 4 reads the raw F, V, l2 of its
    centre from memory, the others from the LDS ring) against the one-lane kernels: 30 steps, fields to
    1e-12, the final energy to 1e-13."""
    print(f"[case] lane pairs on tile_z={tile_z}", flush=True)
    mod, rp, p_i, p_f = make_params("pion", 22, 19, 70)
    F = smooth_field(rp, 4, seed=41, normalise=True)
    res = {}
    for name, flags in (("two", 0), ("one", _engine.FLAG_ONE_LANE)):
        with _engine.Engine(_engine.THEORY_PION, _engine.FP64, p_i, p_f, tile_z=tile_z, flags=flags) as eng:
            eng.upload(_engine.FIELD, F)
            run = eng.anf_run(30, -1.0, 10)
            res[name] = (run, eng.download(_engine.FIELD))
    (r2, F2), (r1, F1) = res["two"], res["one"]
    assert r2.steps == r1.steps == 30
    assert abs(r2.energy - r1.energy) <= 1e-13 * abs(r1.energy)
    assert rel_l2(F2, F1) < 1e-12


def test_one_lane_batch_equals_one_lane_solo_bitwise():
    """A batch created with FLAG_ONE_LANE runs the one-lane kernels for every lattice, bit-identical to the
    lattices run alone with the same flag (the default batch against the default solo run is covered by
    test_gpu_batch.py)."""
    print("[case] one-lane batch vs one-lane solo", flush=True)
    lats = []
    for k, shape in enumerate([(20, 18, 16), (16, 16, 16), (24, 21, 34)]):
        mod, rp, p_i, p_f = make_params("pion", *shape)
        lats.append((p_i, p_f, smooth_field(rp, 4, seed=60 + k, normalise=True)))
    with _engine.Batch(_engine.THEORY_PION, _engine.FP64, max_slots=4, max_blocks=2048,
                       flags=_engine.FLAG_ONE_LANE) as b:
        slots = [b.add(p_i, p_f, F, max_steps=24) for p_i, p_f, F in lats]
        for _ in range(3):
            b.run(8)
        res = [b.remove(s) for s in slots]
    for (p_i, p_f, F), (stb, Fb, Vb) in zip(lats, res):
        with _engine.Engine(_engine.THEORY_PION, _engine.FP64, p_i, p_f, flags=_engine.FLAG_ONE_LANE) as e:
            e.upload(_engine.FIELD, F)
            st = e.anf_run(24, -1.0, 8)
            Fs, Vs = e.download(_engine.FIELD), e.download(_engine.VELOCITY)
        assert stb["steps"] == st.steps == 24
        assert stb["energy"] == st.energy
        assert np.array_equal(Fb, Fs)
        assert np.array_equal(Vb, Vs)


def test_batch_rejects_a_lattice_beyond_the_lane_pair_range():
    """A 650^3 FP64 pion lattice (N + Y Z over 2^28 sites) cannot be addressed by the two-lane kernels:
    a default FP64 pion batch refuses it with SKX_EINVAL before allocating anything."""
    print("[case] lane-pair range check", flush=True)
    mod, rp, p_i, p_f = make_params("pion", 650, 650, 650)
    with _engine.Batch(_engine.THEORY_PION, _engine.FP64, max_slots=2, max_blocks=1 << 20) as b:
        with pytest.raises(_engine.SkxError) as ex:
            b.add(p_i, p_f, None, max_steps=1)
        assert ex.value.code == -1  # SKX_EINVAL
        assert "SKX_FLAG_ONE_LANE" in str(ex.value)
