"""GPU: the device-resident loop launches exactly the steps it is asked for.

A run replays captured chunks of `check_every` steps only while a whole
chunk remains; the remainder runs as direct launches, so no step is ever
enqueued past `max_steps`.  Both parities' graphs are captured before the
first launch of a run (or by `prepare`), never between launches.  The
reference advances one step per `Simulation.step` call
(core/simulation.py:287-308 -> core/integrator.py:284-410).
"""

from __future__ import annotations

import numpy as np
import pytest

from helpers import make_params, smooth_field
from skyrmion_solver_amd import _engine

pytestmark = pytest.mark.gpu


def _setup():
    mod, rp, p_i, p_f = make_params("pion", 18, 20, 22, time_step=0.04)
    F = smooth_field(rp, 4, seed=31, normalise=True)
    return p_i, p_f, F


def test_short_run_issues_exactly_its_steps():
    print("[case] anf_run(20, -1, 100) step accounting", flush=True)
    p_i, p_f, F = _setup()
    with _engine.Engine(_engine.THEORY_PION, _engine.FP32, p_i, p_f) as e:
        e.upload(_engine.FIELD, F)
        st = e.anf_run(20, -1.0, 100)
        assert (st.steps, st.issued, st.graph_launches) == (20, 20, 0)
        st = e.anf_run(250, -1.0, 100)  # two captured chunks, then 50 direct steps
        assert (st.steps, st.issued, st.graph_launches) == (250, 250, 2)
        e.prepare(40)
        st = e.anf_run(120, -1.0, 40)
        assert (st.steps, st.issued, st.graph_launches) == (120, 120, 3)


def test_traced_run_counts_four_stage_launches_per_step():
    print("[case] traced anf_run(20, -1, 100)", flush=True)
    p_i, p_f, F = _setup()
    with _engine.Engine(_engine.THEORY_PION, _engine.FP32, p_i, p_f, flags=_engine.FLAG_TRACE) as e:
        e.upload(_engine.FIELD, F)
        e.trace_reset()
        st = e.anf_run(20, -1.0, 100)
        ms, cnt = e.trace_read()
    assert st.steps == st.issued == 20
    # stages 1..4 once per step; the energy pass once at the end of the run
    assert list(cnt) == [1, 20, 20, 20, 20]
    assert all(m > 0 for m in ms)


@pytest.mark.parametrize("prec", [_engine.FP32, _engine.FP64])
def test_chunked_graph_run_equals_direct_launches_bitwise(prec):
    """270 steps as 2 graph chunks + 70 direct steps (graph engine), as 27 direct chunks of 10 (no-graph
    engine) and as 270 traced launches: identical fields, velocities and energies."""
    print(f"[case] graph+tail vs direct prec={prec}", flush=True)
    p_i, p_f, F = _setup()
    out = []
    for flags, chunk in ((0, 100), (_engine.FLAG_NO_GRAPH, 10), (_engine.FLAG_TRACE, 64)):
        with _engine.Engine(_engine.THEORY_PION, prec, p_i, p_f, flags=flags) as e:
            e.upload(_engine.FIELD, F)
            st = e.anf_run(270, -1.0, chunk)
            out.append((st.steps, st.energy, st.err_signed, e.download(_engine.FIELD), e.download(_engine.VELOCITY)))
    for o in out[1:]:
        assert o[:3] == out[0][:3]
        assert np.array_equal(o[3], out[0][3])
        assert np.array_equal(o[4], out[0][4])


def test_converged_run_stops_at_the_first_converged_step():
    """A tolerance the first step meets: the run takes one step; the rest of the chunk skips its work."""
    print("[case] converged run accounting", flush=True)
    p_i, p_f, F = _setup()
    with _engine.Engine(_engine.THEORY_PION, _engine.FP64, p_i, p_f) as e:
        e.upload(_engine.FIELD, F)
        st = e.anf_run(50, 1.0e9, 16)
    assert st.steps == 1 and st.converged == 1
    assert st.issued == 16 and st.graph_launches == 1


def test_nonfinite_run_stops_at_the_same_step_on_every_path():
    """An omega run with dt far beyond RK4's stability limit goes non-finite within a few steps (the
    oracle at step 6).  A run without a tolerance enqueues its final energy pass right after its last
    chunk; when a non-finite value stopped it earlier, it must still report SKX_ENONFINITE and leave the
    handle on the field it stopped at, as a run with a tolerance (read after every chunk) and a run of
    direct launches do."""
    print("[case] non-finite omega run, three paths", flush=True)
    mod, rp, p_i, p_f = make_params("omega", 16, 16, 16, time_step=0.6)
    F = smooth_field(rp, 5, seed=1, normalise=True)
    outs = []
    for flags, tol in ((0, -1.0), (0, 0.0), (_engine.FLAG_NO_GRAPH, -1.0)):
        with _engine.Engine(_engine.THEORY_OMEGA, _engine.FP64, p_i, p_f, flags=flags) as e:
            e.upload(_engine.FIELD, F)
            e.history_attach(32)
            with pytest.raises(_engine.SkxError) as err:
                e.anf_run(16, tol, 16)
            assert err.value.code == -4
            hist = e.history_read(32)
            outs.append((len(hist), e.download(_engine.FIELD), e.download(_engine.VELOCITY)))
    print("  steps before the stop:", [o[0] for o in outs], flush=True)
    assert 0 < outs[0][0] < 16
    for o in outs[1:]:
        assert o[0] == outs[0][0]
        assert np.array_equal(o[1], outs[0][1], equal_nan=True)
        assert np.array_equal(o[2], outs[0][2], equal_nan=True)
