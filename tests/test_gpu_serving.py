"""GPU: requests served in a mix (continuous batching, two theories on two
streams, staggered arrivals) give each request exactly its solo result."""

from __future__ import annotations

import numpy as np
import pytest

from helpers import make_params, smooth_field
from skyrmion_solver_amd import _engine, workloads
from skyrmion_solver_amd.serving import THEORY_IDS, Request, Server

pytestmark = pytest.mark.gpu


def build(theory, shape, seed, prec):
    kw = {"mpi": 0.6} if theory == "pion" else {"mass": 0.5, "c_omega": 0.5}
    mod, rp, p_i, p_f = make_params(theory, *shape, **kw)
    F = smooth_field(rp, int(p_i[5]), seed=seed, normalise=True)
    if theory == "omega":
        with _engine.Engine(THEORY_IDS[theory], prec, p_i, p_f) as e:
            e.upload(_engine.FIELD, F)
            F = workloads.set_omega_from_chern_simons(e, rp)
    return p_i, p_f, F


@pytest.mark.parametrize("prec", [_engine.FP32, _engine.FP64])
def test_served_results_equal_solo(prec):
    print(f"[case] serving mix prec={prec}", flush=True)
    specs = [("pion", (20, 18, 16)), ("omega", (18, 18, 20)), ("pion", (24, 16, 18)), ("pion", (16, 20, 22)),
             ("omega", (20, 20, 16))]
    reqs = []
    for k, (th, sh) in enumerate(specs):
        p_i, p_f, F = build(th, sh, 40 + k, prec)
        reqs.append(Request(rid=k, theory=th, p_i=p_i, p_f=p_f, F0=F, tol=2.0, max_steps=120 + 10 * k,
                            arrival=0.003 * k))
    srv = Server(precision=prec, hbm_budget_bytes=1 << 30, chunk_steps=8, max_slots=8, max_blocks=2048,
                 tile_z=32, x_chunk_target=0, x_chunk_blocks=0)  # the solo runs below: same tile and chunk, bit for bit
    rep = srv.serve(reqs)
    srv.close()
    print("  report", rep, flush=True)
    assert rep["requests"] == len(reqs)
    for r in reqs:
        with _engine.Engine(THEORY_IDS[r.theory], prec, r.p_i, r.p_f, tile_z=32, x_chunk=int(r.p_i[0])) as e:
            e.upload(_engine.FIELD, r.F0)
            st = e.anf_run(r.max_steps, r.tol, 8)
            F = e.download(_engine.FIELD)
        assert r.status["steps"] == st.steps, r.rid
        assert r.status["energy"] == st.energy, r.rid
        assert np.array_equal(r.F, F), r.rid


def test_fp32_refinement_equals_the_split_run_by_hand():
    """FP32 server with FP64 refinement: each request equals, bit for bit, an FP32 run in the server's
    chunks until max|g| < factor * tol, continued by an FP64 run from the downloaded field and velocity."""
    print("[case] serving fp32 -> fp64 refinement", flush=True)
    specs = [("pion", (20, 18, 16)), ("pion", (16, 20, 22)), ("omega", (18, 18, 20))]
    chunk, factor, tol = 8, 4.0, 2.0
    reqs = []
    for k, (th, sh) in enumerate(specs):
        p_i, p_f, F = build(th, sh, 60 + k, _engine.FP32)
        reqs.append(Request(rid=k, theory=th, p_i=p_i, p_f=p_f, F0=F, tol=tol, max_steps=160, arrival=0.002 * k))
    srv = Server(precision=_engine.FP32, hbm_budget_bytes=1 << 30, chunk_steps=chunk, max_slots=4, max_blocks=512,
                 tile_z=32, x_chunk_target=0, x_chunk_blocks=0, refine_fp64=True, refine_factor=factor)
    rep = srv.serve(reqs)
    srv.close()
    print("  report", rep, flush=True)
    assert rep["requests"] == len(reqs) and rep["refined"] >= 1
    for r in reqs:
        tid, X = THEORY_IDS[r.theory], int(r.p_i[0])
        with _engine.Engine(tid, _engine.FP32, r.p_i, r.p_f, tile_z=32, x_chunk=X) as e:
            e.upload(_engine.FIELD, r.F0)
            steps, split = 0, False
            while steps < r.max_steps:
                st = e.anf_run(min(chunk, r.max_steps - steps), r.tol, chunk)
                steps += st.steps
                if st.converged:
                    break
                if st.err_abs < factor * r.tol and steps < r.max_steps:
                    split = True
                    break
            F, V, last = e.download(_engine.FIELD), e.download(_engine.VELOCITY), st
        assert r.refined == split, r.rid
        if split:
            assert r.steps_fp32 == steps, r.rid
            with _engine.Engine(tid, _engine.FP64, r.p_i, r.p_f, tile_z=32, x_chunk=X) as e:
                e.upload(_engine.FIELD, F)
                e.upload(_engine.VELOCITY, V)
                last = e.anf_run(r.max_steps - steps, r.tol, chunk)
                F = e.download(_engine.FIELD)
            steps += last.steps
        assert r.status["steps"] == steps, r.rid
        assert r.status["energy"] == last.energy, r.rid
        assert np.array_equal(r.F, F), r.rid
