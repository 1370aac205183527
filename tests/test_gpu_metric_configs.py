"""GPU parity at the configurations the measurements are quoted on.

* 128^3 B=4 pion (the headline workload): two FP64 ANF steps against the
  FP64 oracle (core/integrator.py:284-410 with
  theories/nuclear_skyrme/kernels.py:1034-1150), and the FP32 form relaxed
  to max|g| < 1e-4 against the FP64 form relaxed the same way, seeds 0..19.
* The 48-lattice batch ({48, 64, 96, 128}^3, B = 1..6, m in [0, 1.5]) with
  the batch's own launch shape (a tile per stage, ~32-plane x-chunks): every
  lattice bit-identical to its solo run after several chunks.
* 160^3 B=8 omega and Coulomb: the device-built initial field against the
  oracle's, then two FP64 ANF steps against the oracle
  (theories/omega_skyrme/kernels.py:586-743, theories/coulomb_skyrme/kernels.py:584-767).
* A reduced serving stream of the theory mix at 64^3..192^3: every request
  served in the mix bit-identical to its solo run, FP32 and FP64.
"""

from __future__ import annotations

import os

import numpy as np
import pytest

from helpers import ill_conditioned_sites, oracle_field, rel_l2
from oracle import initial_config as OIC
from oracle import integrator as OI
from oracle import parallel as OP
from skyrmion_solver_amd import _engine, autotune, workloads
from skyrmion_solver_amd.serving import THEORY_IDS, Server

pytestmark = pytest.mark.gpu


def test_c2_fp64_steps_match_oracle():
    """128^3 B=4: two ANF steps (host-side arrest, the reference's step()) and the same two steps of the
    device-resident loop; energy, signed err and max|g| of each step and the final field to 1e-12."""
    print("[case] 128^3 B=4 pion fp64: 2 steps vs oracle", flush=True)
    th, rp, p_i, p_f, F = workloads.c2_b4(seed=0)
    st = OI.OracleState("pion", p_i, p_f, F)
    E0 = OI.energy(st)
    ref = []
    E = E0
    for _ in range(2):
        E, err = OI.anf_step(st, E)
        ref.append((E, err, OI.max_abs_force(st)))
    print(f"  oracle E {ref[0][0]:.15g} {ref[1][0]:.15g}", flush=True)
    with _engine.Engine(_engine.THEORY_PION, _engine.FP64, p_i, p_f) as eng:
        eng.upload(_engine.FIELD, F)
        assert abs(eng.energy() - E0) <= 1e-12 * abs(E0)
        Eg = E0
        for k in range(2):
            Eg, err, errabs = eng.anf_step(Eg)
            assert abs(Eg - ref[k][0]) <= 1e-12 * abs(ref[k][0]), k
            assert abs(err - ref[k][1]) <= 1e-10 * abs(ref[k][1]), k
            assert abs(errabs - ref[k][2]) <= 1e-10 * ref[k][2], k
        Fg = eng.download(_engine.FIELD)
        Vg = eng.download(_engine.VELOCITY)
    assert rel_l2(Fg, st.F) < 1e-12
    assert rel_l2(Vg, st.V) < 1e-10
    with _engine.Engine(_engine.THEORY_PION, _engine.FP64, p_i, p_f) as eng:
        eng.upload(_engine.FIELD, F)
        run = eng.anf_run(2, -1.0, 2)
        assert run.steps == 2
        assert abs(run.energy - ref[1][0]) <= 1e-12 * abs(ref[1][0])
        assert abs(run.err_signed - ref[1][1]) <= 1e-10 * abs(ref[1][1])
        assert rel_l2(eng.download(_engine.FIELD), st.F) < 1e-12


def test_c2_fp32_form_converges_to_the_fp64_solution():
    """128^3 B=4, seeds 0..19 (the noise of each start), both forms relaxed to max|g| < 1e-4.  The FP32
    form against the FP64 form (which matches the oracle step for step above), every seed: |dE|/E <= 1e-6,
    |dB| <= 1e-4, field rel-L2 <= 1e-4."""
    worst = dict(dE=0.0, dB=0.0, L2=0.0)
    for seed in range(20):
        print(f"[case] 128^3 B=4 pion fp32 vs fp64 to convergence, seed {seed}", flush=True)
        th, rp, p_i, p_f, F = workloads.c2_b4(seed=seed)
        out = {}
        for prec in (_engine.FP64, _engine.FP32):
            with _engine.Engine(_engine.THEORY_PION, prec, p_i, p_f) as eng:
                eng.upload(_engine.FIELD, F)
                st = eng.anf_run(20000, 1.0e-4, 64)
                out[prec] = (st, eng.download(_engine.FIELD))
        (s64, F64), (s32, F32) = out[_engine.FP64], out[_engine.FP32]
        dE = abs(s32.energy - s64.energy) / abs(s64.energy)
        dB = abs(s32.baryon - s64.baryon)
        L2 = rel_l2(F32, F64)
        print(f"  steps {s64.steps}/{s32.steps} E {s64.energy:.12g}/{s32.energy:.12g} dE/E {dE:.2e} dB {dB:.2e} "
              f"L2 {L2:.2e}", flush=True)
        assert s64.converged == 1 and s32.converged == 1, seed
        assert abs(s64.baryon - 4.0) < 0.05, seed
        assert dE <= 1e-6 and dB <= 1e-4 and L2 <= 1e-4, seed
        worst = dict(dE=max(worst["dE"], dE), dB=max(worst["dB"], dB), L2=max(worst["L2"], L2))
    print(f"  worst over 20 seeds: dE/E {worst['dE']:.2e} dB {worst['dB']:.2e} L2 {worst['L2']:.2e}", flush=True)


def _oracle_threads() -> int:
    return max(1, min(16, len(os.sched_getaffinity(0))))


@pytest.mark.parametrize("theory", ["omega", "coulomb"])
def test_c3_fp64_steps_match_oracle(theory):
    """160^3 B=8 (product of two B=4 rational maps): the device-built field (omega = -c CS for the omega
    theory) equals the oracle's to 1e-12 on every site where the ansatz is well conditioned (the lump
    centres sit half a spacing off the sites, so some sites lie exactly on the B = 4 map's pole
    directions, where the value is set by rounding: at most 0.1% of the sites, and for omega their
    stencil neighbourhoods); from the device's field two FP64 ANF steps match the oracle (its
    slab-parallel driver, bit-identical to the single-threaded one) in energy and signed err to
    1e-12 / 1e-10 and in the field to 1e-12 relative L2."""
    print(f"[case] 160^3 B=8 {theory} fp64: init and 2 steps vs oracle", flush=True)
    tid = _engine.THEORY_OMEGA if theory == "omega" else _engine.THEORY_COULOMB
    _, rp, p_i, p_f, init = workloads.c3_b8(theory, device_init=True)
    ref0 = oracle_field(rp, p_i, p_f, 5, workloads.c3_b8_config())
    if theory == "omega":
        OIC.create_omega_field(ref0, p_i, p_f, float(rp.c_omega))
    pool = OP.SlabPool(int(p_i[0]), _oracle_threads())
    try:
        with _engine.Engine(tid, _engine.FP64, p_i, p_f) as eng:
            eng.init_field(init)
            F0 = eng.download(_engine.FIELD)
            bad = ill_conditioned_sites(rp, p_i, p_f, workloads.c3_b8_config())
            print(f"  ill-conditioned sites: {int(bad.sum())} of {bad.size}", flush=True)
            assert bad.sum() <= 1e-3 * bad.size
            d = np.abs(F0.reshape(ref0.shape) - ref0)
            assert np.max(d[:4][:, ~bad]) <= 1e-12
            if theory == "omega":  # omega = -c CS: a stencil over the pion field
                wide = ill_conditioned_sites(rp, p_i, p_f, workloads.c3_b8_config(), dilate=2)
                assert np.max(d[4][~wide]) <= 1e-12 * max(1.0, np.max(np.abs(ref0[4])))
            st = OI.OracleState(theory, p_i, p_f, F0)
            E = OP.energy(st, pool)
            assert abs(eng.energy() - E) <= 1e-12 * abs(E)
            Eg = E
            for k in range(2):
                E, err = OP.anf_step(st, E, pool)
                Eg, errg, _ = eng.anf_step(Eg)
                print(f"  step {k}: E {E:.15g} engine {Eg:.15g}", flush=True)
                assert abs(Eg - E) <= 1e-12 * abs(E), k
                assert abs(errg - err) <= 1e-10 * abs(err), k
            assert rel_l2(eng.download(_engine.FIELD), st.F) < 1e-12
    finally:
        pool.close()


def test_c4_batch_of_48_equals_solo_bitwise():
    """The 48 lattices of the batched workload in one batch with its launch shape (tile per stage,
    ~32-plane x-chunks), fields built on the device; after 3 chunks of 8 steps every lattice's field,
    velocity, energy and forces equal its solo run with the same tiles and x-chunk, bit for bit."""
    print("[case] c4 48-lattice batch vs solo", flush=True)
    inst = workloads.c4_instances(n=48, seed=0, device_init=True)
    target, steps = 32, 24
    dims = [(d["L"],) * 3 for d in inst]
    blocks = autotune.batch_blocks(dims, 0, target)
    with _engine.Batch(_engine.THEORY_PION, _engine.FP32, max_slots=len(inst), max_blocks=blocks, tile_z=0) as b:
        slots = [b.add(d["p_i"], d["p_f"], init=d["init"], max_steps=steps,
                       x_chunk=autotune.chunk_rule(d["L"], target)) for d in inst]
        for _ in range(3):
            b.run(8)
        res = [b.remove(s) for s in slots]
    for k, (d, (stb, Fb, Vb)) in enumerate(zip(inst, res)):
        xc = autotune.chunk_rule(d["L"], target) or d["L"]
        with _engine.Engine(_engine.THEORY_PION, _engine.FP32, d["p_i"], d["p_f"], x_chunk=xc) as e:
            e.init_field(d["init"])
            st = e.anf_run(steps, -1.0, 8)
            Fs, Vs = e.download(_engine.FIELD), e.download(_engine.VELOCITY)
        assert stb["steps"] == st.steps == steps, k
        assert stb["energy"] == st.energy, k
        assert stb["err_abs"] == st.err_abs and stb["err_signed"] == st.err_signed, k
        assert np.array_equal(Fb, Fs), k
        assert np.array_equal(Vb, Vs), k


@pytest.mark.parametrize("prec", [_engine.FP32, _engine.FP64])
def test_c5_reduced_stream_equals_solo(prec):
    """Nine requests of the serving stream's theory mix (pion / omega / rho) at 64^3..192^3 with staggered
    arrivals and different step budgets, served by continuous batching with the server's defaults (tile
    per stage, about 256 blocks per lattice, fields built on the device at admission); each result equals
    the same request run alone with the same x-chunk, bit for bit."""
    print(f"[case] c5 reduced stream 64^3..192^3 prec={prec}", flush=True)
    reqs = workloads.c5_requests(n=9, rate=50.0, seed=3, sizes=(64, 96, 128, 160, 192), device_init=True)
    for k, r in enumerate(reqs):
        r.max_steps = 24 + 8 * (k % 4)
    print("  requests", [(r.theory, int(r.p_i[0])) for r in reqs], flush=True)
    srv = Server(precision=prec, hbm_budget_bytes=48 << 30, chunk_steps=8, max_slots=16, max_blocks=32768)
    rep = srv.serve(reqs)
    srv.close()
    assert rep["requests"] == len(reqs)
    for r in reqs:
        L = int(r.p_i[0])
        xc = srv.x_chunk(r) or L
        with _engine.Engine(THEORY_IDS[r.theory], prec, r.p_i, r.p_f, x_chunk=xc) as e:
            e.init_field(r.init)
            st = e.anf_run(r.max_steps, r.tol, 8)
            F = e.download(_engine.FIELD)
        assert r.status["steps"] == st.steps == r.max_steps, r.rid
        assert r.status["energy"] == st.energy, r.rid
        assert r.status["err_abs"] == st.err_abs, r.rid
        assert np.array_equal(r.F, F), r.rid
