"""GPU: batched launches give every lattice exactly its solo result.

A batch runs lattices of different sizes and parameters with one launch
per RK4 stage; each lattice must end bit-identical to the same lattice run
alone (the solo engine uses the same one-chunk x decomposition, so even the
FP64 reductions are summed in the same order), and lattices that join while
others are half-way through must not perturb them.
"""

from __future__ import annotations

import numpy as np
import pytest

from helpers import make_params, smooth_field
from skyrmion_solver_amd import _engine

pytestmark = pytest.mark.gpu

TID = {"pion": _engine.THEORY_PION, "omega": _engine.THEORY_OMEGA, "coulomb": _engine.THEORY_COULOMB,
       "rho": _engine.THEORY_RHO, "berger": _engine.THEORY_BERGER, "lightly": _engine.THEORY_LIGHTLY_BOUND}


def lattices(theory, shapes, seed0=0):
    out = []
    for k, sh in enumerate(shapes):
        kw = {"mpi": 0.5 + 0.2 * k} if theory in ("pion", "rho", "berger", "lightly") else {"mass": 0.4 + 0.1 * k}
        mod, rp, p_i, p_f = make_params(theory, *sh, **kw)
        F = smooth_field(rp, int(p_i[5]), seed=seed0 + k, normalise=True)
        out.append((p_i, p_f, F))
    return out


def solo(theory, prec, p_i, p_f, F, steps, tol):
    with _engine.Engine(TID[theory], prec, p_i, p_f, tile_z=32, x_chunk=int(p_i[0])) as e:
        e.upload(_engine.FIELD, F)
        st = e.anf_run(steps, tol, 16)
        return st, e.download(_engine.FIELD), e.download(_engine.VELOCITY)


@pytest.mark.parametrize("theory,prec", [("pion", _engine.FP64), ("pion", _engine.FP32),
                                         ("coulomb", _engine.FP32), ("omega", _engine.FP64),
                                         ("rho", _engine.FP64), ("rho", _engine.FP32),
                                         ("berger", _engine.FP32), ("lightly", _engine.FP64)])
def test_batch_equals_solo_bitwise(theory, prec):
    print(f"[case] batch theory={theory} prec={prec}", flush=True)
    lat = lattices(theory, [(20, 18, 16), (16, 16, 16), (24, 21, 34)])
    steps = [30, 24, 30]
    with _engine.Batch(TID[theory], prec, max_slots=8, max_blocks=2048) as b:
        slots = [b.add(p_i, p_f, F, max_steps=s) for (p_i, p_f, F), s in zip(lat, steps)]
        for _ in range(4):
            b.run(8)
        res = [b.remove(s) for s in slots]
    for (p_i, p_f, F), s, (st, Fb, Vb) in zip(lat, steps, res):
        ref, Fs, Vs = solo(theory, prec, p_i, p_f, F, s, -1.0)
        assert st["steps"] == ref.steps == s
        assert st["energy"] == ref.energy
        assert st["err_abs"] == ref.err_abs
        assert np.array_equal(Fb, Fs)
        assert np.array_equal(Vb, Vs)


@pytest.mark.parametrize("theory,prec", [("pion", _engine.FP32), ("berger", _engine.FP32),
                                         ("lightly", _engine.FP32), ("omega", _engine.FP32),
                                         ("coulomb", _engine.FP32), ("rho", _engine.FP32),
                                         ("pion", _engine.FP64)])
def test_batch_tile_per_stage_equals_solo(theory, prec):
    """tile_z 0: FP32 stages on 32 x 16 tiles through the second block map (lattices with Y below
    and above 32), a lattice leaving half-way through and another joining.  Fields and forces are
    bit-identical to the lattice run alone with its default tiles; the energy agrees to reduction
    order (a lattice with Y < 32 runs 32 x 8 alone, so its blocks partition the sum differently)."""
    print(f"[case] batch tile per stage theory={theory} prec={prec}", flush=True)
    lat = lattices(theory, [(20, 40, 36), (16, 18, 16), (24, 33, 34), (18, 36, 20)], seed0=20)
    with _engine.Batch(TID[theory], prec, max_slots=4, max_blocks=96, tile_z=0) as b:
        s0 = b.add(*lat[0], max_steps=32)
        s1 = b.add(*lat[1], max_steps=16)
        s2 = b.add(*lat[2], max_steps=32)
        b.run(16)
        out = {1: b.remove(s1)}
        s3 = b.add(*lat[3], max_steps=16)  # takes free blocks of both maps after lattice 1 left
        b.run(16)
        out.update({0: b.remove(s0), 2: b.remove(s2), 3: b.remove(s3)})
    for k, steps in ((0, 32), (1, 16), (2, 32), (3, 16)):
        p_i, p_f, F = lat[k]
        with _engine.Engine(TID[theory], prec, p_i, p_f, x_chunk=int(p_i[0])) as e:
            e.upload(_engine.FIELD, F)
            ref = e.anf_run(steps, -1.0, 16)
            Fs, Vs = e.download(_engine.FIELD), e.download(_engine.VELOCITY)
        st, Fb, Vb = out[k]
        assert st["steps"] == ref.steps == steps
        assert abs(st["energy"] - ref.energy) <= 1e-13 * abs(ref.energy)
        assert st["err_abs"] == ref.err_abs
        assert np.array_equal(Fb, Fs)
        assert np.array_equal(Vb, Vs)


def test_late_joiner_does_not_perturb_running_lattices():
    print("[case] batch late joiner", flush=True)
    lat = lattices("pion", [(18, 18, 18), (22, 16, 20), (16, 20, 18)], seed0=10)
    with _engine.Batch(TID["pion"], _engine.FP32, max_slots=4, max_blocks=1024) as b:
        s0 = b.add(*lat[0], max_steps=40)
        b.run(10)
        s1 = b.add(*lat[1], max_steps=40)
        b.run(10)
        s2 = b.add(*lat[2], max_steps=40)
        for _ in range(4):
            b.run(10)
        out = {s: b.remove(s) for s in (s0, s1, s2)}
    for (p_i, p_f, F), s in zip(lat, (s0, s1, s2)):
        ref, Fs, _ = solo("pion", _engine.FP32, p_i, p_f, F, 40, -1.0)
        st, Fb, _ = out[s]
        assert st["steps"] == 40
        assert st["energy"] == ref.energy
        assert np.array_equal(Fb, Fs)


def test_batch_convergence_stops_each_lattice():
    print("[case] batch convergence", flush=True)
    lat = lattices("pion", [(16, 16, 16), (18, 18, 18)], seed0=20)
    tol = 5.0
    with _engine.Batch(TID["pion"], _engine.FP64, max_slots=4, max_blocks=1024) as b:
        slots = [b.add(p_i, p_f, F, max_steps=2000, tol=tol) for p_i, p_f, F in lat]
        for _ in range(20):
            b.run(16)
            if all(s["done"] for s in b.poll()):
                break
        res = [b.remove(s) for s in slots]
    for (p_i, p_f, F), (st, Fb, _) in zip(lat, res):
        ref, Fs, _ = solo("pion", _engine.FP64, p_i, p_f, F, 2000, tol)
        assert st["converged"] == 1 and ref.converged == 1
        assert st["steps"] == ref.steps
        assert np.array_equal(Fb, Fs)


def test_batch_rejects_when_full():
    lat = lattices("pion", [(16, 16, 16)])
    with _engine.Batch(TID["pion"], _engine.FP32, max_slots=1, max_blocks=64) as b:
        b.add(*lat[0])
        with pytest.raises(_engine.SkxError) as e:
            b.add(*lat[0])
        assert e.value.code == -5


def test_batch_rejects_a_lattice_larger_than_its_grid():
    lat = lattices("pion", [(16, 40, 64)])  # 2 x 5 tiles of 32 x 8 per x-chunk
    with _engine.Batch(TID["pion"], _engine.FP32, max_slots=2, max_blocks=8, tile_z=32) as b:
        with pytest.raises(_engine.SkxError) as e:
            b.add(*lat[0])
        assert e.value.code == -1  # SKX_EINVAL: it could never fit, unlike a busy grid (SKX_EBUSY)


def test_block_maps_under_churn():
    """Ten lattices of random shapes through a batch whose grid holds two or three at a time, tile per
    stage (two block maps): they join as blocks free up, so first fit reuses and fragments both maps.
    Every result equals the lattice run alone (energy to reduction order: Y < 32 runs 32 x 8 alone)."""
    print("[case] batch block maps under churn", flush=True)
    rng = np.random.default_rng(4)
    shapes = [(int(rng.integers(14, 22)), int(rng.choice([18, 34, 40])), int(rng.integers(16, 40))) for _ in range(10)]
    lat = lattices("pion", shapes, seed0=40)
    steps = [16 + 8 * int(rng.integers(0, 3)) for _ in shapes]
    results, order = {}, []
    with _engine.Batch(TID["pion"], _engine.FP32, max_slots=4, max_blocks=24, tile_z=0) as b:
        queue, live = list(range(len(lat))), {}
        for _ in range(200):
            while queue:
                k = queue[0]
                try:
                    s = b.add(*lat[k], max_steps=steps[k])
                except _engine.SkxError as e:
                    if e.code == -5:
                        break
                    raise
                live[s] = k
                order.append(k)
                queue.pop(0)
            if not live:
                break
            b.run(8)
            for st in b.poll():
                if st["done"] and st["slot"] in live:
                    k = live.pop(st["slot"])
                    results[k] = b.remove(st["slot"])
    assert sorted(results) == list(range(len(lat))) and order == list(range(len(lat)))
    for k, (p_i, p_f, F) in enumerate(lat):
        with _engine.Engine(TID["pion"], _engine.FP32, p_i, p_f, x_chunk=int(p_i[0])) as e:
            e.upload(_engine.FIELD, F)
            ref = e.anf_run(steps[k], -1.0, 8)
            Fs = e.download(_engine.FIELD)
        st, Fb, _ = results[k]
        assert st["steps"] == ref.steps == steps[k], k
        assert abs(st["energy"] - ref.energy) <= 1e-13 * abs(ref.energy), k
        assert np.array_equal(Fb, Fs), k
