"""GPU: the autotuner picks a timed candidate and caches it; batch lattices with an x-chunk equal the
solo engine with the same x-chunk bit for bit."""

from __future__ import annotations

import numpy as np
import pytest

from helpers import make_params, smooth_field
from skyrmion_solver_amd import _engine, autotune

pytestmark = pytest.mark.gpu


def test_tune_engine_picks_fastest_and_caches(tmp_path, monkeypatch):
    print("[case] autotune engine 24x20x40 pion fp32", flush=True)
    monkeypatch.setenv("SKX_TUNE_CACHE", str(tmp_path / "tune.json"))
    mod, rp, p_i, p_f = make_params("pion", 24, 20, 40)
    r = autotune.tune_engine(_engine.THEORY_PION, _engine.FP32, p_i, p_f, steps=4, xchunks=[0, 8])
    assert len(r["table"]) == len(autotune.ENGINE_TILES) * 2
    best = min(r["table"], key=lambda t: t["us_per_step"])
    assert (r["tile_z"], r["x_chunk"]) == (best["tile_z"], best["x_chunk"])
    again = autotune.tune_engine(_engine.THEORY_PION, _engine.FP32, p_i, p_f, steps=4, xchunks=[0, 8])
    assert again == r  # served from the cache


@pytest.mark.parametrize("theory", ["pion", "rho"])
def test_batch_xchunk_equals_solo_xchunk(theory):
    print(f"[case] batch x_chunk theory={theory}", flush=True)
    shapes = [(26, 18, 16), (21, 16, 34)]
    lat = []
    for k, sh in enumerate(shapes):
        mod, rp, p_i, p_f = make_params(theory, *sh)
        lat.append((p_i, p_f, smooth_field(rp, int(p_i[5]), seed=30 + k, normalise=True)))
    xc = 8
    with _engine.Batch(_engine.THEORY_PION if theory == "pion" else _engine.THEORY_RHO, _engine.FP64,
                       max_slots=4, max_blocks=512) as b:
        slots = [b.add(p_i, p_f, F, max_steps=12, x_chunk=xc) for p_i, p_f, F in lat]
        b.run(12)
        res = [b.remove(s) for s in slots]
    tid = _engine.THEORY_PION if theory == "pion" else _engine.THEORY_RHO
    for (p_i, p_f, F), (st, Fb, Vb) in zip(lat, res):
        with _engine.Engine(tid, _engine.FP64, p_i, p_f, tile_z=32, x_chunk=xc) as e:
            e.upload(_engine.FIELD, F)
            ref = e.anf_run(12, -1.0, 4)
            Fs = e.download(_engine.FIELD)
        assert st["steps"] == ref.steps == 12
        assert st["energy"] == ref.energy
        assert np.array_equal(Fb, Fs)
