"""GPU: the allocation-free observable entry point equals the allocating one.

skx_observable_noalloc runs the same observable kernel as skx_observable but
keeps its block partials in the library's static per-device array instead
of a buffer allocated on the handle's first observable call.  Two calls in a
row on lattices of decreasing size (the second leaves stale partials of the
first beyond its own blocks) must give exactly skx_observable's values, for
every observable kind, in FP64 and with FP32 storage.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from helpers import make_params, smooth_field
from skyrmion_solver_amd import _engine

pytestmark = pytest.mark.gpu

KINDS = (_engine.OBS_COM, _engine.OBS_RMS, _engine.OBS_QUADRUPOLE, _engine.OBS_U, _engine.OBS_V, _engine.OBS_W,
         _engine.OBS_D)


def _noalloc(eng, kind, com):
    out = np.full(9, np.nan)
    c = np.ascontiguousarray(com, dtype=np.float64)
    rc = eng._lib.skx_observable_noalloc(eng._h, int(kind), _engine._dptr(c), _engine._dptr(out), out.size)
    _engine._check(rc, "skx_observable_noalloc")
    return out[:_engine.OBS_COUNT[kind]]


@pytest.mark.parametrize("prec", [_engine.FP64, _engine.FP32])
def test_noalloc_observables_equal_allocating_entry(prec):
    for shape, seed in (((40, 36, 44), 3), ((20, 18, 16), 4)):  # second call: a smaller lattice
        print(f"[case] observable_noalloc prec={prec} lattice={shape}", flush=True)
        mod, rp, p_i, p_f = make_params("pion", *shape)
        F = smooth_field(rp, 4, seed=seed, normalise=True, band=True)
        with _engine.Engine(_engine.THEORY_PION, prec, p_i, p_f) as eng:
            eng.upload(_engine.FIELD, F)
            com = eng.observable(_engine.OBS_COM)
            com3 = com[1:4] / com[0]
            for kind in KINDS:
                ref = eng.observable(kind, com3)
                got = _noalloc(eng, kind, com3)
                assert np.array_equal(got, ref), (shape, kind, got, ref)
                again = _noalloc(eng, kind, com3)
                assert np.array_equal(again, ref), (shape, kind)


def test_noalloc_argument_errors():
    print("[case] observable_noalloc argument errors", flush=True)
    mod, rp, p_i, p_f = make_params("pion", 12, 12, 12)
    with _engine.Engine(_engine.THEORY_PION, _engine.FP64, p_i, p_f) as eng:
        out = np.zeros(9)
        com = np.zeros(3)
        lib = eng._lib
        assert lib.skx_observable_noalloc(eng._h, 99, _engine._dptr(com), _engine._dptr(out), 9) == -1
        assert lib.skx_observable_noalloc(eng._h, _engine.OBS_U, _engine._dptr(com), _engine._dptr(out), 4) == -1
        assert lib.skx_observable_noalloc(eng._h, _engine.OBS_RMS, C.POINTER(C.c_double)(), _engine._dptr(out),
                                          9) == -1
        assert lib.skx_observable_noalloc(None, _engine.OBS_COM, _engine._dptr(com), _engine._dptr(out), 9) == -1
