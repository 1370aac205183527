"""Host-side checks: packed parameter ABI, theory registry, and the C-ABI
library's exported symbols (no device calls)."""

from __future__ import annotations

import os
import re

import numpy as np
import pytest

from skyrmion_solver_amd import _engine
from skyrmion_solver_amd.theories import get_theory_spec, list_theories, load_theory

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_registry_names_and_aliases():
    assert set(list_theories()) == {"Nuclear Skyrme model", "Omega-Skyrme model", "Coulomb-Skyrme model",
                                    "Rho-Skyrme model", "Berger Skyrme model", "Lightly bound Skyrme model"}
    assert get_theory_spec("skyrmions").name == "Nuclear Skyrme model"
    assert get_theory_spec("  MAXWELL skyrme ").name == "Coulomb-Skyrme model"
    assert get_theory_spec("squashed sphere skyrme").name == "Berger Skyrme model"
    assert get_theory_spec("Loosely bound Skyrme model").name == "Lightly bound Skyrme model"
    with pytest.raises(KeyError):
        get_theory_spec("baby skyrme")


def test_core_packing_layout():
    th = load_theory("Nuclear Skyrme model")
    p = th.params.default_params(xlen=10, ylen=11, zlen=12, xsize=9.0, ysize=10.0, zsize=11.0, courant=0.5,
                                 mpi=0.3, kappa=1.7, ansatz="rational_map")
    rp = p.resolved()
    p_i, p_f = th.params.pack_device_params(rp)
    assert p_i.dtype == np.int32 and p_f.dtype == np.float64
    assert list(p_i[:11]) == [10, 11, 12, 2, 3, 4, 1320, 5280, 1, 1, 1]
    assert list(p_i[11:]) == [0, 1, 0]
    assert p_f[3] == 9.0 / 9 and p_f[4] == 10.0 / 10 and p_f[5] == 11.0 / 11
    assert p_f[6] == p_f[3] * p_f[4] * p_f[5]
    assert p_f[7] == 0.5 * p_f[3]
    assert p_f[8] == 0.3 and p_f[9] == 1.7


def test_berger_and_lightly_bound_packing():
    """p_f[8..12] = mpi, kappa, B, isospin_rotation, alpha (theories/berger_skyrme/params.py:237-243);
    p_f[8..11] = mpi, alpha, B, isospin_rotation (theories/lightly_bound_skyrme/params.py:234-239)."""
    be = load_theory("Berger Skyrme model").params
    p_i, p_f = be.pack_device_params(be.default_params(mpi=0.4, kappa=1.3, alpha=0.7, baryon_number=2.0,
                                                       isospin_rotation=0.5, ansatz="rational_map").resolved())
    assert list(p_f[8:]) == [0.4, 1.3, 2.0, 0.5, 0.7] and list(p_i[11:]) == [0, 1, 0] and p_i[5] == 4
    lb = load_theory("Lightly bound Skyrme model").params
    p_i, p_f = lb.pack_device_params(lb.default_params(mpi=0.6, alpha=0.2, baryon_number=3.0).resolved())
    assert list(p_f[8:]) == [0.6, 0.2, 3.0, 0.0] and list(p_i[11:]) == [1, 0, 0] and p_i[5] == 4


def test_theory_derived_couplings():
    om = load_theory("Omega-Skyrme model").params.default_params().resolved()
    assert om.mass == 138.0 / 782.0 and om.c_omega == 782.0 * 8.253 / 186.0
    co = load_theory("Coulomb-Skyrme model").params.default_params().resolved()
    assert co.mass == 2 * 138.0 / (4.0 * 105.9) and co.kappa == 4.0 * 4.0 * 0.302822 * 0.302822 / 2.0
    rh = load_theory("Rho-Skyrme model").params
    r = rh.default_params().resolved()
    p_i, p_f = rh.pack_device_params(r)
    assert p_i[5] == 13 and p_f[10] == r.c_alpha == 0.0617
    with pytest.raises(ValueError):
        load_theory("Nuclear Skyrme model").params.default_params(ansatz="torus").resolved()


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "skx.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(skx_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = _engine.load_library()
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_engine.EXPORTED)
    assert lib.skx_version() == 1


def test_invalid_descriptor_is_rejected_without_a_device():
    th = load_theory("Nuclear Skyrme model")
    rp = th.params.default_params(xlen=8, ylen=8, zlen=8).resolved()
    p_i, p_f = th.params.pack_device_params(rp)
    bad = p_i.copy()
    bad[3] = 1  # halo below the stencil radius
    with pytest.raises(_engine.SkxError) as e:
        _engine.footprint_bytes(_engine.THEORY_PION, _engine.FP64, bad, p_f)
    assert e.value.code == -1
    n = _engine.footprint_bytes(_engine.THEORY_PION, _engine.FP32, p_i, p_f)
    assert n >= 7 * 4 * 512 * 4
    # an empty lattice, a size that disagrees with its dimensions, a non-positive spacing
    for k, v in ((0, 0), (6, 511), (7, 4 * 511)):
        bad = p_i.copy()
        bad[k] = v
        with pytest.raises(_engine.SkxError) as e:
            _engine.footprint_bytes(_engine.THEORY_PION, _engine.FP64, bad, p_f)
        assert e.value.code == -1, k
    bad_f = p_f.copy()
    bad_f[4] = 0.0
    with pytest.raises(_engine.SkxError) as e:
        _engine.footprint_bytes(_engine.THEORY_PION, _engine.FP64, p_i, bad_f)
    assert e.value.code == -1


def test_host_chern_simons_density_matches_oracle():
    """initial_data.chern_simons_density (omega starts at -c CS) against the oracle's CS of the same field."""
    import numpy as np

    from helpers import make_params, smooth_field
    from oracle import stencils as OS
    from oracle import theories as OT
    from skyrmion_solver_amd import initial_data

    mod, rp, p_i, p_f = make_params("omega", 14, 13, 12)
    T = smooth_field(rp, 5, seed=3, normalise=True)
    d1 = OS.first_derivatives(T, (rp.lsx, rp.lsy, rp.lsz))
    ref = OT.chern_simons(T, d1)
    got = initial_data.chern_simons_density(T, rp)
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))
