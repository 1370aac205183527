"""``Simulation.save_output`` / ``load_output`` / checkpoints on the GPU engine.

The bundle's field, grid and lattice files must be byte-identical to the
oracle's writer fed with the engine's own field; the energy and baryon
density files hold device FP64 values, so they are parsed and compared with
the oracle's densities to 1e-12 relative.
"""

from __future__ import annotations

import os

import numpy as np
import pytest

from helpers import THEORY_NAMES, make_params, smooth_field
from oracle import integrator as OI
from oracle import io as OIO
from skyrmion_solver_amd.core.simulation import Simulation
from skyrmion_solver_amd.theories import load_theory

pytestmark = pytest.mark.gpu


def _sim(theory, shape):
    mod = load_theory(THEORY_NAMES[theory])
    _, rp, p_i, p_f = make_params(theory, *shape)
    return Simulation(rp, mod), mod, rp, p_i, p_f


@pytest.mark.parametrize("theory", ["pion", "rho"])
def test_save_output_matches_reference_format(tmp_path, theory):
    shape = (9, 7, 11)
    sim, mod, rp, p_i, p_f = _sim(theory, shape)
    try:
        nf = int(p_i[5])
        F = smooth_field(rp, nf, seed=13, normalise=True)
        sim.Field.copy_to_device(F)
        print(f"[case] save_output {theory} {shape}", flush=True)
        sim.save_output(str(tmp_path / "n"), precision=32)
        st = OI.OracleState(theory, p_i, p_f, F)
        E, B = OI.energy_density(st).ravel(), OI.baryon_density(st).ravel()
        X, Y, Z = shape
        g = OIO.grid(X, Y, Z, p_f[3], p_f[4], p_f[5])
        OIO.output_data_bundle_core(str(tmp_path / "o"), h_Field=F.ravel(), h_grid=g, xlen=X, ylen=Y, zlen=Z,
                                    number_coordinates=3, number_total_fields=nf, precision=32,
                                    bundle_spec=mod.io.BUNDLE_SPEC,
                                    arrays={"grid": g, "Field": F.ravel(), "EnergyDensity": E, "BaryonDensity": B},
                                    field_dump_name="Field.dat")
        names = sorted(os.listdir(tmp_path / "o"))
        assert sorted(os.listdir(tmp_path / "n")) == names
        for fn in names:
            a = (tmp_path / "n" / fn).read_bytes()
            b = (tmp_path / "o" / fn).read_bytes()
            if fn in ("EnergyDensity.dat", "BaryonDensity.dat"):
                va, vb = np.array(a.split(), dtype=np.float64), np.array(b.split(), dtype=np.float64)
                assert a.count(b"\n") == b.count(b"\n") and va.size == vb.size == X * Y * Z
                assert np.max(np.abs(va - vb)) <= 1e-12 * np.max(np.abs(vb)), fn
            else:
                assert a == b, fn
    finally:
        sim.close()


def test_load_output_and_checkpoint_round_trip(tmp_path):
    sim, mod, rp, p_i, p_f = _sim("pion", (10, 9, 8))
    try:
        F = smooth_field(rp, 4, seed=2, normalise=True)
        V = 0.01 * smooth_field(rp, 4, seed=3)
        sim.Field.copy_to_device(F)
        sim.Velocity.copy_to_device(V)
        print("[case] load_output / checkpoint pion 10x9x8", flush=True)
        sim.save_output(str(tmp_path / "t"))
        sim.save_output(str(tmp_path / "b"), binary=True)
        sim.save_checkpoint(str(tmp_path / "c.npz"))
        for d in ("t", "b"):
            sim.Field.copy_to_device(np.zeros_like(F))
            sim.load_output(str(tmp_path / d))
            np.testing.assert_array_equal(sim.Field.copy_to_host(), F.ravel())
            assert not sim.Velocity.copy_to_host().any()
        sim.load_checkpoint(str(tmp_path / "c.npz"))
        np.testing.assert_array_equal(sim.Velocity.copy_to_host(), V.ravel())
        E0 = sim.observables()["energy"]
        assert E0 == pytest.approx(OI.energy(OI.OracleState("pion", p_i, p_f, F)), rel=1e-12)
    finally:
        sim.close()
