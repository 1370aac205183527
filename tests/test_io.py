"""The .dat text output (``core/io.py`` of the reference) and its binary fast path.

The oracle (``oracle/io.py``) is pinned against files written out by hand from
the format the reference's loops produce; the native writer
(``skx_write_dat``, host code, no GPU) must then match the oracle byte for
byte on values chosen to exercise Python's ``.{p}g`` formatting: exact
integers, negative zero, values that switch to exponent form, subnormals,
infinities, NaN of both signs and random doubles, at several precisions.
"""

from __future__ import annotations

import os

import numpy as np
import pytest

from oracle import io as OIO
from skyrmion_solver_amd.core import io as SIO
from skyrmion_solver_amd.theories import load_theory


def _read(path):
    with open(path, "rb") as f:
        return f.read()


def test_oracle_field_file_by_hand(tmp_path):
    # two components of a 1 x 2 x 3 lattice, z fastest
    F = np.arange(12, dtype=np.float64)
    F[7] = 0.5
    OIO.output_field_dat(F, tmp_path / "f.dat", 1, 2, 3, 2, precision=32)
    want = "0\t1\t2\t\n3\t4\t5\t\n\n\n" "6\t0.5\t8\t\n9\t10\t11\t\n\n\n"
    assert _read(tmp_path / "f.dat") == want.encode()


def test_oracle_density_and_scalar_files_by_hand(tmp_path):
    d = np.array([1.0, -0.0, 1e-5, 123456.0, 0.1, 2.5, 3.0, 4.0])  # 2 x 2 x 2
    OIO.output_density_data_dat(d, 0, tmp_path / "d.dat", 2, 2, 2, precision=3)
    want = "1\t-0\t\n1e-05\t1.23e+05\t\n\n0.1\t2.5\t\n3\t4\t\n\n"
    assert _read(tmp_path / "d.dat") == want.encode()
    OIO.output_iteration_data_dat([128.0, 64.0, 0.1], tmp_path / "s.dat", precision=32)
    assert _read(tmp_path / "s.dat") == b"128\t64\t0.10000000000000000555111512312578\t"


def test_oracle_lattice_vectors_by_hand():
    g = OIO.grid(3, 4, 5, 0.5, 0.25, 2.0)
    np.testing.assert_array_equal(OIO.lattice_vectors(g, 3, 4, 5, 3),
                                  [1.0, 0.0, 0.0, 0.0, 0.75, 0.0, 0.0, 0.0, 8.0])


def special_values(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=n) * 10.0 ** rng.integers(-12, 12, size=n)
    specials = [0.0, -0.0, 1.0, -1.0, 0.1, 1 / 3, 1e-5, 1e-4, 9.999999e-5, 1e16, 1e17, 12345678901234567890.0,
                5e-324, -2.2250738585072014e-308, np.inf, -np.inf, np.nan, -np.nan, 2.0 ** 52, 0.5, 1e300]
    v[: len(specials)] = specials
    return v


@pytest.mark.parametrize("precision", [1, 3, 6, 16, 17, 32, 60, 100])
def test_native_writer_matches_oracle_field(tmp_path, precision):
    X, Y, Z, nf = 3, 5, 7, 2
    F = special_values(nf * X * Y * Z, precision)
    OIO.output_field_dat(F, tmp_path / "o.dat", X, Y, Z, nf, precision)
    SIO.output_field_dat(F, str(tmp_path / "n.dat"), X, Y, Z, nf, precision)
    assert _read(tmp_path / "n.dat") == _read(tmp_path / "o.dat")


@pytest.mark.parametrize("comp", [0, 2])
def test_native_writer_matches_oracle_density(tmp_path, comp):
    X, Y, Z = 6, 1, 9
    d = special_values(3 * X * Y * Z, 7 + comp)
    OIO.output_density_data_dat(d, comp, tmp_path / "o.dat", X, Y, Z, 32)
    SIO.output_density_data_dat(d, comp, str(tmp_path / "n.dat"), X, Y, Z, 32)
    assert _read(tmp_path / "n.dat") == _read(tmp_path / "o.dat")


def test_native_writer_many_slabs(tmp_path):
    # enough rows that the writer splits them across threads and slabs
    X, Y, Z = 70, 130, 3
    F = np.random.default_rng(1).normal(size=X * Y * Z)
    OIO.output_field_dat(F, tmp_path / "o.dat", X, Y, Z, 1, 17)
    SIO.output_field_dat(F, str(tmp_path / "n.dat"), X, Y, Z, 1, 17)
    assert _read(tmp_path / "n.dat") == _read(tmp_path / "o.dat")


def test_native_writer_errors(tmp_path):
    from skyrmion_solver_amd import _engine

    with pytest.raises(_engine.SkxError):
        SIO.output_field_dat(np.zeros(8), str(tmp_path / "p.dat"), 2, 2, 2, 1, precision=101)
    with pytest.raises(_engine.SkxError):
        SIO.output_field_dat(np.zeros(8), str(tmp_path / "no_such_dir" / "f.dat"), 2, 2, 2, 1)
    with pytest.raises(ValueError):
        SIO.output_field_dat(np.zeros(7), str(tmp_path / "f.dat"), 2, 2, 2, 1)


@pytest.mark.parametrize("name", ["Nuclear Skyrme model", "Omega-Skyrme model", "Coulomb-Skyrme model",
                                  "Rho-Skyrme model", "Berger Skyrme model", "Lightly bound Skyrme model"])
def test_theory_bundles_match_oracle(tmp_path, name):
    mod = load_theory(name)
    nf = {"Omega-Skyrme model": 5, "Coulomb-Skyrme model": 5, "Rho-Skyrme model": 13}.get(name, 4)
    X, Y, Z = 4, 3, 5
    rng = np.random.default_rng(3)
    F = rng.normal(size=nf * X * Y * Z)
    E = rng.normal(size=X * Y * Z)
    B = rng.normal(size=X * Y * Z)
    g = SIO.lattice_grid(X, Y, Z, 0.3, 0.2, 0.1)
    np.testing.assert_array_equal(g, OIO.grid(X, Y, Z, 0.3, 0.2, 0.1))
    mod.io.output_data_bundle(str(tmp_path / "n"), F, E, B, g, X, Y, Z, number_coordinates=3,
                              number_total_fields=nf, precision=32)
    OIO.output_data_bundle_core(str(tmp_path / "o"), h_Field=F, h_grid=g, xlen=X, ylen=Y, zlen=Z,
                                number_coordinates=3, number_total_fields=nf, precision=32,
                                bundle_spec=mod.io.BUNDLE_SPEC,
                                arrays={"grid": g, "Field": F, "EnergyDensity": E, "BaryonDensity": B},
                                lattice_points_name="LatticePoints.dat", lattice_vectors_name="LatticeVectors.dat",
                                field_dump_name="Field.dat")
    names = sorted(os.listdir(tmp_path / "o"))
    assert sorted(os.listdir(tmp_path / "n")) == names
    assert len(names) == 3 + 3 + nf + 2  # points, vectors, dump; grids; components; densities
    for fn in names:
        assert _read(tmp_path / "n" / fn) == _read(tmp_path / "o" / fn), fn


def test_binary_bundle_and_read_back(tmp_path):
    mod = load_theory("Nuclear Skyrme model")
    X, Y, Z = 5, 4, 3
    F = np.random.default_rng(9).normal(size=4 * X * Y * Z)
    g = SIO.lattice_grid(X, Y, Z, 0.1, 0.1, 0.1)
    mod.io.output_data_bundle(str(tmp_path / "b"), F, F[: X * Y * Z], F[: X * Y * Z], g, X, Y, Z, binary=True)
    mod.io.output_data_bundle(str(tmp_path / "t"), F, F[: X * Y * Z], F[: X * Y * Z], g, X, Y, Z)
    assert np.load(tmp_path / "b" / "Field.npy").shape == (4, X, Y, Z)
    np.testing.assert_array_equal(np.load(tmp_path / "b" / "PionField2.npy").ravel(), F[2 * X * Y * Z:3 * X * Y * Z])
    # the text dump at 32 significant digits reads back exactly, as does the binary one
    np.testing.assert_array_equal(SIO.read_field(str(tmp_path / "t"), X, Y, Z, 4), F)
    np.testing.assert_array_equal(SIO.read_field(str(tmp_path / "b"), X, Y, Z, 4), F)
    with pytest.raises(ValueError):
        SIO.read_field(str(tmp_path / "t"), X, Y, Z, 3)
