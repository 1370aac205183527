"""Nuclear Skyrme output bundle (``theories/nuclear_skyrme/io.py:16``): the reference's file names and order."""

from __future__ import annotations

from skyrmion_solver_amd.core.io import output_data_bundle_core

BUNDLE_SPEC = [  # theories/nuclear_skyrme/io.py:74
    ("grid", 0, "xGrid.dat"),
    ("grid", 1, "yGrid.dat"),
    ("grid", 2, "zGrid.dat"),
    ("Field", 0, "SigmaField.dat"),
    ("Field", 1, "PionField1.dat"),
    ("Field", 2, "PionField2.dat"),
    ("Field", 3, "PionField3.dat"),
    ("BaryonDensity", 0, "BaryonDensity.dat"),
    ("EnergyDensity", 0, "EnergyDensity.dat"),
]


def output_data_bundle(output_dir, h_Field, h_EnergyDensity, h_BaryonDensity, h_grid, xlen, ylen, zlen,
                       number_coordinates=3, number_total_fields=4, precision=32, binary=False):
    arrays = {"grid": h_grid, "Field": h_Field, "EnergyDensity": h_EnergyDensity,
              "BaryonDensity": h_BaryonDensity}
    output_data_bundle_core(output_dir, h_Field=h_Field, h_grid=h_grid, xlen=xlen, ylen=ylen, zlen=zlen,
                            number_coordinates=number_coordinates, number_total_fields=number_total_fields,
                            precision=precision, bundle_spec=BUNDLE_SPEC, arrays=arrays,
                            lattice_points_name="LatticePoints.dat", lattice_vectors_name="LatticeVectors.dat",
                            field_dump_name="Field.dat", binary=binary)
