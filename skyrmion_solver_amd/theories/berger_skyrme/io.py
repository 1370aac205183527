"""Berger Skyrme model output bundle (``theories/berger_skyrme/io.py:10``): the nuclear file set."""

from __future__ import annotations

from skyrmion_solver_amd.theories.nuclear_skyrme.io import BUNDLE_SPEC, output_data_bundle  # noqa: F401
