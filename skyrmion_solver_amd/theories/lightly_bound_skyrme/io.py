"""Lightly bound Skyrme model output bundle (``theories/lightly_bound_skyrme/io.py:10``): the nuclear file set."""

from __future__ import annotations

from skyrmion_solver_amd.theories.nuclear_skyrme.io import BUNDLE_SPEC, output_data_bundle  # noqa: F401
