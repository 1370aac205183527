"""Public theory interface: ``load_theory``, ``list``, ``print_table``.

Mirrors ``theories/__init__.py`` of the reference.  The six theories whose
hot path the engine runs are registered here with the reference's names and
aliases (``theories/*/__init__.py:20-26``).
"""

from __future__ import annotations

from .registry import (TheorySpec, describe_theory, get_theory_spec, list_theories, load_theory,
                       print_theory_table, register_theory)

_BUILTIN = (
    TheorySpec(name="Nuclear Skyrme model", aliases=("Skyrme Model", "3D skyrmions", "Skyrmions"),
               import_path="skyrmion_solver_amd.theories.nuclear_skyrme",
               description="Skyrmions in the 3D nuclear Skyrme model, with the standard pion mass potential.",
               version="1.0"),
    TheorySpec(name="Omega-Skyrme model",
               aliases=("Omega-Skyrme", "Omega-meson model", "Omega-meson Skyrme", "Omega-sigma model"),
               import_path="skyrmion_solver_amd.theories.omega_skyrme",
               description="Skyrmions coupled to the omega meson.", version="1.0"),
    TheorySpec(name="Coulomb-Skyrme model", aliases=("Coulomb Skyrme", "Maxwell Skyrme"),
               import_path="skyrmion_solver_amd.theories.coulomb_skyrme",
               description="Skyrmions with an electrostatic (Coulomb) potential.", version="1.0"),
    TheorySpec(name="Rho-Skyrme model", aliases=("Rho-Skyrme", "Rho-skyrmions", "Skyrmions with rho mesons"),
               import_path="skyrmion_solver_amd.theories.rho_skyrme",
               description="Skyrmions coupled to rho mesons.", version="1.0"),
    TheorySpec(name="Berger Skyrme model", aliases=("Berger Skyrme", "Squashed sphere Skyrme", "Berger skyrmions"),
               import_path="skyrmion_solver_amd.theories.berger_skyrme",
               description="Skyrmions in the 3D Berger-Skyrme model, with the standard pion mass potential.", version="1.0"),
    TheorySpec(name="Lightly bound Skyrme model", aliases=("Lightly bound Skyrme", "Loosely bound Skyrme", "Loosely bound Skyrme model"),
               import_path="skyrmion_solver_amd.theories.lightly_bound_skyrme",
               description="Skyrmions in the lightly bound Skyrme model.", version="1.0"),
)

for _spec in _BUILTIN:
    register_theory(_spec)


def list() -> tuple:  # noqa: A001 - same name as the reference's API
    return list_theories()


def print_description(name: str) -> None:
    describe_theory(name)


def print_table() -> None:
    print_theory_table()
