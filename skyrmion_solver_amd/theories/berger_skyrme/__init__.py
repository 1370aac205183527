"""Berger Skyrme model plugin (engine theory ``THEORY_BERGER``)."""

from __future__ import annotations

from skyrmion_solver_amd.theories.registry import get_theory_spec

from . import initial_config, io, kernels, observables, params  # noqa: F401

THEORY_SPEC = get_theory_spec("Berger Skyrme model")
threads3d = (8, 8, 4)


def describe() -> None:
    print(f"{THEORY_SPEC.name} (version {THEORY_SPEC.version})")
    print(THEORY_SPEC.description)
    params.describe()


def print_instructions() -> None:
    describe()
