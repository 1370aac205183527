"""Theory registry: discovery and loading of theory plugins.

Mirrors ``theories/registry.py`` of the reference (``TheorySpec`` ``:19-54``,
``register_theory``, ``list_theories``, ``get_theory_spec``, ``load_theory``
``:168-196``, ``describe_theory``, ``print_theory_table``): the same names,
case-insensitive aliases and errors.
"""

from __future__ import annotations

import importlib
from dataclasses import dataclass
from types import ModuleType
from typing import Dict


@dataclass(frozen=True)
class TheorySpec:
    name: str
    import_path: str
    description: str = ""
    version: str = "0.0"
    aliases: tuple = ()
    required_submodules: tuple = ("params", "kernels", "initial_config", "observables")


_REGISTRY: Dict[str, TheorySpec] = {}


def register_theory(spec: TheorySpec) -> None:
    for k in [spec.name, *spec.aliases]:
        key = (k or "").strip().lower()
        if not key:
            raise ValueError("TheorySpec names must be non-empty")
        if key in _REGISTRY:
            raise ValueError(f"Theory '{key}' is already registered")
        _REGISTRY[key] = spec


def list_theories() -> tuple:
    return tuple(sorted({spec.name for spec in _REGISTRY.values()}))


def get_theory_spec(name: str) -> TheorySpec:
    key = (name or "").strip().lower()
    if key not in _REGISTRY:
        raise KeyError(f"Unknown theory '{name}'. Available: {list_theories()}")
    return _REGISTRY[key]


def load_theory(name: str) -> ModuleType:
    spec = get_theory_spec(name)
    mod = importlib.import_module(spec.import_path)
    for sub in spec.required_submodules:
        importlib.import_module(f"{spec.import_path}.{sub}")
    return mod


def describe_theory(name: str) -> None:
    spec = get_theory_spec(name)
    print(f"Theory: {spec.name}")
    print(f"Import path: {spec.import_path}")
    print(f"Version: {spec.version}")
    if spec.aliases:
        print("Aliases:", ", ".join(spec.aliases))
    if spec.description:
        print(f"Description: {spec.description}")


def print_theory_table() -> None:
    specs = sorted({s for s in _REGISTRY.values()}, key=lambda s: s.name.lower())
    if not specs:
        print("No theories registered.")
        return
    w = max(len("Theory"), max(len(s.name) for s in specs))
    print(f"{'Theory':<{w}}  Version  Description")
    print("-" * (w + 24))
    for s in specs:
        print(f"{s.name:<{w}}  {s.version:<7}  {s.description}")
