"""Coulomb-Skyrme model parameters.

Mirrors ``theories/coulomb_skyrme/params.py`` of the reference: defaults,
``mass = 2 m_pi/(g f_pi)`` and ``kappa = g^2 e^2 / 2`` (``:186-199``),
packing (``:255-298``).  The engine reads ``nsk = p_i[11]``,
``mass = p_f[8]`` and ``kappa = p_f[9]``.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from skyrmion_solver_amd.core._ansatz import resolve_ansatz
from skyrmion_solver_amd.core.params import Params as CoreParams
from skyrmion_solver_amd.core.params import ResolvedParams as CoreResolvedParams
from skyrmion_solver_amd.core.params import pack_device_params as pack_core_device_params


@dataclass(frozen=True)
class Params(CoreParams):
    number_total_fields: int = 5
    number_skyrme_fields: int = 4
    f_pi: float = 105.9
    m_pi: float = 138.0
    g: float = 4.0
    e: float = 0.302822
    mass: float | None = None
    kappa: float | None = None
    baryon_number: float = 1.0
    isospin_rotation: float = 0.0
    ansatz: str = "hedgehog"

    def resolved(self) -> "ResolvedParams":
        return ResolvedParams.from_params(self)


@dataclass(frozen=True)
class ResolvedParams(CoreResolvedParams):
    number_skyrme_fields: int
    f_pi: float
    m_pi: float
    g: float
    e: float
    mass: float
    kappa: float
    baryon_number: float
    isospin_rotation: float
    ansatz_hedgehog: bool
    ansatz_rational_map: bool
    ansatz_uniform: bool

    @staticmethod
    def from_params(p: Params) -> "ResolvedParams":
        hh, rm, uni = resolve_ansatz(p.ansatz)
        if float(p.g) == 0.0:
            raise ValueError("g must be nonzero when deriving mass = 2 m_pi / (g f_pi).")
        if float(p.f_pi) == 0.0:
            raise ValueError("f_pi must be nonzero when deriving mass = 2 m_pi / (g f_pi).")
        mass = 2.0 * float(p.m_pi) / (float(p.g) * float(p.f_pi)) if p.mass is None else p.mass
        kappa = float(p.g) * float(p.g) * float(p.e) * float(p.e) / 2.0 if p.kappa is None else p.kappa
        return ResolvedParams(**CoreResolvedParams.core_fields(p), number_skyrme_fields=int(p.number_skyrme_fields),
                              f_pi=float(p.f_pi), m_pi=float(p.m_pi), g=float(p.g), e=float(p.e),
                              mass=mass, kappa=kappa,
                              baryon_number=float(p.baryon_number), isospin_rotation=float(p.isospin_rotation),
                              ansatz_hedgehog=hh, ansatz_rational_map=rm, ansatz_uniform=uni)


def default_params(**overrides) -> Params:
    return Params().with_(**overrides)


def pack_device_params(rp: ResolvedParams):
    """``p_i += [nsk, hedgehog, rational_map, uniform]``; ``p_f += [mass, kappa, B, isospin_rotation]``."""
    p_i_core, p_f_core = pack_core_device_params(rp)
    p_i = np.concatenate((p_i_core, np.array([rp.number_skyrme_fields, int(rp.ansatz_hedgehog),
                                              int(rp.ansatz_rational_map), int(rp.ansatz_uniform)], dtype=np.int32)))
    p_f = np.concatenate((p_f_core, np.array([rp.mass, rp.kappa, rp.baryon_number, rp.isospin_rotation],
                                             dtype=np.float64)))
    return p_i, p_f


def describe() -> None:
    print("Coulomb-Skyrme model: 5 components (sigma, pi1, pi2, pi3, V).")
    print("  mass  = 2 m_pi / (g f_pi) -> p_f[8]")
    print("  kappa = g^2 e^2 / 2       -> p_f[9]")
