"""Omega-meson Skyrme model parameters.

Mirrors ``theories/omega_skyrme/params.py`` of the reference: defaults,
``mass = m_pi/m_omega`` and ``c_omega = m_omega*beta_omega/f_pi``
(``:192-204``), packing (``:262-305``).  The engine reads ``nsk = p_i[11]``,
``mass = p_f[8]`` and ``c_omega = p_f[9]``.
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
    f_pi: float = 186.0
    m_pi: float = 138.0
    m_omega: float = 782.0
    beta_omega: float = 8.253
    mass: float | None = None
    c_omega: float | None = None
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
    m_omega: float
    beta_omega: float
    mass: float
    c_omega: float
    baryon_number: float
    isospin_rotation: float
    ansatz_hedgehog: bool
    ansatz_rational_map: bool
    ansatz_uniform: bool

    @staticmethod
    def from_params(p: Params) -> "ResolvedParams":
        hh, rm, uni = resolve_ansatz(p.ansatz)
        if float(p.m_omega) == 0.0:
            raise ValueError("m_omega must be nonzero when deriving mass = m_pi / m_omega.")
        if float(p.f_pi) == 0.0:
            raise ValueError("f_pi must be nonzero when deriving c_omega = m_omega * beta_omega / f_pi.")
        mass = float(p.m_pi) / float(p.m_omega) if p.mass is None else p.mass
        c_omega = float(p.m_omega) * float(p.beta_omega) / float(p.f_pi) if p.c_omega is None else p.c_omega
        return ResolvedParams(**CoreResolvedParams.core_fields(p), number_skyrme_fields=int(p.number_skyrme_fields),
                              f_pi=float(p.f_pi), m_pi=float(p.m_pi), m_omega=float(p.m_omega),
                              beta_omega=float(p.beta_omega), mass=mass, c_omega=c_omega,
                              baryon_number=float(p.baryon_number), isospin_rotation=float(p.isospin_rotation),
                              ansatz_hedgehog=hh, ansatz_rational_map=rm, ansatz_uniform=uni)


def default_params(**overrides) -> Params:
    return Params().with_(**overrides)


def pack_device_params(rp: ResolvedParams):
    """``p_i += [nsk, hedgehog, rational_map, uniform]``; ``p_f += [mass, c_omega, B, isospin_rotation]``."""
    p_i_core, p_f_core = pack_core_device_params(rp)
    p_i = np.concatenate((p_i_core, np.array([rp.number_skyrme_fields, int(rp.ansatz_hedgehog),
                                              int(rp.ansatz_rational_map), int(rp.ansatz_uniform)], dtype=np.int32)))
    p_f = np.concatenate((p_f_core, np.array([rp.mass, rp.c_omega, rp.baryon_number, rp.isospin_rotation],
                                             dtype=np.float64)))
    return p_i, p_f


def describe() -> None:
    print("Omega-Skyrme model: 5 components (sigma, pi1, pi2, pi3, omega0).")
    print("  mass    = m_pi / m_omega              -> p_f[8]")
    print("  c_omega = m_omega * beta_omega / f_pi -> p_f[9]")
