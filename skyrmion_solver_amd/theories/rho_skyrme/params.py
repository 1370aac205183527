"""Rho-meson Skyrme model parameters.

Mirrors ``theories/rho_skyrme/params.py`` of the reference: 13 components
(``:63``), ``mpi = 2 m_pi/(e f_pi)``, ``mrho = 2 m_rho/(e f_pi)``,
``c_alpha = alpha e`` (``:176-190``), packing (``:248-288``).  The engine
reads ``mpi = p_f[8]``, ``mrho = p_f[9]`` and ``c_alpha = p_f[10]``.
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
    number_total_fields: int = 13
    mass_pion: float = 138.0
    mass_rho: float = 775.0
    f_pi: float = 129.0
    e_skyrme: float = 1.0
    alpha: float = 0.0617
    mpi: float | None = None
    mrho: float | None = None
    c_alpha: float | None = None
    baryon_number: float = 1.0
    isospin_rotation: float = 0.0
    ansatz: str = "hedgehog"

    def resolved(self) -> "ResolvedParams":
        return ResolvedParams.from_params(self)


@dataclass(frozen=True)
class ResolvedParams(CoreResolvedParams):
    mass_pion: float
    mass_rho: float
    f_pi: float
    e_skyrme: float
    alpha: float
    mpi: float
    mrho: float
    c_alpha: float
    mpi_is_derived: bool
    mrho_is_derived: bool
    c_alpha_is_derived: bool
    baryon_number: float
    isospin_rotation: float
    ansatz_hedgehog: bool
    ansatz_rational_map: bool
    ansatz_uniform: bool

    @staticmethod
    def from_params(p: Params) -> "ResolvedParams":
        hh, rm, uni = resolve_ansatz(p.ansatz)
        need = (p.mpi is None) or (p.mrho is None) or (p.c_alpha is None)
        if need and float(p.e_skyrme) == 0.0:
            raise ValueError("e_skyrme must be nonzero when deriving mpi, mrho, or c_alpha.")
        if need and float(p.f_pi) == 0.0:
            raise ValueError("f_pi must be nonzero when deriving mpi or mrho.")
        denom = float(p.e_skyrme) * float(p.f_pi)
        mpi = 2.0 * float(p.mass_pion) / denom if p.mpi is None else float(p.mpi)
        mrho = 2.0 * float(p.mass_rho) / denom if p.mrho is None else float(p.mrho)
        c_alpha = float(p.alpha) * float(p.e_skyrme) if p.c_alpha is None else float(p.c_alpha)
        return ResolvedParams(**CoreResolvedParams.core_fields(p), mass_pion=float(p.mass_pion),
                              mass_rho=float(p.mass_rho), f_pi=float(p.f_pi), e_skyrme=float(p.e_skyrme),
                              alpha=float(p.alpha), mpi=mpi, mrho=mrho, c_alpha=c_alpha,
                              mpi_is_derived=p.mpi is None, mrho_is_derived=p.mrho is None,
                              c_alpha_is_derived=p.c_alpha is None,
                              baryon_number=float(p.baryon_number), isospin_rotation=float(p.isospin_rotation),
                              ansatz_hedgehog=hh, ansatz_rational_map=rm, ansatz_uniform=uni)


def default_params(**overrides) -> Params:
    return Params().with_(**overrides)


def pack_device_params(rp: ResolvedParams):
    """``p_i += [hedgehog, rational_map, uniform]``; ``p_f += [mpi, mrho, c_alpha, B, isospin_rotation]``."""
    p_i_core, p_f_core = pack_core_device_params(rp)
    p_i = np.concatenate((p_i_core, np.array([int(rp.ansatz_hedgehog), int(rp.ansatz_rational_map),
                                              int(rp.ansatz_uniform)], dtype=np.int32)))
    p_f = np.concatenate((p_f_core, np.array([rp.mpi, rp.mrho, rp.c_alpha, rp.baryon_number, rp.isospin_rotation],
                                             dtype=np.float64)))
    return p_i, p_f


def describe() -> None:
    print("Rho-Skyrme model: 13 components (sigma, pi1..3, rho^a_i at 4 + 3 i + a).")
    print("  mpi     = 2 m_pi  / (e f_pi) -> p_f[8]")
    print("  mrho    = 2 m_rho / (e f_pi) -> p_f[9]")
    print("  c_alpha = alpha e            -> p_f[10]")
