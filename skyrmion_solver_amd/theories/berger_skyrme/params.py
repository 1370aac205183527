"""Berger-Skyrme model parameters (Skyrme field on a squashed 3-sphere).

Mirrors ``theories/berger_skyrme/params.py`` of the reference: fields
(``:51-59``), resolution (``:78-183``) and packing (``:206-250``).  The
engine reads ``mpi = p_f[8]``, ``kappa = p_f[9]`` and ``alpha = p_f[12]``
(the squashing; ``alpha = 1`` is the nuclear model).
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
    number_total_fields: int = 4
    mpi: float = 1.0
    kappa: float = 1.0
    alpha: float = 1.0
    baryon_number: float = 1.0
    isospin_rotation: float = 0.0
    ansatz: str = "hedgehog"

    def resolved(self) -> "ResolvedParams":
        return ResolvedParams.from_params(self)


@dataclass(frozen=True)
class ResolvedParams(CoreResolvedParams):
    mpi: float
    kappa: float
    alpha: float
    baryon_number: float
    isospin_rotation: float
    ansatz_hedgehog: bool
    ansatz_rational_map: bool
    ansatz_uniform: bool

    @staticmethod
    def from_params(p: Params) -> "ResolvedParams":
        hh, rm, uni = resolve_ansatz(p.ansatz)
        return ResolvedParams(**CoreResolvedParams.core_fields(p), mpi=float(p.mpi), kappa=float(p.kappa), alpha=float(p.alpha),
                              baryon_number=float(p.baryon_number), isospin_rotation=float(p.isospin_rotation),
                              ansatz_hedgehog=hh, ansatz_rational_map=rm, ansatz_uniform=uni)


def default_params(**overrides) -> Params:
    return Params().with_(**overrides)


def pack_device_params(rp: ResolvedParams):
    """``p_i += [hedgehog, rational_map, uniform]``; ``p_f += [mpi, kappa, B, isospin_rotation, alpha]``."""
    p_i_core, p_f_core = pack_core_device_params(rp)
    p_i = np.concatenate((p_i_core, np.array([int(rp.ansatz_hedgehog), int(rp.ansatz_rational_map),
                                              int(rp.ansatz_uniform)], dtype=np.int32)))
    p_f = np.concatenate((p_f_core, np.array([rp.mpi, rp.kappa, rp.baryon_number, rp.isospin_rotation, rp.alpha], dtype=np.float64)))
    return p_i, p_f


def describe() -> None:
    print("Berger-Skyrme model: 4 components (sigma, pi1, pi2, pi3) on a squashed S^3.")
    print("  mpi   : pion mass in the potential mpi^2 (1 - sigma)          -> p_f[8]")
    print("  kappa : Skyrme quartic-term coefficient                       -> p_f[9]")
    print("  alpha : Berger squashing, beta = alpha^2 - 1 (alpha = 1: nuclear) -> p_f[12]")
    print("  baryon_number, isospin_rotation, ansatz : initial-data controls")
