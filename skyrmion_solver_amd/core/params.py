"""Theory-agnostic solver parameters and the packed parameter ABI.

Mirrors ``core/params.py`` of the reference (``Params`` ``:15-107``,
``ResolvedParams`` ``:124-226``, ``pack_device_params`` ``:248-268``): same
field names, defaults, derivation rules and packed layout, because the packed
``p_i`` (int32) / ``p_f`` (float64) arrays are what the engine's C-ABI
consumes verbatim:

``p_i = [X, Y, Z, halo, ncoords, nf, N, nf*N, killkinen, newtonflow, unit_length, <theory ints>]``
``p_f = [xsize, ysize, zsize, lsx, lsy, lsz, vol, dt, <theory floats>]``
"""

from __future__ import annotations

from dataclasses import dataclass, replace

import numpy as np


@dataclass(frozen=True)
class Params:
    """User-facing core parameters shared by all theories.

    ``lsx``/``lsy``/``lsz`` default to ``size / (len - 1)`` and ``time_step``
    to ``courant * lsx`` (``core/params.py:219-223``).
    """

    xlen: int = 64
    ylen: int = 64
    zlen: int = 64
    halo: int = 2
    number_coordinates: int = 3
    number_total_fields: int = 3

    xsize: float = 10.0
    ysize: float = 10.0
    zsize: float = 10.0

    lsx: float | None = None
    lsy: float | None = None
    lsz: float | None = None

    courant: float = 0.2
    time_step: float | None = None
    killkinen: bool = True
    newtonflow: bool = True
    unit_length: bool = True

    def with_(self, **kwargs) -> "Params":
        """Copy with selected fields overridden."""
        return replace(self, **kwargs)

    def resolved(self) -> "ResolvedParams":
        return ResolvedParams.from_params(self)


@dataclass(frozen=True)
class ResolvedParams:
    """Solver-ready parameters derived from :class:`Params`."""

    xlen: int
    ylen: int
    zlen: int
    halo: int
    number_coordinates: int
    number_total_fields: int
    dim_grid: int
    dim_fields: int
    killkinen: int
    newtonflow: int
    unit_length: int

    xsize: float
    ysize: float
    zsize: float
    lsx: float
    lsy: float
    lsz: float
    grid_volume: float
    time_step: float

    @staticmethod
    def core_fields(p: Params) -> dict:
        """The derived core fields (``core/params.py:219-226``)."""
        lsx = p.lsx if p.lsx is not None else p.xsize / (p.xlen - 1)
        lsy = p.lsy if p.lsy is not None else p.ysize / (p.ylen - 1)
        lsz = p.lsz if p.lsz is not None else p.zsize / (p.zlen - 1)
        grid_volume = lsx * lsy * lsz
        time_step = p.time_step if p.time_step is not None else (p.courant * lsx)
        dim_grid = p.xlen * p.ylen * p.zlen
        return dict(
            xlen=p.xlen, ylen=p.ylen, zlen=p.zlen, halo=p.halo,
            number_coordinates=p.number_coordinates,
            number_total_fields=int(p.number_total_fields),
            dim_grid=dim_grid, dim_fields=int(p.number_total_fields) * dim_grid,
            killkinen=1 if p.killkinen else 0, newtonflow=1 if p.newtonflow else 0,
            unit_length=1 if p.unit_length else 0,
            xsize=float(p.xsize), ysize=float(p.ysize), zsize=float(p.zsize),
            lsx=float(lsx), lsy=float(lsy), lsz=float(lsz),
            grid_volume=float(grid_volume), time_step=float(time_step),
        )

    @staticmethod
    def from_params(p: Params) -> "ResolvedParams":
        return ResolvedParams(**ResolvedParams.core_fields(p))


def default_params(**overrides) -> Params:
    return Params().with_(**overrides)


def pack_device_params(rp: ResolvedParams):
    """Core packed arrays ``(p_i int32[11], p_f float64[8])`` (``core/params.py:266-267``)."""
    p_i = np.array([rp.xlen, rp.ylen, rp.zlen, rp.halo, rp.number_coordinates, rp.number_total_fields,
                    rp.dim_grid, rp.dim_fields, rp.killkinen, rp.newtonflow, rp.unit_length], dtype=np.int32)
    p_f = np.array([rp.xsize, rp.ysize, rp.zsize, rp.lsx, rp.lsy, rp.lsz, rp.grid_volume, rp.time_step],
                   dtype=np.float64)
    return p_i, p_f
