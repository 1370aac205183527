"""Launch-shape autotuner for the stage kernels.

The stage kernels are built for three (y, z) tiles -- 16x16, 32x8 and 64x4
threads -- and march along x over a chunk of planes per block.  The best
pair depends on the lattice shape (a short z axis wastes lanes of a wide
tile; short x-chunks fill the device but re-read 4 halo planes per chunk) and
on the theory's register and LDS budget.  The tuner times short
device-resident runs of each candidate through the C ABI and keeps the
fastest:

* ``tune_engine``: one lattice (``skx_desc.tile_z`` / ``x_chunk``);
* ``tune_batch``: a set of lattices run as one batch (the batch's tile and a
  per-size-class x-chunk rule).

Results are cached in a JSON file (``$SKX_TUNE_CACHE``, default
``~/.cache/skyrmion_solver_amd/autotune.json``) keyed by device architecture
and CU count, theory, precision and lattice dimensions, so a size class is
timed once per device.
"""

from __future__ import annotations

import hashlib
import json
import os
import time
from pathlib import Path

import numpy as np

from skyrmion_solver_amd import _engine

TILES = (16, 32, 64)
# candidates of both tuners: 0 is the engine default (a tile per stage: 32 x 16 for the FP32
# stages that gain from it, 32 x 8 elsewhere), the others force one tile on every stage
ENGINE_TILES = (0,) + TILES


def cache_path() -> Path:
    return Path(os.environ.get("SKX_TUNE_CACHE", Path.home() / ".cache" / "skyrmion_solver_amd" / "autotune.json"))


def _load_cache() -> dict:
    p = cache_path()
    if p.exists():
        try:
            return json.loads(p.read_text())
        except (OSError, ValueError):
            return {}
    return {}


def _store(key: str, value: dict):
    p = cache_path()
    data = _load_cache()
    data[key] = value
    p.parent.mkdir(parents=True, exist_ok=True)
    tmp = p.with_suffix(".tmp")
    tmp.write_text(json.dumps(data, indent=1, sort_keys=True))
    os.replace(tmp, p)


def _device_key(device: int) -> str:
    arch, cus = _engine.device_info(device)
    return f"{arch}/{cus}cu"


def xchunk_candidates(X: int) -> list:
    """0 (the engine's occupancy rule) plus power-of-two chunks shorter than X and X itself."""
    return [0] + sorted({c for c in (8, 16, 32, 64) if c < X} | {X})


def _vacuum_field(nf: int, N: int) -> np.ndarray:
    F = np.zeros(nf * N)
    F[:N] = 1.0
    return F


def tune_engine(theory: int, precision: int, p_i, p_f, F=None, steps: int = 16, device: int = 0,
                use_cache: bool = True, tiles=ENGINE_TILES, xchunks=None) -> dict:
    """Fastest (tile_z, x_chunk) for one lattice; returns ``{"tile_z", "x_chunk", "us_per_step", "table"}``.

    `F` is the field timed (default: the vacuum, which runs the same
    instructions as any finite field)."""
    p_i = np.asarray(p_i, dtype=np.int32)
    X, Y, Z, nf = (int(v) for v in (p_i[0], p_i[1], p_i[2], p_i[5]))
    key = f"engine|{_device_key(device)}|th{theory}|p{precision}|{X}x{Y}x{Z}"
    if use_cache:
        hit = _load_cache().get(key)
        if hit:
            return hit
    N = X * Y * Z
    F = _vacuum_field(nf, N) if F is None else np.ascontiguousarray(F, dtype=np.float64).ravel()
    table = []
    for tz in tiles:
        for xc in (xchunks or xchunk_candidates(X)):
            with _engine.Engine(theory, precision, p_i, p_f, device=device, tile_z=tz, x_chunk=xc) as eng:
                eng.upload(_engine.FIELD, F)
                eng.anf_run(2, -1.0, 2)  # warm-up: graph capture, caches
                eng.upload(_engine.FIELD, F)
                st = eng.anf_run(steps, -1.0, steps)
            table.append(dict(tile_z=tz, x_chunk=xc, us_per_step=1e6 * st.seconds / max(st.steps, 1)))
    best = min(table, key=lambda r: r["us_per_step"])
    out = dict(tile_z=best["tile_z"], x_chunk=best["x_chunk"], us_per_step=best["us_per_step"], table=table)
    if use_cache:
        _store(key, out)
    return out


def chunk_rule(L: int, target: int) -> int:
    """x-chunk of a batch lattice of x-length L under the rule "about `target` planes per block"
    (0 = the whole lattice)."""
    if target <= 0 or L <= target:
        return 0
    n = max(1, round(L / target))
    return -(-L // n)


def chunk_for_blocks(X: int, Y: int, Z: int, target_blocks: int, tile_z: int = 0, min_planes: int = 8) -> int:
    """x-chunk that gives an X x Y x Z batch lattice about `target_blocks` blocks of the (tile_z x 256/tile_z)
    tile (tile_z 0: the 32 x 8 base tile), with at least `min_planes` planes per block (0 = the whole
    lattice).  Small lattices get short chunks, so that a batch of a few of them still spreads over the
    CUs; large ones get long chunks, which re-read fewer x-halo planes."""
    tz = tile_z or 32
    ty = 256 // tz
    cols = max(1, -(-Z // tz) * -(-Y // ty))
    n = max(1, round(target_blocks / cols))
    xc = max(min_planes, -(-X // n))
    return 0 if xc >= X else xc


def batch_blocks(dims, tile_z: int, target: int) -> int:
    """Blocks of a batch of lattices `dims` (X, Y, Z) with the (tile_z x 256/tile_z) tile and the
    ``chunk_rule(X, target)`` x-chunks.  tile_z 0 (a tile per stage) counts its base 32 x 8 tile,
    the larger of its two block maps."""
    tile_z = tile_z or 32
    ty = 256 // tile_z
    total = 0
    for X, Y, Z in dims:
        xc = chunk_rule(X, target) or X
        total += -(-Z // tile_z) * -(-Y // ty) * -(-X // xc)
    return total


def tune_batch(theory: int, precision: int, lattices, steps: int = 8, device: int = 0, use_cache: bool = True,
               tiles=ENGINE_TILES, targets=(0, 64, 32, 16), max_slots: int | None = None) -> dict:
    """Fastest (tile_z, x-chunk target) for running `lattices` (a list of ``(p_i, p_f)``) as one batch.

    Returns ``{"tile_z", "target", "site_updates_per_s", "table"}``; the
    x-chunk of each lattice is ``chunk_rule(X, target)``."""
    dims = sorted(tuple(int(v) for v in np.asarray(p_i)[:3]) for p_i, _ in lattices)
    sig = ",".join("x".join(map(str, d)) for d in dims)
    key = (f"batch|{_device_key(device)}|th{theory}|p{precision}|{len(dims)}|{hashlib.sha1(sig.encode()).hexdigest()[:12]}"
           f"|t{'-'.join(map(str, tiles))}|c{'-'.join(map(str, targets))}")
    if use_cache:
        hit = _load_cache().get(key)
        if hit:
            return hit
    slots = max_slots or len(lattices)
    sites = sum(int(np.asarray(p_i)[6]) for p_i, _ in lattices)
    table = []
    for tz in tiles:
        for tgt in targets:
            blocks = batch_blocks([tuple(int(v) for v in np.asarray(p_i)[:3]) for p_i, _ in lattices], tz, tgt)
            with _engine.Batch(theory, precision, max_slots=slots, max_blocks=blocks, tile_z=tz,
                               device=device) as b:
                for p_i, p_f in lattices:
                    p_i = np.asarray(p_i)
                    nf, N = int(p_i[5]), int(p_i[6])
                    b.add(p_i, p_f, _vacuum_field(nf, N), max_steps=1 << 30, x_chunk=chunk_rule(int(p_i[0]), tgt))
                b.run(2)
                t0 = time.perf_counter()
                b.run(steps)
                dt = time.perf_counter() - t0
            table.append(dict(tile_z=tz, target=tgt, site_updates_per_s=sites * steps / dt))
    best = max(table, key=lambda r: r["site_updates_per_s"])
    out = dict(tile_z=best["tile_z"], target=best["target"], site_updates_per_s=best["site_updates_per_s"],
               table=table)
    if use_cache:
        _store(key, out)
    return out
