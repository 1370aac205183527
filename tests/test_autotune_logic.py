"""CPU: the autotuner's x-chunk rule, block counts and cache file (no GPU calls)."""

from __future__ import annotations

import json

from skyrmion_solver_amd import autotune


def test_chunk_rule_splits_evenly_and_never_exceeds_target_by_much():
    assert autotune.chunk_rule(48, 0) == 0          # whole lattice
    assert autotune.chunk_rule(48, 64) == 0         # shorter than the target: whole lattice
    assert autotune.chunk_rule(128, 32) == 32
    assert autotune.chunk_rule(96, 32) == 32
    assert autotune.chunk_rule(100, 32) == 34       # 3 chunks of <= 34 planes cover 100
    for L in range(17, 300):
        for tgt in (16, 32, 64):
            c = autotune.chunk_rule(L, tgt)
            if L <= tgt:
                assert c == 0
            else:
                n = -(-L // c)
                assert n * c >= L and (n - 1) * c < L
                assert c <= 2 * tgt


def test_chunk_for_blocks_targets_the_block_count():
    # 64^3, 32x8 tiles: 2 x 8 = 16 columns; 256 blocks would need 4-plane chunks, the floor is 8
    assert autotune.chunk_for_blocks(64, 64, 64, 256) == 8
    # 128^3: 64 columns -> 4 chunks of 32
    assert autotune.chunk_for_blocks(128, 128, 128, 256) == 32
    # 192^3: 144 columns -> 2 chunks of 96
    assert autotune.chunk_for_blocks(192, 192, 192, 256) == 96
    # 16x16 tiles: 64^3 has 16 columns too
    assert autotune.chunk_for_blocks(64, 64, 64, 256, tile_z=16) == 8
    # a lattice whose columns alone reach the target runs whole
    assert autotune.chunk_for_blocks(64, 512, 512, 256) == 0
    for L in range(8, 260, 7):
        for tgt in (64, 128, 256, 512):
            c = autotune.chunk_for_blocks(L, L, L, tgt)
            if c:
                assert 8 <= c < L
                cols = -(-L // 32) * -(-L // 8)
                nb = cols * -(-L // c)
                assert nb <= max(2 * tgt, cols * -(-L // 8))


def test_batch_blocks_counts_tiles_and_chunks():
    # 128^3 with 32x8 tiles: 4 z-tiles x 16 y-tiles; 4 x-chunks of 32
    assert autotune.batch_blocks([(128, 128, 128)], 32, 32) == 4 * 16 * 4
    assert autotune.batch_blocks([(128, 128, 128)], 32, 0) == 4 * 16
    # 48^3 with 16x16 tiles: 3 x 3 tiles, whole lattice per block
    assert autotune.batch_blocks([(48, 48, 48)], 16, 64) == 9
    assert autotune.batch_blocks([(48, 48, 48), (64, 64, 64)], 64, 0) == 1 * 12 + 1 * 16
    # a tile per stage: its base 32 x 8 map is the larger one
    assert autotune.batch_blocks([(128, 128, 128)], 0, 32) == autotune.batch_blocks([(128, 128, 128)], 32, 32)


def test_cache_round_trip(tmp_path, monkeypatch):
    monkeypatch.setenv("SKX_TUNE_CACHE", str(tmp_path / "t.json"))
    autotune._store("k1", {"tile_z": 32})
    autotune._store("k2", {"tile_z": 16})
    data = json.loads((tmp_path / "t.json").read_text())
    assert data == {"k1": {"tile_z": 32}, "k2": {"tile_z": 16}}
    assert autotune._load_cache()["k2"]["tile_z"] == 16


def test_xchunk_candidates():
    assert autotune.xchunk_candidates(128) == [0, 8, 16, 32, 64, 128]
    assert autotune.xchunk_candidates(20) == [0, 8, 16, 20]
