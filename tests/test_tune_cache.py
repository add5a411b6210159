"""The autotuner's choice cache (mpm_amd.engine.TuneCache / autotune(cache=...)), on the host only: the key
separates what the tuned shape depends on, the file round-trips, an unreadable file counts as empty, and a
cached choice is returned without calling the engine."""
import json

import numpy as np

from mpm_amd import abi, scenes
from mpm_amd.engine import TuneCache, autotune, tune_key


def test_tune_key_separates_what_the_choice_depends_on():
    p = abi.preset(abi.PRESET_MIXED1024)
    sc = scenes.jelly128(n=4096)
    k = tune_key(p, sc)
    assert k == tune_key(p.copy(), sc)
    q = p.copy()
    q.grid_w += 16
    assert tune_key(q, sc) != k
    q = p.copy()
    q.mode = abi.MODE_TILED
    assert tune_key(q, sc) != k
    assert tune_key(p, sc, substeps=7) != k
    assert tune_key(p, scenes.jelly128(n=16384)) != k
    assert json.loads(k)["particles_log2"] == 12


def test_cache_round_trip_and_unreadable_file(tmp_path):
    path = tmp_path / "tune.json"
    cache = TuneCache(path)
    assert cache.get("k") is None
    cache.put("k", abi.MpmTuneResult(16, 8, 2048, 100, 0.07, 16))
    cache.put("j", abi.MpmTuneResult(32, 16, 512, 50, 0.5, 16))
    assert cache.get("k") == {"block_x": 16, "block_y": 8, "chunk": 2048, "sort_interval": 100,
                              "ms_per_substep": np.float32(0.07).item()}
    assert TuneCache(path).get("j")["chunk"] == 512
    path.write_text("{not json")
    assert cache.get("k") is None
    cache.put("k", abi.MpmTuneResult(8, 8, 1024, 25, 0.1, 16))
    assert cache.get("k")["block_x"] == 8 and cache.get("j") is None


def test_cached_choice_needs_no_engine(tmp_path):
    p = abi.preset(abi.PRESET_MIXED1024)
    sc = scenes.jelly128(n=4096)
    path = tmp_path / "tune.json"
    TuneCache(path).put(tune_key(p, sc), abi.MpmTuneResult(32, 16, 4096, 50, 0.25, 16))
    tuned, res = autotune(p, sc, cache=path)  # a miss would need a GPU; a hit returns at once
    assert (tuned.block_x, tuned.block_y, tuned.chunk, tuned.sort_interval) == (32, 16, 4096, 50)
    assert res.candidates == 0
    assert (p.block_x, p.block_y, p.chunk) == (16, 16, 4096)  # the caller's params (the preset's) are left alone
