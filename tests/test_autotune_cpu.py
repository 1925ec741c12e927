"""CPU checks of the autotune table: the library's per-bucket knob values
(host memory only, so no GPU is needed), argument validation, and the JSON
cache round trip."""
from __future__ import annotations

import json

import pytest

from flash_ph import autotune, _lib


@pytest.fixture(autouse=True)
def _clean_table():
    autotune.reset()
    yield
    autotune.reset()


def test_candidates_and_defaults():
    assert autotune.candidates("screen_tile") == [32, 64, 128]
    assert autotune.candidates("enum_block") == [64, 128, 256, 512, 1024]
    assert autotune.candidates("apparent_block") == [64, 128, 256, 512, 1024]
    assert autotune.candidates("queue_wgs") == [1, 2, 4, 8, 16, 32, 64]
    # the device-wide queue from 8192 columns on, one workgroup below
    assert autotune.query("queue_wgs", 26_000) == 16
    assert autotune.query("queue_wgs", 8192) == 16
    assert autotune.query("queue_wgs", 8191) == 1
    assert autotune.query("screen_tile", 10_000, 10) == 128
    assert autotune.query("enum_block", 26_000) == 256
    assert autotune.query("apparent_block", 26_000) == 256


def test_buckets_are_log2_of_size_and_d():
    autotune.set_value("screen_tile", 10_000, 10, 32)
    # same bucket: floor(log2) of 8192..16383 and of 8..15
    assert autotune.query("screen_tile", 8192, 8) == 32
    assert autotune.query("screen_tile", 16383, 15) == 32
    # neighbouring buckets keep the default
    assert autotune.query("screen_tile", 16384, 10) == 128
    assert autotune.query("screen_tile", 10_000, 16) == 128
    assert autotune.query("screen_tile", 8191, 10) == 128
    # knobs are independent
    assert autotune.query("enum_block", 10_000, 10) == 256
    autotune.set_value("screen_tile", 10_000, 10, 0)
    assert autotune.query("screen_tile", 10_000, 10) == 128


def test_invalid_values_rejected():
    with pytest.raises(_lib.FlashPHError):
        autotune.set_value("screen_tile", 1000, 10, 96)
    with pytest.raises(_lib.FlashPHError):
        autotune.set_value("enum_block", 1000, 0, 100)
    with pytest.raises(_lib.FlashPHError):
        autotune.set_value("enum_block", -1, 0, 256)
    with pytest.raises(ValueError):
        autotune.query("tile", 10)
    lib = _lib.load()
    assert lib.fph_autotune_query(7, 10, 0) == _lib.FPH_ERR_ARG
    assert lib.fph_autotune_set(4, 10, 0, 64) == _lib.FPH_ERR_ARG
    assert lib.fph_autotune_set(3, 10, 0, 3) == _lib.FPH_ERR_ARG
    assert autotune.query("screen_tile", 1000, 10) == 128


def test_reset_restores_defaults():
    autotune.set_value("enum_block", 5000, 0, 64)
    autotune.set_value("apparent_block", 5000, 0, 1024)
    autotune.reset()
    assert autotune.query("enum_block", 5000) == 256
    assert autotune.query("apparent_block", 5000) == 256


def test_cache_round_trip(tmp_path):
    path = str(tmp_path / "tune.json")
    autotune.save_cache([{"knob": "screen_tile", "size": 10_000, "d": 10, "value": 32},
                         {"knob": "enum_block", "size": 26_000, "d": 0, "value": 128}], path)
    # same bucket replaces, new bucket appends
    autotune.save_cache([{"knob": "screen_tile", "size": 12_000, "d": 12, "value": 128},
                         {"knob": "apparent_block", "size": 300, "d": 0, "value": 64}], path)
    data = json.load(open(path))
    assert data["arch"] == "gfx950" and len(data["entries"]) == 3
    assert autotune.load_cache(path) == 3
    assert autotune.query("screen_tile", 10_000, 10) == 128
    assert autotune.query("enum_block", 26_000) == 128
    assert autotune.query("apparent_block", 300) == 64
    assert autotune.load_cache(str(tmp_path / "missing.json")) == 0


def test_cache_path_from_environment(monkeypatch, tmp_path):
    monkeypatch.setenv("FLASH_PH_TUNE_CACHE", str(tmp_path / "x.json"))
    assert autotune.cache_path() == str(tmp_path / "x.json")
    monkeypatch.delenv("FLASH_PH_TUNE_CACHE")
    assert autotune.cache_path().endswith("autotune-gfx950.json")


def test_tune_needs_gpu():
    with pytest.raises(RuntimeError):
        autotune.tune([[0.0, 0.0], [1.0, 0.0]], 2.0)
