"""Host logic of the launch-shape autotuner's cache (rphys.py): keys by (config, body-count
power-of-two bucket), tolerant loading, atomic stores.  The search itself runs on the device
(tests/test_gpu_tuning.py)."""
import json
import os

import rphys


def test_key_buckets_by_power_of_two():
    assert rphys.tuning_key("batch", 605152) == "batch/n2^19"
    assert rphys.tuning_key("batch", 524288) == "batch/n2^19"
    assert rphys.tuning_key("batch", 524287) == "batch/n2^18"
    assert rphys.tuning_key("x", 0) == "x/n2^0"


def test_missing_or_corrupt_cache_loads_empty(tmp_path):
    assert rphys.load_tuning_cache(str(tmp_path / "none.json")) == {}
    bad = tmp_path / "bad.json"
    bad.write_text("{not json")
    assert rphys.load_tuning_cache(str(bad)) == {}
    lst = tmp_path / "list.json"
    lst.write_text("[1, 2]")
    assert rphys.load_tuning_cache(str(lst)) == {}


def test_store_round_trip_is_atomic(tmp_path):
    path = str(tmp_path / "sub" / "tuning.json")
    data = {"batch/n2^19": {"grid_small": 4096, "grid_large": 128, "grid_huge": 64, "ms_per_tick": 61.5, "ticks": 2}}
    rphys.store_tuning_cache(path, data)
    assert rphys.load_tuning_cache(path) == data
    assert json.load(open(path)) == data
    assert [f for f in os.listdir(os.path.dirname(path)) if f.startswith(".tuning")] == []


def test_env_overrides_default_location(monkeypatch, tmp_path):
    monkeypatch.setenv("RPHYS_TUNING_CACHE", str(tmp_path / "t.json"))
    assert rphys.default_tuning_cache() == str(tmp_path / "t.json")
