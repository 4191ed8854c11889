"""Autotuner bookkeeping on the CPU: cache keys, store/lookup round trip, capacity rule."""
import json

from shapefromcaustics_amd import autotune


def test_cache_round_trip(tmp_path, monkeypatch):
    monkeypatch.setenv("PDS_TUNE_CACHE", str(tmp_path / "sub" / "t.json"))
    assert autotune.lookup(1 << 20, (3, 512, 512), 12, "auto") is None
    entry = {"tile": 32, "seg_photons": 1024, "block_photons": 4096, "capacity_pct": 160, "us": 12.5}
    autotune.store(1 << 20, (3, 512, 512), 12, "auto", entry)
    assert autotune.lookup(1 << 20, (3, 512, 512), 12, "auto") == {k: entry[k] for k in
                                                                    ("tile", "seg_photons", "block_photons",
                                                                     "capacity_pct")}
    # the same photon-count bucket hits, other sizes / caps / modes miss
    assert autotune.lookup((1 << 20) + 1000, (3, 512, 512), 12, "auto") is not None
    assert autotune.lookup(1 << 20, (3, 512, 512), 8, "auto") is None
    assert autotune.lookup(1 << 20, (3, 512, 512), 12, "deterministic") is None
    assert autotune.lookup(1 << 22, (3, 512, 512), 12, "auto") is None
    d = json.loads((tmp_path / "sub" / "t.json").read_text())
    assert len(d) == 1


def test_problem_key_accepts_mode_codes():
    assert autotune.problem_key(100, (1, 2, 3), 4, 3, arch="gfx950") == autotune.problem_key(
        100, (1, 2, 3), 4, "deterministic", arch="gfx950")


def test_capacity_rule():
    assert autotune.capacity_for(0, 0) == 200
    assert autotune.capacity_for(100, 100) == 130      # 1.25 x 100 % + 5
    assert autotune.capacity_for(10, 100) == 105       # floor of 105 %
    assert autotune.capacity_for(250, 100) == 318
