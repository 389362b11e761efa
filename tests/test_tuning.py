"""CPU checks of the launch-shape table: the autotuner's decision rules on
constructed timings, the table validation, and the committed table."""
import json
import os

import pytest

from schnellfmm_amd import autotune as A
from schnellfmm_amd import tuning as T


def test_two_tile_threshold():
    # two tiles never slower: threshold 0 (always two)
    assert A.pick_two_tile_threshold([(78, 1.0, 0.9), (390, 2.0, 1.5), (3906, 20.0, 15.0)]) == 0
    # one tile wins on the small grids only
    rows = [(78, 1.0, 1.3), (195, 2.0, 2.1), (390, 3.0, 2.9), (3906, 20.0, 15.0)]
    assert A.pick_two_tile_threshold(rows) == 390
    # within the 1 % tolerance counts as not slower
    assert A.pick_two_tile_threshold([(78, 1.0, 1.005), (390, 3.0, 2.9)]) == 0
    # a slower large grid pushes the threshold past every smaller one
    assert A.pick_two_tile_threshold([(78, 1.0, 0.8), (390, 3.0, 3.5), (3906, 20.0, 15.0)]) == 3906
    # two tiles slower everywhere: past the largest grid measured
    assert A.pick_two_tile_threshold([(78, 1.0, 1.2), (390, 3.0, 3.5)]) == 391


def test_min_total_and_batch_rules():
    times = {256: [1.0, 10.0], 512: [1.1, 9.0], 1024: [1.5, 9.1]}
    # totals: 256 -> 11.0, 512 -> 10.1, 1024 -> 10.6
    assert A.pick_min_total(times) == 512
    # a large relative loss on a small shape does not outweigh a larger shape
    assert A.pick_min_total({256: [0.070, 0.285], 512: [0.087, 0.228]}) == 512
    assert A.pick_min_total({256: [1.0], 512: [1.0]}) == 256   # tie: the smaller
    b = {(4, 1.4): 50.0, (8, 1.4): 44.0, (16, 1.4): 44.3, (8, 0.7): 45.0}
    assert A.pick_batch(b) == (8, 1.4)
    assert A.pick_batch({(16, 1.4): 44.0, (8, 2.8): 44.2}) == (8, 2.8)   # within 1 %: fewer streams


def test_validate():
    p = T.validate({"fwd_target_wg": 1024})
    assert p["fwd_target_wg"] == 1024 and p["gram_max_blocks"] == T.DEFAULTS["gram_max_blocks"]
    for bad in ({"nope": 1}, {"fwd_target_wg": 0}, {"batch_streams": 17}, {"batch_gb_per_ms": 0.0},
                {"gram_max_blocks": 70000}, {"snp_two_tiles_min_blocks": -1}):
        with pytest.raises(ValueError):
            T.validate(bad)


def test_committed_table_is_valid():
    if not os.path.exists(T.DEFAULT_PATH):
        pytest.skip("no tuning table committed")
    with open(T.DEFAULT_PATH) as f:
        table = json.load(f)
    p = T.validate(table["params"])
    assert set(p) == set(T.DEFAULTS)
    assert "measurements" in table and table["command"].startswith("python -m schnellfmm_amd.autotune")
    # the two-tile threshold is one of the measured 2-tile grid works (or 0 / past the largest)
    works = sorted({r["blocks_two_tiles"] for r in table["measurements"]["snp_kernel_ms"]})
    assert p["snp_two_tiles_min_blocks"] in [0, *works, works[-1] + 1]


def test_env_disables_table(monkeypatch):
    monkeypatch.setenv("SLFMM_TUNING", "none")
    try:
        assert T.reload() == T.DEFAULTS
    finally:
        monkeypatch.delenv("SLFMM_TUNING")
        T.reload()
