"""Tuning table: band lookup, defaults, validation (host logic of beat_hip/tuning.py)."""

import json

import pytest

from beat_hip import tuning as TU


def _table():
    return TU.validate({
        "lu": [{"precision": "mixed", "n_max": 4096, "batch_max": 64, "choice": {"split_tile": 1, "trsv_wg_min_batch": 16,
                                                                                 "lookahead_wgs": -1}},
               {"precision": "fp64", "n_max": 1024, "batch_max": 1, "choice": {"split_tile": 0, "trsv_wg_min_batch": 0,
                                                                               "lookahead_wgs": 128}}],
        "assembly": [{"precision": "mixed", "faces_max": 9000, "choice": {"strip_nf": 8, "field_slice_budget": 4096}}]})


def test_bands():
    assert TU.band(1, TU.N_BANDS) == 1024 and TU.band(1024, TU.N_BANDS) == 1024 and TU.band(1025, TU.N_BANDS) == 2048
    assert TU.band(30000, TU.N_BANDS) is None
    assert TU.band(8, TU.BATCH_BANDS) == 23 and TU.band(24, TU.BATCH_BANDS) == 64 and TU.band(1, TU.BATCH_BANDS) == 1


def test_lookup_hits_its_band_and_defaults_elsewhere():
    t = _table()
    assert t.lu_choice("mixed", 3002, 40) == TU.LuChoice(1, 16, -1)
    assert t.lu_choice("mixed", 2049, 24) == TU.LuChoice(1, 16, -1)
    assert t.lu_choice("mixed", 3002, 23) == TU.LuChoice()  # other batch band
    assert t.lu_choice("mixed", 5000, 40) == TU.LuChoice()  # other N band
    assert t.lu_choice("fp64", 642, 1) == TU.LuChoice(0, 0, 128)
    assert t.lu_choice("fp64", 642, 2) == TU.LuChoice()
    assert t.lu_choice("mixed", 30000, 40) == TU.LuChoice()  # beyond every band
    assert t.assembly_choice("mixed", 6000) == TU.AssemblyChoice(8, 4096)
    assert t.assembly_choice("fp64", 6000) == TU.AssemblyChoice()
    assert TU.LuChoice() == TU.LuChoice(0, 8, 0) and TU.AssemblyChoice() == TU.AssemblyChoice(4, 2048)


@pytest.mark.parametrize("bad", [
    {"lu": [{"precision": "mixed", "n_max": 3000, "batch_max": 64, "choice": {}}]},
    {"lu": [{"precision": "mixed", "n_max": 4096, "batch_max": 64, "choice": {"split_tile": 2}}]},
    {"lu": [{"precision": "bf16", "n_max": 4096, "batch_max": 64, "choice": {}}]},
    {"assembly": [{"precision": "mixed", "faces_max": 9000, "choice": {"strip_nf": 0}}]},
])
def test_validation_rejects_bad_entries(bad):
    with pytest.raises((ValueError, TypeError)):
        TU.validate(bad)


def test_committed_table_is_valid():
    if TU.TABLE_PATH.exists():
        t = TU.validate(json.loads(TU.TABLE_PATH.read_text()))
        assert t.lu and t.assembly
        for e in t.lu:  # every entry was measured and is no slower than the defaults
            assert e["ms"] <= e["default_ms"] * 1.0001
